"""Host-array calls from many threads at once.

The host-array entry points stage caller arrays through a pool of contexts (device
buffers, a stream and a pinned ring each), so calls from several threads run side
by side.  Every result of a concurrent run must equal, bit for bit, what the same
call gives alone (the kernels are deterministic), across real, complex, mixed and
matrix-vector calls in FP64, STRICT and the modular form, with padded leading
dimensions and with arrays larger than one staging chunk.
"""
import ctypes
import threading

import numpy as np
import pytest

from flag4j_amd import _native as L

pytestmark = pytest.mark.gpu


def p(a):
    return ctypes.c_void_p(a.ctypes.data)


def make_cases():
    r = np.random.default_rng(77)
    cases = []
    for i in range(16):
        kind = ["real", "cplx", "mixed", "gemv"][i % 4]
        prec = ["fp64", "strict", "split_int8", "fp64"][(i // 4) % 4]
        M, N, K = int(r.integers(40, 700)), int(r.integers(40, 700)), int(r.integers(40, 900))
        if i == 5:
            M, N, K = 2300, 2100, 1900               # A, B and C above one 16 MiB staging chunk
        pad = int(r.integers(0, 5))
        if kind == "cplx":
            A = r.standard_normal((M, K + pad)) + 1j * r.standard_normal((M, K + pad))
            B = r.standard_normal((K, N + pad)) + 1j * r.standard_normal((K, N + pad))
        elif kind == "mixed":
            A = r.standard_normal((M, K + pad))
            B = r.standard_normal((K, N + pad)) + 1j * r.standard_normal((K, N + pad))
        elif kind == "gemv":
            A = r.standard_normal((M, K + pad))
            B = r.standard_normal(K * 3)                 # x with stride 3
        else:
            A = r.standard_normal((M, K + pad))
            B = r.standard_normal((K, N + pad))
        cases.append((kind, prec if kind != "gemv" else ("fp64" if prec != "strict" else "strict"), M, N, K, A, B))
    return cases


def run_case(case):
    kind, prec, M, N, K, A, B = case
    lib = L.load()
    pr = L.PREC[prec]
    if kind == "real":
        C = np.full((M, N + 3), np.nan)
        st = lib.f4j_dgemm(0, 0, M, N, K, p(A), A.shape[1], p(B), B.shape[1], p(C), N + 3, pr, 0)
    elif kind == "cplx":
        C = np.full((M, N + 3), np.nan + 0j)
        st = lib.f4j_zgemm(0, 0, M, N, K, p(A), A.shape[1], p(B), B.shape[1], p(C), N + 3, pr, 0, 1)
    elif kind == "mixed":
        C = np.full((M, N + 3), np.nan + 0j)
        st = lib.f4j_dzgemm(M, N, K, p(A), A.shape[1], p(B), B.shape[1], p(C), N + 3, pr, 0)
    else:
        C = np.full(2 * M, np.nan)
        st = lib.f4j_dgemv(M, K, p(A), A.shape[1], p(B), 3, p(C), 2, pr)
    if st != L.OK:
        raise L.NativeError(st, lib.f4j_last_error().decode())
    return C


def test_concurrent_host_calls_match_serial_results():
    cases = make_cases()
    serial = [run_case(c) for c in cases]
    errors = []
    results = {}

    def worker(t):
        try:
            for rep in range(3):
                order = [(t + 5 * j + rep) % len(cases) for j in range(len(cases))]
                for i in order[:6]:
                    results[(t, rep, i)] = run_case(cases[i])
        except Exception as e:           # surfaced below with the thread's id
            errors.append((t, repr(e)))

    threads = [threading.Thread(target=worker, args=(t,)) for t in range(8)]
    for th in threads:
        th.start()
    for th in threads:
        th.join(timeout=240)
    assert not any(th.is_alive() for th in threads), "a host call did not return"
    assert not errors, errors
    assert len(results) == 8 * 3 * 6
    for (t, rep, i), C in results.items():
        assert np.array_equal(C.view(np.float64), serial[i].view(np.float64), equal_nan=True), (t, rep, i, cases[i][:5])
