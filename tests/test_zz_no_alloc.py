"""The allocation-free entry f4j_dgemm_dev_noalloc against f4j_dgemm_dev.

Same kernels on the same inputs, so the results must be equal bit for bit: the
modular INT8 form on the library's fixed scratch array (two calls in a row, the
second smaller, so stale scratch from the first call cannot pass), FP64, and the
errors for a product whose scratch does not fit and for the digit forms.
"""
import numpy as np
import pytest

from flag4j_amd import _native as L
from flag4j_amd import device, rng
from oracle import oracle as o

pytestmark = pytest.mark.gpu


def run(entry, opA, opB, A, B, M, N, K, prec, slices=0):
    lda, ldb = A.shape[1], B.shape[1]
    with device.DeviceBuffer.from_array(A) as dA, device.DeviceBuffer.from_array(B) as dB, \
            device.DeviceBuffer(M * N * 8) as dC:
        dC.upload(np.full(M * N, np.nan))
        st = getattr(L.load(), entry)(L.OP[opA], L.OP[opB], M, N, K, dA.ptr, lda, dB.ptr, ldb, dC.ptr, N,
                                      L.PREC[prec], slices, None)
        L.check(st)
        device.sync()
        rep = L.last_split_report()
        return dC.download(np.float64, M * N).reshape(M, N), rep


@pytest.fixture(autouse=True)
def split_config():
    L.config_set("splitCheckRows", 4)
    L.config_set("splitTolerance", 1e-15)
    L.config_set("splitFallback", 1)
    yield


def test_noalloc_int8_equals_dev_entry_twice():
    L.check(L.load().f4j_init(0))
    cases = [("N", "N", 1024, 1024, 1024, 61), ("T", "N", 300, 200, 500, 62)]
    for opA, opB, M, N, K, seed in cases:
        print(f"[noalloc] split_int8 {opA}{opB} {M}x{N}x{K}", flush=True)
        A = rng.uniform(seed, *((M, K) if opA == "N" else (K, M)))
        B = rng.uniform(seed + 100, *((K, N) if opB == "N" else (N, K)))
        ref, rep_ref = run("f4j_dgemm_dev", opA, opB, A, B, M, N, K, "split_int8")
        got, rep = run("f4j_dgemm_dev_noalloc", opA, opB, A, B, M, N, K, "split_int8")
        assert rep.precision == rep_ref.precision == L.PREC["split_int8"]
        assert rep.slices == rep_ref.slices and rep.max_rel_err == rep_ref.max_rel_err
        assert np.array_equal(got.view(np.int64), ref.view(np.int64)), (M, N, K)
        Aop = A if opA == "N" else np.ascontiguousarray(A.T)
        r = np.random.default_rng(seed)
        I, J = r.integers(0, M, 256), r.integers(0, N, 256)
        hi, lo, ab = o.real_dot2(Aop, B, I, J)
        assert float(np.max(np.abs((got[I, J] - hi) - lo) / ab)) <= 1e-15


def test_noalloc_fp64_equals_dev_entry():
    M, N, K = 700, 500, 300
    A, B = rng.uniform(63, M, K), rng.uniform(64, K, N)
    ref, _ = run("f4j_dgemm_dev", "N", "N", A, B, M, N, K, "fp64")
    got, _ = run("f4j_dgemm_dev_noalloc", "N", "N", A, B, M, N, K, "fp64")
    assert np.array_equal(got.view(np.int64), ref.view(np.int64))


def test_noalloc_rejects_what_does_not_fit():
    M = N = K = 2048                         # 16 moduli: 3 x 64 MiB of residues
    with device.DeviceBuffer(M * K * 8) as dA, device.DeviceBuffer(K * N * 8) as dB, \
            device.DeviceBuffer(M * N * 8) as dC:
        st = L.load().f4j_dgemm_dev_noalloc(0, 0, M, N, K, dA.ptr, K, dB.ptr, N, dC.ptr, N, L.PREC["split_int8"], 0,
                                            None)
        assert st == L.ERR_INVALID_ARG
        assert b"scratch" in L.load().f4j_last_error()
        for prec in ("split_bf16", "split_fp16"):
            st = L.load().f4j_dgemm_dev_noalloc(0, 0, 64, 64, 64, dA.ptr, 64, dB.ptr, 64, dC.ptr, 64, L.PREC[prec], 0,
                                                None)
            assert st == L.ERR_UNSUPPORTED
    # the entry still works after the rejections
    A, B = rng.uniform(65, 128, 96), rng.uniform(66, 96, 160)
    ref, _ = run("f4j_dgemm_dev", "N", "N", A, B, 128, 160, 96, "split_int8")
    got, _ = run("f4j_dgemm_dev_noalloc", "N", "N", A, B, 128, 160, 96, "split_int8")
    assert np.array_equal(got.view(np.int64), ref.view(np.int64))
