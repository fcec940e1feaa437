"""The C ABI without a GPU: exported symbols, argument validation, configuration,
and the dispatcher's decisions (none of these touch the device)."""
import ctypes
import os
import re

import numpy as np
import pytest

import flag4j_amd as f4j
from flag4j_amd import _native
from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "flag4j_amd.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(f4j_\w+)\s*\(", text)))


def test_exports_every_declared_symbol():
    lib = _native.load()
    syms = declared_symbols()
    assert len(syms) >= 18
    for s in syms:
        assert hasattr(lib, s), s
    assert set(syms) == set(_native.SIGNATURES)


def test_abi_version():
    assert _native.load().f4j_abi_version() == 4


def _dgemm(*args):
    return _native.load().f4j_dgemm(*args)


def test_invalid_arguments_rejected_before_device_use():
    a = np.zeros(4)
    p = ctypes.c_void_p(a.ctypes.data)
    L = _native
    assert _dgemm(7, 0, 2, 2, 2, p, 2, p, 2, p, 2, 0, 0) == L.ERR_INVALID_ARG      # bad op
    assert _dgemm(2, 0, 2, 2, 2, p, 2, p, 2, p, 2, 0, 0) == L.ERR_INVALID_ARG      # op C on real
    assert _dgemm(0, 0, -1, 2, 2, p, 2, p, 2, p, 2, 0, 0) == L.ERR_INVALID_ARG     # negative size
    assert _dgemm(0, 0, 2, 2, 2, None, 2, p, 2, p, 2, 0, 0) == L.ERR_INVALID_ARG   # null A
    assert _dgemm(0, 0, 2, 2, 3, p, 2, p, 2, p, 2, 0, 0) == L.ERR_INVALID_ARG      # lda < K
    assert _dgemm(0, 0, 2, 3, 2, p, 2, p, 3, p, 2, 0, 0) == L.ERR_INVALID_ARG      # ldc < N
    assert b"ldc" in _native.load().f4j_last_error()
    assert _dgemm(0, 0, 2, 2, 2, p, 2, p, 2, p, 2, 9, 0) == L.ERR_INVALID_ARG      # precision
    assert _native.load().f4j_zgemm(0, 3, 2, 2, 2, p, 2, p, 2, p, 2, 0, 0, 0) == L.ERR_INVALID_ARG


def test_lstsq_arguments_rejected_before_device_use():
    """f4j_dlstsq validates before it touches the device: R of the reduced QR must be
    square (ensureSquare, as RealBackSolver requires: M >= N >= 1), sizes non-negative,
    pointers present, leading dimensions large enough, the precision code known."""
    a = np.zeros(64)
    p = ctypes.c_void_p(a.ctypes.data)
    L = _native
    f = _native.load().f4j_dlstsq
    assert f(3, 4, 1, p, 4, p, 1, p, 1, -1) == L.ERR_SHAPE                 # M < N
    assert f(3, 0, 1, p, 1, p, 1, p, 1, -1) == L.ERR_SHAPE                 # N = 0
    assert b"square" in _native.load().f4j_last_error()
    assert f(-1, 2, 1, p, 2, p, 1, p, 1, -1) == L.ERR_INVALID_ARG
    assert f(4, 2, 1, None, 2, p, 1, p, 1, -1) == L.ERR_INVALID_ARG        # null A
    assert f(4, 2, 1, p, 1, p, 1, p, 1, -1) == L.ERR_INVALID_ARG           # lda < N
    assert f(4, 2, 3, p, 2, p, 2, p, 3, -1) == L.ERR_INVALID_ARG           # ldb < R
    assert f(4, 2, 1, p, 2, None, 1, p, 1, -1) == L.ERR_INVALID_ARG        # null B
    assert f(4, 2, 1, p, 2, p, 1, p, 1, 9) == L.ERR_INVALID_ARG            # precision
    assert _native.load().f4j_dlstsq_dev(3, 4, 1, p, 4, p, 1, p, 1, -1, None) == L.ERR_SHAPE


def test_pinv_arguments_rejected_before_device_use():
    """f4j_dpinv validates before it touches the device: non-empty shape, pointers present,
    leading dimensions large enough (X is N x M: ldx >= M)."""
    a = np.zeros(64)
    p = ctypes.c_void_p(a.ctypes.data)
    L = _native
    f = _native.load().f4j_dpinv
    assert f(0, 3, p, 3, p, 1) == L.ERR_SHAPE
    assert f(3, 0, p, 1, p, 3) == L.ERR_SHAPE
    assert f(-1, 2, p, 2, p, 1) == L.ERR_INVALID_ARG
    assert f(4, 2, None, 2, p, 4) == L.ERR_INVALID_ARG                    # null A
    assert f(4, 2, p, 2, None, 4) == L.ERR_INVALID_ARG                    # null X
    assert f(4, 2, p, 1, p, 4) == L.ERR_INVALID_ARG                       # lda < N
    assert f(4, 2, p, 2, p, 3) == L.ERR_INVALID_ARG                       # ldx < M
    assert _native.load().f4j_dpinv_dev(0, 3, p, 3, p, 1, None) == L.ERR_SHAPE
    with pytest.raises(f4j.IllegalArgumentException):
        f4j.Invert.pInv(f4j.Matrix(np.ones((0, 3))))


def test_lstsq_solver_errors_match_reference_exceptions():
    """RealLstsqSolver.solve: Qh.mult(b) with b's rows != A's rows -> LinearAlgebraException;
    a wide A leaves R non-square -> IllegalArgumentException (ValidateParameters.ensureSquare)."""
    s = f4j.RealLstsqSolver()
    with pytest.raises(f4j.LinearAlgebraException):
        s.solve(f4j.Matrix(np.ones((5, 3))), f4j.Vector(np.ones(4)))
    with pytest.raises(f4j.LinearAlgebraException):
        s.solve(f4j.Matrix(np.ones((5, 3))), f4j.Matrix(np.ones((3, 2))))
    with pytest.raises(f4j.IllegalArgumentException):
        s.solve(f4j.Matrix(np.ones((3, 5))), f4j.Vector(np.ones(3)))


def test_empty_products_need_no_device():
    a = np.zeros(1)
    p = ctypes.c_void_p(a.ctypes.data)
    assert _dgemm(0, 0, 0, 5, 5, None, 5, p, 5, p, 5, 0, 0) == 0
    assert _dgemm(0, 0, 5, 0, 5, p, 5, None, 1, None, 1, 0, 0) == 0


def test_shape_errors_match_reference_exceptions():
    """Matrix.mult -> LinearAlgebraException (MatrixMultTests.java:61-62);
    multTranspose -> IllegalArgumentException (MatrixMultTests.java:281-282);
    CMatrix.mult -> LinearAlgebraException (CMatrixMultTests.java:100-101)."""
    A = f4j.Matrix(np.ones((4, 3)))
    with pytest.raises(f4j.LinearAlgebraException):
        A.mult(f4j.Matrix(np.ones((4, 2))))
    with pytest.raises(f4j.IllegalArgumentException):
        A.multTranspose(f4j.Matrix(np.ones((2, 4))))
    C = f4j.CMatrix(np.ones((4, 3), dtype=complex))
    with pytest.raises(f4j.LinearAlgebraException):
        C.mult(f4j.CMatrix(np.ones((2, 2), dtype=complex)))
    with pytest.raises(f4j.LinearAlgebraException):
        f4j.Matrix(np.ones((3, 2))).pow(2)
    with pytest.raises(f4j.IllegalArgumentException):
        f4j.Matrix(np.ones((3, 3))).pow(-1)
    with pytest.raises(f4j.IllegalArgumentException):
        f4j.CMatrix(np.ones((2, 3), dtype=complex)).pow(2)
    with pytest.raises(f4j.IllegalArgumentException):
        f4j.Matrix(f4j.Shape(2, 3), np.ones(5))
    with pytest.raises(ArithmeticError):
        f4j.Shape(70000, 70000)


def _name(cplx, opA, opB, M, N, K, prec="auto"):
    buf = ctypes.create_string_buffer(64)
    _native.check(_native.load().f4j_dispatch_name(int(cplx), opA, opB, M, N, K, _native.PREC[prec], buf, 64))
    return buf.value.decode()


def test_dispatcher_small_path_threshold():
    """Products with A.len + B.len < smallThreshold take the small path
    (ReDeMatMultDispatcher.java:159-168); here that is the bitwise-strict kernel."""
    assert _name(0, 0, 0, 4, 2, 3) == "STRICT"
    assert _name(0, 0, 0, 34, 35, 35) == "STRICT"             # 34*35 + 35*35 = 2415 < 2450
    assert _name(0, 0, 0, 35, 35, 35).startswith("MFMA")      # 2450 entries: not small
    assert _name(0, 0, 0, 512, 512, 512).startswith("MFMA_F64")
    assert _name(1, 2, 0, 300, 300, 300).startswith("MFMA_F64")
    assert _name(0, 0, 0, 8192, 8192, 8192).startswith("MFMA_F64_128x128x16")
    assert _name(0, 0, 0, 0, 3, 3) == "EMPTY"
    assert _name(0, 0, 0, 3, 3, 0) == "ZERO_FILL"
    assert _name(0, 0, 0, 4, 2, 3, "fp64").startswith("MFMA_F64")


def test_dispatcher_matrix_vector_routes():
    """N = 1 with A stored M x K is a matrix-vector product: the reference picks
    standard (small) or the 64-block vector kernels (ReDeMatMultDispatcher.java:158-198),
    multTranspose a k-ascending sum (RealDenseMatMultTranspose.java:53-82), complex
    products a k-ascending sum (DenseSemiringMatMult.java:435-574)."""
    assert _name(0, 0, 0, 40, 1, 40) == "GEMV_CHAIN"             # 1600 + 40 < 2450: small path
    assert _name(0, 0, 0, 100, 1, 100) == "GEMV_BLOCKED"
    assert _name(0, 0, 0, 100, 1, 100, "strict") == "GEMV_BLOCKED"
    assert _name(0, 0, 0, 100, 1, 100, "fp64") == "GEMV_BLOCKED"
    assert _name(0, 0, 1, 100, 1, 100) == "GEMV_CHAIN"           # multTranspose
    assert _name(1, 0, 0, 100, 1, 100) == "GEMV_COMPLEX"
    assert _name(1, 0, 2, 100, 1, 100, "strict") == "GEMV_COMPLEX"
    assert _name(0, 1, 0, 100, 1, 100).startswith("MFMA")        # op(A) = T: not the vector path
    assert _name(0, 0, 0, 100, 1, 100, "split_int8") == "SPLIT_split_int8"


def test_gemv_arguments_checked_before_device_use():
    L = _native.load()
    a = np.zeros(8)
    p = ctypes.c_void_p(a.ctypes.data)
    assert L.f4j_dgemv(-1, 2, p, 2, p, 1, p, 1, 0) == _native.ERR_INVALID_ARG
    assert L.f4j_dgemv(2, 3, p, 2, p, 1, p, 1, 0) == _native.ERR_INVALID_ARG     # lda < K
    assert L.f4j_dgemv(2, 2, p, 2, p, 0, p, 1, 0) == _native.ERR_INVALID_ARG     # incx < 1
    assert L.f4j_dgemv(2, 2, None, 2, p, 1, p, 1, 0) == _native.ERR_INVALID_ARG
    assert L.f4j_zgemv(2, 2, p, 2, p, 1, None, 1, 0) == _native.ERR_INVALID_ARG
    assert L.f4j_dgemv(2, 2, p, 2, p, 1, p, 1, 9) == _native.ERR_INVALID_ARG
    assert L.f4j_dgemv(2, 2, p, 2, p, 1, p, 1, _native.PREC["split_bf16"]) == _native.ERR_UNSUPPORTED
    assert b"FP64" in L.f4j_last_error()
    assert L.f4j_dgemv(0, 5, None, 5, None, 1, None, 1, 0) == 0                 # empty: no device


def test_config_keys_validated():
    L = _native.load()
    buf = ctypes.create_string_buffer(32)
    assert L.f4j_config_set(b"noSuchKey", b"1") == _native.ERR_INVALID_ARG
    assert L.f4j_config_set(b"smallThreshold", b"abc") == _native.ERR_INVALID_ARG
    assert L.f4j_config_set(b"defaultPrecision", b"fp128") == _native.ERR_INVALID_ARG
    assert L.f4j_config_set(b"splitTiming", b"2") == _native.ERR_INVALID_ARG
    _native.check(L.f4j_config_set(b"splitTiming", b"1"))
    _native.check(L.f4j_config_get(b"splitTiming", buf, 32))
    assert buf.value == b"1"
    _native.check(L.f4j_config_set(b"splitTiming", b"0"))
    _native.check(L.f4j_config_get(b"smallThreshold", buf, 32))
    old = buf.value
    try:
        _native.check(L.f4j_config_set(b"smallThreshold", b"0"))
        assert _name(0, 0, 0, 4, 2, 3).startswith("MFMA")
    finally:
        _native.check(L.f4j_config_set(b"smallThreshold", old))


def test_config_file_all_or_nothing(tmp_path):
    L = _native.load()
    good = tmp_path / "GpuMatMult.properties"
    good.write_text("# comment\nsmallThreshold = 2450\ndefaultPrecision=fp64\n")
    _native.check(L.f4j_config_load(str(good).encode()))
    bad = tmp_path / "bad.properties"
    bad.write_text("smallThreshold = 10\nsquareStandardVecThreshold = 256\n")
    assert L.f4j_config_load(str(bad).encode()) == _native.ERR_INVALID_ARG
    assert b"bad.properties:2" in L.f4j_last_error()
    buf = ctypes.create_string_buffer(32)
    _native.check(L.f4j_config_get(b"smallThreshold", buf, 32))
    assert buf.value == b"2450"


def test_strict_kernel_has_no_fused_multiply_add():
    """The strict kernel must round every product and sum on its own (Java
    semantics): its gfx950 assembly may hold no v_fma_f64."""
    import subprocess
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "flag4j_amd", "csrc"), "strict-asm"], check=True)
    asm = open(os.path.join(ROOT, "_build", "obj", "gemm_strict.s")).read()
    assert "v_mul_f64" in asm and "v_add_f64" in asm
    assert "v_fma_f64" not in asm
    gemv = open(os.path.join(ROOT, "_build", "obj", "gemv.s")).read()
    assert "gemv_dma_kernel" in gemv and "v_fma_f64" not in gemv
    # The least-squares kernels: fused multiply-adds only inside the correctly rounded
    # division (v_rcp_f64 .. v_div_fixup_f64) and square root (v_rsq_f64 .. v_ldexp_f64)
    # expansions, never in the reference's products and sums.
    lines = open(os.path.join(ROOT, "_build", "obj", "lstsq.s")).read().splitlines()
    depth, outside, kernels = 0, [], set()
    for ln in lines:
        t = ln.strip()
        if t.startswith("_ZN3f4j") and ":" in t:
            kernels.add(t.split(":")[0])
        if t.startswith("v_rcp_f64") or t.startswith("v_rsq_f64"):
            depth += 1
        elif (t.startswith("v_div_fixup_f64") or t.startswith("v_ldexp_f64")) and depth > 0:
            depth -= 1
        elif depth == 0 and (t.startswith("v_fma_f64") or t.startswith("v_fmac_f64")):
            outside.append(t)
    assert any("hh_dot_kernel" in k for k in kernels) and any("back_solve_kernel" in k for k in kernels)
    assert not outside, outside


def test_tuner_cache_file_roundtrip(tmp_path):
    L = _native.load()
    good = tmp_path / "tune.txt"
    good.write_text("# cache\ngfx950 0,0 2 7\ngfx950 3,2 3 4\n")
    _native.check(L.f4j_tuner_load(str(good).encode()))
    out = tmp_path / "out.txt"
    _native.check(L.f4j_tuner_save(str(out).encode()))
    lines = [l for l in out.read_text().splitlines() if not l.startswith("#")]
    assert "gfx950 0,0 2 7" in lines and "gfx950 3,2 3 4" in lines
    bad = tmp_path / "bad.txt"
    bad.write_text("gfx950 0,0 9 7\n")                   # bucket out of range
    assert L.f4j_tuner_load(str(bad).encode()) == _native.ERR_INVALID_ARG
    bad.write_text("gfx950 0,0 1 99\n")                  # unknown tile config
    assert L.f4j_tuner_load(str(bad).encode()) == _native.ERR_INVALID_ARG
    _native.check(L.f4j_tuner_clear())


def test_grouped_plan_argument_checks():
    import ctypes
    L = _native.load()
    h = ctypes.c_void_p()
    assert L.f4j_grouped_plan_create(0, 2, 0, 1, None, None, None, None, None, None, None, None, None, 0,
                                     ctypes.byref(h)) == _native.ERR_INVALID_ARG
    assert L.f4j_grouped_plan_create(0, 0, 0, -1, None, None, None, None, None, None, None, None, None, 0,
                                     ctypes.byref(h)) == _native.ERR_INVALID_ARG
    assert L.f4j_grouped_plan_create(0, 0, 0, 1, None, None, None, None, None, None, None, None, None, 0,
                                     ctypes.byref(h)) == _native.ERR_INVALID_ARG


def test_tuner_cache_config_key(tmp_path):
    """tunerCache names a tuning cache that is read when the key is set (and at
    initialisation): entries land in the process-wide cache; a missing file is a
    first run, not an error; the key round-trips and loads from a properties file."""
    L = _native.load()
    _native.check(L.f4j_tuner_clear())
    cache = tmp_path / "tune.txt"
    cache.write_text("gfx950 1,1 0 8\n")
    try:
        _native.config_set("tunerCache", str(cache))
        assert _native.config_get("tunerCache") == str(cache)
        out = tmp_path / "out.txt"
        _native.check(L.f4j_tuner_save(str(out).encode()))
        assert "gfx950 1,1 0 8" in out.read_text().splitlines()
        _native.config_set("tunerCache", str(tmp_path / "missing.txt"))      # not an error
        props = tmp_path / "c.properties"
        props.write_text(f"tunerCache = {cache}\n")
        _native.check(L.f4j_config_load(str(props).encode()))
        assert _native.config_get("tunerCache") == str(cache)
    finally:
        _native.config_set("tunerCache", "")
        _native.check(L.f4j_tuner_clear())
