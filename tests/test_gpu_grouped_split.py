"""GPU tests of the batched split-precision products (f4j_grouped_plan_create_split):
real batches on BF16/FP16 digits and complex batches in the 3M and 4M forms,
every product against the compensated oracle (Dot2) on sampled entries,
normalised by (|A||B|)_ij (|z| = |re| + |im| for complex entries).

The digits are exact relative to each line's largest entry, so the error is
bounded by ~2^-(digit bits) 2^(eA_i + eB_j); for very short inner sizes
(|A||B|)_ij can sit far below that scale.  Shapes follow the batched workloads
(M, N, K >= 16) and the slice counts carry one digit of margin."""
import ctypes

import numpy as np
import pytest

from flag4j_amd import _native as L
from flag4j_amd import grouped, rng
from oracle import oracle as o

pytestmark = pytest.mark.gpu


def op(X, o_):
    return {"N": X, "T": X.T, "C": X.conj().T}[o_]


def problems(seed, count, lo, hi, opA, opB, cplx, phi=0):
    dims = rng.log_uniform_ints(seed, 3 * count, lo, hi).reshape(count, 3)
    As, Bs = [], []
    for t, (M, N, K) in enumerate(dims):
        sa = (M, K) if opA == "N" else (K, M)
        sb = (K, N) if opB == "N" else (N, K)
        if cplx:
            A = rng.scaled_uniform(seed + 10 + 4 * t, *sa, phi) + 1j * rng.scaled_uniform(seed + 11 + 4 * t, *sa, phi)
            B = rng.scaled_uniform(seed + 12 + 4 * t, *sb, phi) + 1j * rng.scaled_uniform(seed + 13 + 4 * t, *sb, phi)
        else:
            A = rng.scaled_uniform(seed + 10 + 4 * t, *sa, phi)
            B = rng.scaled_uniform(seed + 12 + 4 * t, *sb, phi)
        As.append(A)
        Bs.append(B)
    return As, Bs


def max_err(C, A, B, opA, opB, cplx, n=64, seed=0):
    M, N = C.shape
    if M == 0 or N == 0:
        return 0.0
    r = np.random.default_rng(seed)
    I, J = r.integers(0, M, n), r.integers(0, N, n)
    if cplx:
        rh, rl, ih, il, ab = o.cplx_dot2(A, B, I, J, opA, opB)
        er = np.abs((C[I, J].real - rh) - rl)
        ei = np.abs((C[I, J].imag - ih) - il)
        e = np.maximum(er, ei)
    else:
        Aop, Bop = np.ascontiguousarray(op(A, opA)), np.ascontiguousarray(op(B, opB))
        hi, lo, ab = o.real_dot2(Aop, Bop, I, J)
        e = np.abs((C[I, J] - hi) - lo)
    with np.errstate(invalid="ignore", divide="ignore"):
        q = np.where(ab > 0, e / ab, np.where(e == 0, 0.0, np.inf))
    return float(np.max(q))


@pytest.fixture(autouse=True)
def check_rows():
    L.config_set("splitCheckRows", 2)
    yield
    L.config_set("splitCheckRows", 4)


@pytest.mark.parametrize("prec,S", [("split_bf16", 7), ("split_fp16", 6)])
@pytest.mark.parametrize("opA,opB", [("N", "N"), ("T", "N"), ("N", "T")])
def test_real_batch_accuracy(prec, S, opA, opB):
    As, Bs = problems(200, 60, 16, 300, opA, opB, False)
    Cs, (errs, prods) = grouped.mult_all(As, Bs, opA, opB, precision=prec, slices=S, report=True)
    assert prods == 60 * S * (S + 1) // 2
    for i, (A, B, C) in enumerate(zip(As, Bs, Cs)):
        assert max_err(C, A, B, opA, opB, False) <= 1e-15, i
        assert 0.0 <= errs[i] <= 1e-15, (i, errs[i])


def test_real_small_integers_exact():
    """Integers |x| <= 127 with every row / column maximum 127: one BF16 digit holds
    them exactly, so the batched product is the exact integer product."""
    r = np.random.default_rng(5)
    As, Bs = [], []
    for M, N, K in [(17, 33, 65), (64, 64, 64), (130, 70, 200), (1, 1, 1)]:
        A = r.integers(-127, 128, (M, K)).astype(np.float64)
        B = r.integers(-127, 128, (K, N)).astype(np.float64)
        A[:, 0] = 127.0
        B[0, :] = 127.0
        As.append(A)
        Bs.append(B)
    Cs = grouped.mult_all(As, Bs, precision="split_bf16", slices=1)
    for A, B, C in zip(As, Bs, Cs):
        exact = (A.astype(np.int64) @ B.astype(np.int64)).astype(np.float64)
        assert np.array_equal(C, exact)


@pytest.mark.parametrize("opA,opB", [("C", "N"), ("N", "N"), ("T", "C"), ("C", "C")])
@pytest.mark.parametrize("use3m", [True, False])
def test_complex_batch_accuracy(opA, opB, use3m):
    As, Bs = problems(300, 40, 16, 260, opA, opB, True)
    Cs, (errs, prods) = grouped.mult_all(As, Bs, opA, opB, precision="split_fp16", slices=6, use3m=use3m,
                                         report=True)
    assert prods == 40 * (3 if use3m else 4) * 21
    for i, (A, B, C) in enumerate(zip(As, Bs, Cs)):
        assert max_err(C, A, B, opA, opB, True) <= 1e-15, i
        assert 0.0 <= errs[i] <= 1e-15, (i, errs[i])


def test_complex_wide_range_3m_fp16():
    """A^H B with entries (2r-1) 10^(4 (2r'-1)) (real and imaginary parts), 3M on
    FP16 digits: 8 slices reach the bar on the wide range."""
    As, Bs = problems(400, 30, 32, 300, "C", "N", True, phi=4)
    Cs, (errs, _) = grouped.mult_all(As, Bs, "C", "N", precision="split_fp16", slices=8, report=True)
    worst = max(max_err(C, A, B, "C", "N", True) for A, B, C in zip(As, Bs, Cs))
    assert worst <= 1e-15, worst
    assert np.all((errs >= 0) & (errs <= 1e-15))


def test_error_falls_with_slices():
    As, Bs = problems(500, 20, 16, 200, "N", "N", True)
    L.config_set("splitCheckRows", 0)
    errs = []
    for S in (2, 3, 4, 5):
        Cs, (rep, _) = grouped.mult_all(As, Bs, precision="split_fp16", slices=S, report=True)
        assert np.all(rep == -1.0)                  # check off
        errs.append(max(max_err(C, A, B, "N", "N", True) for A, B, C in zip(As, Bs, Cs)))
    for a, b in zip(errs, errs[1:]):
        assert b <= max(a / 64, 4e-16), errs


def test_empty_and_zero_k_problems():
    As = [np.ones((0, 5)), np.ones((3, 0)), np.random.default_rng(1).random((40, 50)), np.ones((4, 6))]
    Bs = [np.ones((5, 2)), np.ones((0, 4)), np.random.default_rng(2).random((50, 30)), np.ones((6, 0))]
    Cs, (errs, _) = grouped.mult_all(As, Bs, precision="split_bf16", report=True)
    assert Cs[0].shape == (0, 2) and Cs[3].shape == (4, 0)
    assert np.all(Cs[1] == 0.0) and Cs[1].shape == (3, 4)
    assert max_err(Cs[2], As[2], Bs[2], "N", "N", False) <= 1e-15
    assert errs[0] == -1.0 and errs[1] == -1.0 and errs[3] == -1.0 and 0 <= errs[2] <= 1e-15


@pytest.mark.parametrize("prec,S", [("split_fp16", 6), ("split_int8", 0)])
def test_plan_replay_is_stable(prec, S):
    """The captured graph replayed several times, with other device work and
    syncs in between, gives the same bits every time (no stale scratch state)."""
    from flag4j_amd import device
    As, Bs = problems(600, 25, 16, 200, "N", "N", True)
    shapes = [(A.shape[0], B.shape[1], A.shape[1]) for A, B in zip(As, Bs)]
    dA = [device.DeviceBuffer.from_array(np.ascontiguousarray(A)) for A in As]
    dB = [device.DeviceBuffer.from_array(np.ascontiguousarray(B)) for B in Bs]
    dC = [device.DeviceBuffer(m * n * 16) for m, n, _ in shapes]
    X = np.random.default_rng(9).random((300, 300))
    dX, dY = device.DeviceBuffer.from_array(X), device.DeviceBuffer(X.nbytes)
    plan = grouped.GroupedPlan(shapes, [a.ptr for a in dA], [b.ptr for b in dB], [c.ptr for c in dC],
                               [s[2] for s in shapes], [s[1] for s in shapes], [s[1] for s in shapes],
                               "N", "N", True, prec, S, True)
    outs = []
    nan = np.full(max(m * n for m, n, _ in shapes), np.nan + 1j * np.nan)
    try:
        for rep in range(3):
            for c, (m, n, _) in zip(dC, shapes):
                c.upload(nan[:m * n])
            device.sync()
            plan.run()
            device.sync()
            outs.append([c.download(np.complex128, m * n).reshape(m, n) for c, (m, n, _) in zip(dC, shapes)])
            device.dgemm_dev("N", "N", 300, 300, 300, dX.ptr, 300, dX.ptr, 300, dY.ptr, 300)   # other work
            device.sync()
    finally:
        plan.close()
        for b in dA + dB + dC + [dX, dY]:
            b.close()
    for a, b in zip(outs[0], outs[1]):
        assert np.array_equal(a.view(np.float64), b.view(np.float64))
    for a, b in zip(outs[0], outs[2]):
        assert np.array_equal(a.view(np.float64), b.view(np.float64))
    for (A, B), C in zip(zip(As, Bs), outs[0]):
        assert max_err(C, A, B, "N", "N", True) <= 1e-15


def test_create_with_split_precision_defaults():
    """f4j_grouped_plan_create with a SPLIT_* precision builds a split plan
    (default slices: 7 for BF16, 28 pairs per real product)."""
    from flag4j_amd import device
    As, Bs = problems(700, 10, 16, 100, "N", "N", False)
    dA = [device.DeviceBuffer.from_array(np.ascontiguousarray(A)) for A in As]
    dB = [device.DeviceBuffer.from_array(np.ascontiguousarray(B)) for B in Bs]
    dC = [device.DeviceBuffer(A.shape[0] * B.shape[1] * 8) for A, B in zip(As, Bs)]
    arr = lambda v: np.ascontiguousarray(np.asarray(v, dtype=np.int64))
    M, N, K = arr([A.shape[0] for A in As]), arr([B.shape[1] for B in Bs]), arr([A.shape[1] for A in As])
    pa, pb, pc = (np.ascontiguousarray(np.array([t.ptr for t in ts], dtype=np.uint64)) for ts in (dA, dB, dC))
    vp = lambda a: ctypes.c_void_p(a.ctypes.data)
    h = ctypes.c_void_p()
    lib = L.load()
    try:
        L.check(lib.f4j_grouped_plan_create(0, 0, 0, len(As), vp(M), vp(N), vp(K), vp(pa), vp(K), vp(pb), vp(N),
                                            vp(pc), vp(N), L.PREC["split_bf16"], ctypes.byref(h)))
        L.check(lib.f4j_grouped_plan_run(h, None))
        errs = np.zeros(len(As))
        prods = ctypes.c_int()
        L.check(lib.f4j_grouped_plan_split_report(h, ctypes.c_void_p(errs.ctypes.data), len(As), ctypes.byref(prods)))
        assert prods.value == 10 * 28
        assert np.all((errs >= 0) & (errs <= 1e-15))
        for A, B, c in zip(As, Bs, dC):
            C = c.download(np.float64, A.shape[0] * B.shape[1]).reshape(A.shape[0], B.shape[1])
            assert max_err(C, A, B, "N", "N", False) <= 1e-15
    finally:
        if h.value:
            L.check(lib.f4j_grouped_plan_destroy(h))
        for b in dA + dB + dC:
            b.close()


def test_nonfinite_operands_are_reported():
    """A product whose operands hold NaN or Inf is reported with a NaN error (its
    digit-slice result is not meaningful); the other products are unaffected."""
    As, Bs = problems(700, 3, 32, 96, "N", "N", False)
    As[1] = As[1].copy()
    As[1][0, 0] = np.nan
    Cs, (errs, _) = grouped.mult_all(As, Bs, "N", "N", "split_bf16", 7, report=True)
    assert np.isnan(errs[1])
    assert np.all(np.isfinite(errs[[0, 2]]))
    for t in (0, 2):
        assert max_err(Cs[t], As[t], Bs[t], "N", "N", False) <= 1e-15


# ---- modular form (SPLIT_INT8): residues modulo coprime moduli on the INT8 matrix cores ----

def moduli_of(prods, jobs):
    assert prods % jobs == 0
    return prods // jobs


@pytest.mark.parametrize("opA,opB", [("N", "N"), ("T", "N"), ("N", "T"), ("T", "T")])
def test_int8_real_batch_accuracy(opA, opB):
    As, Bs = problems(800, 60, 16, 300, opA, opB, False)
    Cs, (errs, prods) = grouped.mult_all(As, Bs, opA, opB, precision="split_int8", report=True)
    assert moduli_of(prods, 60) == 18                 # the fewest giving t >= 64 at K <= 300
    for i, (A, B, C) in enumerate(zip(As, Bs, Cs)):
        assert max_err(C, A, B, opA, opB, False) <= 1e-15, i
        assert 0.0 <= errs[i] <= 1e-15, (i, errs[i])


@pytest.mark.parametrize("opA,opB", [("C", "N"), ("N", "N"), ("T", "C")])
@pytest.mark.parametrize("use3m", [True, False])
def test_int8_complex_batch_accuracy(opA, opB, use3m):
    As, Bs = problems(900, 40, 16, 260, opA, opB, True)
    Cs, (errs, prods) = grouped.mult_all(As, Bs, opA, opB, precision="split_int8", use3m=use3m, report=True)
    assert moduli_of(prods, 40 * (3 if use3m else 4)) == 18
    for i, (A, B, C) in enumerate(zip(As, Bs, Cs)):
        assert max_err(C, A, B, opA, opB, True) <= 1e-15, i
        assert 0.0 <= errs[i] <= 1e-15, (i, errs[i])


def test_int8_complex_wide_range_3m():
    """A^H B with parts (2r-1) 10^(4 (2r'-1)), 3M: the small entries of a line keep
    t - 27 bits, so the wide range takes t = 58 (17 moduli) to reach the bar, as
    the digit form takes 8 slices instead of its default."""
    As, Bs = problems(400, 30, 32, 300, "C", "N", True, phi=4)
    Cs, (errs, _) = grouped.mult_all(As, Bs, "C", "N", precision="split_int8", slices=17, report=True)
    worst = max(max_err(C, A, B, "C", "N", True) for A, B, C in zip(As, Bs, Cs))
    assert worst <= 1e-15, worst
    assert np.all((errs >= 0) & (errs <= 1e-15))


def test_int8_small_integers_exact():
    """Integers |x| <= 127: scaling by a power of two and the residue products are
    exact, and the reconstruction returns the exact integer product."""
    r = np.random.default_rng(6)
    As, Bs = [], []
    for M, N, K in [(17, 33, 65), (64, 64, 64), (130, 70, 200), (1, 1, 1), (3, 200, 5)]:
        As.append(r.integers(-127, 128, (M, K)).astype(np.float64))
        Bs.append(r.integers(-127, 128, (K, N)).astype(np.float64))
    Cs = grouped.mult_all(As, Bs, precision="split_int8")
    for A, B, C in zip(As, Bs, Cs):
        exact = (A.astype(np.int64) @ B.astype(np.int64)).astype(np.float64)
        assert np.array_equal(C, exact)


def test_int8_moduli_bounds():
    A, B = np.ones((4, 4)), np.ones((4, 4))
    for bad in (1, 21):
        with pytest.raises(L.NativeError):
            grouped.mult_all([A], [B], precision="split_int8", slices=bad)
    Cs, (_, prods) = grouped.mult_all([A], [B], precision="split_int8", slices=20, report=True)
    assert prods == 20 and np.array_equal(Cs[0], A @ B)


def test_int8_empty_zero_k_and_nonfinite():
    g = np.random.default_rng(3)
    As = [np.ones((0, 5)), np.ones((3, 0)), g.random((40, 50)), g.random((70, 20)), g.random((9, 9))]
    Bs = [np.ones((5, 2)), np.ones((0, 4)), g.random((50, 30)), g.random((20, 65)), g.random((9, 9))]
    As[4][2, 3] = np.inf
    Cs, (errs, _) = grouped.mult_all(As, Bs, precision="split_int8", report=True)
    assert Cs[0].shape == (0, 2) and np.all(Cs[1] == 0.0)
    for t in (2, 3):
        assert max_err(Cs[t], As[t], Bs[t], "N", "N", False) <= 1e-15
        assert 0 <= errs[t] <= 1e-15
    assert errs[0] == -1.0 and errs[1] == -1.0 and np.isnan(errs[4])


# ---- FP64 recomputation inside the plan (splitFallback = 1) ----

def test_fallback_recomputes_rejected_products_in_fp64():
    """With splitFallback = 1 a split plan recomputes in FP64, inside its own graph,
    the products whose checked error is above splitTolerance: those equal the FP64
    grouped plan's result bit for bit, the others keep the split result bit for bit,
    on every replay of the graph."""
    from flag4j_amd import device
    As, Bs = problems(1100, 30, 16, 200, "N", "N", False)
    ref = grouped.mult_all(As, Bs, precision="fp64")
    L.config_set("splitFallback", 0)
    try:
        split, (errs, _) = grouped.mult_all(As, Bs, precision="split_bf16", slices=3, report=True)
    finally:
        L.config_set("splitFallback", 1)
    tol = float(np.median(errs))
    assert np.sum(errs > tol) >= 5 and np.sum(errs <= tol) >= 5
    dA = [device.DeviceBuffer.from_array(np.ascontiguousarray(A)) for A in As]
    dB = [device.DeviceBuffer.from_array(np.ascontiguousarray(B)) for B in Bs]
    dC = [device.DeviceBuffer(A.shape[0] * B.shape[1] * 8) for A, B in zip(As, Bs)]
    L.config_set("splitTolerance", repr(tol))
    try:
        plan = grouped.GroupedPlan([(A.shape[0], B.shape[1], A.shape[1]) for A, B in zip(As, Bs)],
                                   [t.ptr for t in dA], [t.ptr for t in dB], [t.ptr for t in dC],
                                   [A.shape[1] for A in As], [B.shape[1] for B in Bs], [B.shape[1] for B in Bs],
                                   precision="split_bf16", slices=3)
    finally:
        L.config_set("splitTolerance", 1e-15)
    try:
        for rep in range(3):
            for c in dC:                       # stale C must not survive: refill with NaN
                c.upload(np.full(c.nbytes // 8, np.nan))
            plan.run()
            device.sync()
            e2, _ = plan.split_report()
            assert np.array_equal(e2, errs), rep
            for t, (A, B) in enumerate(zip(As, Bs)):
                C = dC[t].download(np.float64, A.shape[0] * B.shape[1]).reshape(A.shape[0], B.shape[1])
                exp = ref[t] if errs[t] > tol else split[t]
                assert np.array_equal(C, exp), (rep, t, errs[t], tol)
    finally:
        plan.close()
        for b in dA + dB + dC:
            b.close()


def test_default_slices_on_wide_range():
    """Default counts on A^H B with parts spread over 10^+-4 (3M): the modular
    form's default (t >= 64) meets 1e-15 by itself; the digit forms' defaults
    meet it after the plan's FP64 recomputation of the products their check
    rejects, and those products equal the FP64 plan's bits."""
    As, Bs = problems(1200, 40, 32, 300, "C", "N", True, phi=4)
    L.config_set("splitFallback", 0)
    try:
        Cs, (errs, prods) = grouped.mult_all(As, Bs, "C", "N", precision="split_int8", report=True)
    finally:
        L.config_set("splitFallback", 1)
    assert prods == 40 * 3 * 18
    assert np.all((errs >= 0) & (errs <= 1e-15))
    assert max(max_err(C, A, B, "C", "N", True) for A, B, C in zip(As, Bs, Cs)) <= 1e-15
    ref = grouped.mult_all(As, Bs, "C", "N", precision="fp64")
    for prec in ("split_bf16", "split_fp16"):
        Cs, (errs, _) = grouped.mult_all(As, Bs, "C", "N", precision=prec, report=True)
        assert np.all(errs >= 0), prec
        for t, (A, B, C) in enumerate(zip(As, Bs, Cs)):
            if errs[t] > 1e-15:
                assert np.array_equal(C.view(np.float64), ref[t].view(np.float64)), (prec, t)
            else:
                assert max_err(C, A, B, "C", "N", True) <= 1e-15, (prec, t)
