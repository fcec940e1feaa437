"""GPU tests of the grouped (batched) products: every product against the
oracle within 2 K u (|A||B|), and invariants of the plan (replay, tuning,
agreement with the single-product path)."""
import numpy as np
import pytest

import flag4j_amd as f4j
from flag4j_amd import _native as L
from flag4j_amd import device, grouped, rng
from oracle import oracle as o

pytestmark = pytest.mark.gpu
U = 2.0**-53


def rand_problems(seed, count, lo, hi, opA="N", opB="N", cplx=False):
    dims = rng.log_uniform_ints(seed, 3 * count, lo, hi).reshape(count, 3)
    g = rng.next_gaussians(seed + 1, int(np.sum(dims[:, 0] * dims[:, 2] + dims[:, 2] * dims[:, 1])) * (2 if cplx else 1))
    As, Bs, pos = [], [], 0
    w = 2 if cplx else 1
    for M, N, K in dims:
        a = g[pos:pos + M * K * w]
        pos += M * K * w
        b = g[pos:pos + K * N * w]
        pos += K * N * w
        if cplx:
            a, b = a.view(np.complex128), b.view(np.complex128)
        A = a.reshape((M, K) if opA == "N" else (K, M))
        B = b.reshape((K, N) if opB == "N" else (N, K))
        As.append(A)
        Bs.append(B)
    return As, Bs


def op(X, o_):
    return {"N": X, "T": X.T, "C": X.conj().T}[o_]


def check_real(C, A, B):
    ref = o.real_reordered(np.ascontiguousarray(A), np.ascontiguousarray(B))
    K = A.shape[1]
    return bool(np.all(np.abs(C - ref) <= 2 * K * U * (np.abs(A) @ np.abs(B))))


@pytest.mark.parametrize("opA,opB", [("N", "N"), ("N", "T"), ("T", "N")])
def test_grouped_real_against_oracle(opA, opB):
    As, Bs = rand_problems(100, 400, 1, 300, opA, opB)
    Cs = grouped.mult_all(As, Bs, opA, opB)
    for A, B, C in zip(As, Bs, Cs):
        assert check_real(C, op(A, opA), op(B, opB))


def test_grouped_complex_hermitian():
    """A_i^H B_i, the batched complex form (the reference forms A.H().mult(B))."""
    As, Bs = rand_problems(200, 120, 1, 200, "C", "N", cplx=True)
    Cs = grouped.mult_all(As, Bs, "C", "N")
    for A, B, C in zip(As, Bs, Cs):
        ref = o.cplx_gemm(A, B, "C", "N")
        nA = np.abs(A.real) + np.abs(A.imag)
        nB = np.abs(B.real) + np.abs(B.imag)
        lim = 2 * 2 * A.shape[0] * U * (nA.T @ nB)
        assert np.all(np.abs(C.real - ref.real) <= lim) and np.all(np.abs(C.imag - ref.imag) <= lim)


def test_grouped_empty_and_degenerate():
    As = [np.zeros((0, 5)), np.ones((3, 0)), np.ones((2, 3)), np.ones((4, 4))]
    Bs = [np.ones((5, 2)), np.ones((0, 6)), np.ones((3, 0)), np.ones((4, 1))]
    Cs = grouped.mult_all(As, Bs)
    assert Cs[0].shape == (0, 2)
    assert Cs[1].shape == (3, 6) and np.all(Cs[1] == 0.0)     # K = 0: C is the zero matrix
    assert Cs[2].shape == (2, 0)
    assert np.all(Cs[3] == 4.0)
    assert grouped.mult_all([], []) == []


def test_grouped_equals_single_product_bitwise():
    """The MFMA accumulation order per element does not depend on the tile
    shape, so a grouped product equals the single-product call exactly."""
    As, Bs = rand_problems(300, 60, 8, 400)
    Cs = grouped.mult_all(As, Bs)
    for A, B, C in zip(As, Bs, Cs):
        single = f4j.Matrix(A).mult(f4j.Matrix(B), precision="fp64").numpy()
        assert np.array_equal(C, single)


def test_plan_replay_tune_and_refill():
    As, Bs = rand_problems(400, 300, 16, 512)
    shapes = [(a.shape[0], b.shape[1], a.shape[1]) for a, b in zip(As, Bs)]
    pa = np.concatenate([a.reshape(-1) for a in As])
    pb = np.concatenate([b.reshape(-1) for b in Bs])
    oa = np.cumsum([0] + [a.size for a in As])
    ob = np.cumsum([0] + [b.size for b in Bs])
    oc = np.cumsum([0] + [m * n for m, n, _ in shapes])
    with device.DeviceBuffer.from_array(pa) as dA, device.DeviceBuffer.from_array(pb) as dB, \
            device.DeviceBuffer(int(oc[-1]) * 8) as dC:
        plan = grouped.GroupedPlan(shapes, [dA.ptr + 8 * int(x) for x in oa[:-1]],
                                   [dB.ptr + 8 * int(x) for x in ob[:-1]], [dC.ptr + 8 * int(x) for x in oc[:-1]],
                                   [a.shape[1] for a in As], [b.shape[1] for b in Bs], [n for _, n, _ in shapes])
        info = plan.info()
        assert sum(info["tiles"]) > 0 and info["flops"] > 0
        plan.run()
        device.sync()
        c1 = dC.download(np.float64, int(oc[-1]))
        plan.run()
        device.sync()
        assert np.array_equal(dC.download(np.float64, int(oc[-1])), c1)
        plan.tune(reps=1)
        plan.run()
        device.sync()
        assert np.array_equal(dC.download(np.float64, int(oc[-1])), c1)
        # refill the same buffers: the captured graph reads the new operands
        dA.upload(-pa)
        plan.run()
        device.sync()
        assert np.array_equal(dC.download(np.float64, int(oc[-1])), -c1)
        plan.close()
    for i in (0, 17, 299):
        C = c1[oc[i]:oc[i + 1]].reshape(shapes[i][0], shapes[i][1])
        assert check_real(C, As[i], Bs[i])


def test_tuner_cache_read_at_init_and_written_by_tune(tmp_path):
    """A tuning cache named by tunerCache is read at f4j_init (a plan then takes
    the cached tile configuration) and rewritten after f4j_grouped_plan_tune."""
    As, Bs = rand_problems(410, 40, 16, 60)           # small products: the 64 x 64 bucket (0)
    lib = L.load()
    cache = tmp_path / "tune.txt"
    cache.write_text("gfx950 0,0 0 7\n")
    L.check(lib.f4j_tuner_clear())
    try:
        L.config_set("tunerCache", str(cache))
        L.check(lib.f4j_shutdown())
        L.check(lib.f4j_tuner_clear())
        L.check(lib.f4j_init(0))                    # reads the cache
        shapes = [(a.shape[0], b.shape[1], a.shape[1]) for a, b in zip(As, Bs)]
        pa = np.concatenate([a.reshape(-1) for a in As])
        pb = np.concatenate([b.reshape(-1) for b in Bs])
        oa = np.cumsum([0] + [a.size for a in As])
        ob = np.cumsum([0] + [b.size for b in Bs])
        oc = np.cumsum([0] + [m * n for m, n, _ in shapes])
        with device.DeviceBuffer.from_array(pa) as dA, device.DeviceBuffer.from_array(pb) as dB, \
                device.DeviceBuffer(int(oc[-1]) * 8) as dC:
            plan = grouped.GroupedPlan(shapes, [dA.ptr + 8 * int(x) for x in oa[:-1]],
                                       [dB.ptr + 8 * int(x) for x in ob[:-1]],
                                       [dC.ptr + 8 * int(x) for x in oc[:-1]],
                                       [a.shape[1] for a in As], [b.shape[1] for b in Bs], [n for _, n, _ in shapes])
            info = plan.info()
            assert info["tiles"][0] > 0 and info["tile_cfg"][0] == 7
            cache.write_text("# emptied\n")
            plan.tune(reps=1)
            lines = [l for l in cache.read_text().splitlines() if not l.startswith("#")]
            assert any(l.startswith("gfx950 0,0 0 ") for l in lines), lines
            plan.run()
            device.sync()
            C = dC.download(np.float64, int(oc[-1]))
            plan.close()
        for i in (0, 39):
            assert check_real(C[oc[i]:oc[i + 1]].reshape(shapes[i][0], shapes[i][1]), As[i], Bs[i])
    finally:
        L.config_set("tunerCache", "")
        L.check(lib.f4j_tuner_clear())


@pytest.mark.parametrize("cfg", list(range(9)) + [100, 101, 102, 103])
@pytest.mark.parametrize("opA,opB", [("N", "N"), ("N", "T"), ("T", "N"), ("T", "T")])
def test_tile_configurations_bitwise(tmp_path, cfg, opA, opB):
    """Every grouped tile configuration -- the LDS kernels 0-8 (8- and 16-byte loads mixed
    within a workgroup) and the direct-load kernels 100-103 (one wave per tile, fragments
    from global memory) -- forced on every bucket through the tuner cache: every
    product, ragged edges, K not a multiple of 4 and K = 0 (no operand element to read:
    every load of its tiles lies outside the buffer records, or is not issued) included,
    equals the single-product FP64 call bit for bit
    (same MFMA k steps per element, in the same order); products with N = 1, which a
    single call runs in the reference's matrix-vector order, are held to the FP64 bound."""
    As, Bs = rand_problems(500 + cfg, 90, 1, 300, opA, opB)
    for i in (5, 41, 42, 89):                          # K = 0 items between the others
        M, N = op(As[i], opA).shape[0], op(Bs[i], opB).shape[1]
        As[i] = np.zeros((M, 0) if opA == "N" else (0, M))
        Bs[i] = np.zeros((0, N) if opB == "N" else (N, 0))
    kinds = f"{'NT'.index(opA)},{'NT'.index(opB)}"
    cache = tmp_path / "tiles.txt"
    cache.write_text("".join(f"gfx950 {kinds} {b} {cfg}\n" for b in range(4)))
    lib = L.load()
    L.check(lib.f4j_tuner_clear())
    try:
        L.check(lib.f4j_tuner_load(str(cache).encode()))
        shapes = [(op(a, opA).shape[0], op(b, opB).shape[1], op(a, opA).shape[1]) for a, b in zip(As, Bs)]
        pa = np.concatenate([a.reshape(-1) for a in As])
        pb = np.concatenate([b.reshape(-1) for b in Bs])
        oa = np.cumsum([0] + [a.size for a in As])
        ob = np.cumsum([0] + [b.size for b in Bs])
        oc = np.cumsum([0] + [m * n for m, n, _ in shapes])
        with device.DeviceBuffer.from_array(pa) as dA, device.DeviceBuffer.from_array(pb) as dB, \
                device.DeviceBuffer(int(oc[-1]) * 8) as dC:
            plan = grouped.GroupedPlan(shapes, [dA.ptr + 8 * int(x) for x in oa[:-1]],
                                       [dB.ptr + 8 * int(x) for x in ob[:-1]],
                                       [dC.ptr + 8 * int(x) for x in oc[:-1]],
                                       [max(a.shape[1], 1) for a in As], [max(b.shape[1], 1) for b in Bs],
                                       [n for _, n, _ in shapes], opA=opA, opB=opB)
            info = plan.info()
            assert all(c == cfg for c, t in zip(info["tile_cfg"], info["tiles"]) if t > 0), info
            plan.run()
            device.sync()
            C = dC.download(np.float64, int(oc[-1]))
            plan.close()
    finally:
        L.check(lib.f4j_tuner_clear())
    for i, (A, B) in enumerate(zip(As, Bs)):
        M, N, K = shapes[i]
        if N == 1:              # a single call with N = 1 runs the reference's matrix-vector order instead
            assert check_real(C[oc[i]:oc[i + 1]].reshape(M, N), op(A, opA), op(B, opB))
            continue
        single = np.full((M, N), np.nan)
        L.check(lib.f4j_dgemm(L.OP[opA], L.OP[opB], M, N, K, A.ctypes.data, max(A.shape[1], 1), B.ctypes.data,
                              max(B.shape[1], 1), single.ctypes.data, max(N, 1), L.PREC["fp64"], 0))
        assert np.array_equal(C[oc[i]:oc[i + 1]].reshape(M, N), single), (i, M, N, K)


@pytest.mark.parametrize("cfg", list(range(9)))
@pytest.mark.parametrize("opA,opB", [("C", "N"), ("N", "N"), ("T", "C"), ("N", "T")])
def test_complex_tile_configurations_bitwise(tmp_path, cfg, opA, opB):
    """Complex grouped products (real-expanded operands: the LDS kernels swap and negate in
    their fragment addressing) with every LDS tile configuration forced on every bucket: each
    product, K = 0 items included, equals the single complex FP64 call bit for bit; N = 1
    products (a matrix-vector order in the single call) are held to the FP64 bound."""
    As, Bs = rand_problems(700 + cfg, 40, 1, 160, opA, opB, cplx=True)
    for i in (3, 21):
        M, N = op(As[i], opA).shape[0], op(Bs[i], opB).shape[1]
        As[i] = np.zeros((M, 0) if opA == "N" else (0, M), dtype=np.complex128)
        Bs[i] = np.zeros((0, N) if opB == "N" else (N, 0), dtype=np.complex128)
    kz = {"N": 2, "T": 3, "C": 4}
    cache = tmp_path / "ctiles.txt"
    cache.write_text("".join(f"gfx950 {kz[opA]},{kz[opB]} {b} {cfg}\n" for b in range(4)))
    lib = L.load()
    L.check(lib.f4j_tuner_clear())
    try:
        L.check(lib.f4j_tuner_load(str(cache).encode()))
        Cs = grouped.mult_all(As, Bs, opA, opB)
    finally:
        L.check(lib.f4j_tuner_clear())
    for i, (A, B, C) in enumerate(zip(As, Bs, Cs)):
        oA, oB = op(A, opA), op(B, opB)
        M, K = oA.shape
        N = oB.shape[1]
        if N == 1:
            ref = o.cplx_gemm(A, B, opA, opB) if K else np.zeros((M, N), dtype=np.complex128)
            nA = np.abs(oA.real) + np.abs(oA.imag)
            nB = np.abs(oB.real) + np.abs(oB.imag)
            lim = 2 * 2 * max(K, 1) * U * (nA @ nB)
            assert np.all(np.abs(C.real - ref.real) <= lim) and np.all(np.abs(C.imag - ref.imag) <= lim)
            continue
        a = np.ascontiguousarray(A)
        b = np.ascontiguousarray(B)
        single = np.full((M, N), np.nan, dtype=np.complex128)
        L.check(lib.f4j_zgemm(L.OP[opA], L.OP[opB], M, N, K, a.ctypes.data, max(a.shape[1], 1), b.ctypes.data,
                              max(b.shape[1], 1), single.ctypes.data, max(N, 1), L.PREC["fp64"], 0, 0))
        assert np.array_equal(C, single), (i, M, N, K)
