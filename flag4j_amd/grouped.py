"""Grouped (batched) products through the C ABI: many independent products
C_i = op(A_i) op(B_i) of uneven shapes in persistent launches replayed from a
hipGraph (f4j_grouped_plan_* in include/flag4j_amd.h).

Flag4j has no batched API; a batch there is a Java loop of Matrix.mult or
A.H().mult(B) calls (M:arrays/dense/Matrix.java:638-640, M: =
src/main/java/org/flag4j/).  `mult_all` is that loop's drop-in replacement.
"""
import ctypes

import numpy as np

from . import _native
from .device import DeviceBuffer


class GroupedPlan:
    """A plan over device-resident operands (device pointers as ints)."""

    def __init__(self, shapes, a_ptrs, b_ptrs, c_ptrs, lda, ldb, ldc, opA="N", opB="N", complex_=False,
                 precision="fp64", slices=0, use3m=True):
        """precision "fp64", or "split_bf16" / "split_fp16" / "split_int8" for the
        batched split forms (`slices` digits or moduli, 0 = default; complex
        products as three real products when `use3m`, else four)."""
        n = len(shapes)
        M = np.array([s[0] for s in shapes], dtype=np.int64)
        N = np.array([s[1] for s in shapes], dtype=np.int64)
        K = np.array([s[2] for s in shapes], dtype=np.int64)
        self._keep = [M, N, K]
        arr = lambda v, t: np.ascontiguousarray(np.asarray(v, dtype=t))
        A, B, C = arr(a_ptrs, np.uint64), arr(b_ptrs, np.uint64), arr(c_ptrs, np.uint64)
        la, lb, lc = arr(lda, np.int64), arr(ldb, np.int64), arr(ldc, np.int64)
        self._keep += [A, B, C, la, lb, lc]
        p = lambda a: ctypes.c_void_p(a.ctypes.data) if a.size else None
        h = ctypes.c_void_p()
        self.split = precision in ("split_bf16", "split_fp16", "split_int8")
        if self.split:
            _native.check(_native.load().f4j_grouped_plan_create_split(
                int(complex_), _native.OP[opA], _native.OP[opB], n, p(M), p(N), p(K), p(A), p(la), p(B), p(lb),
                p(C), p(lc), _native.PREC[precision], int(slices), int(bool(use3m)), ctypes.byref(h)))
        else:
            _native.check(_native.load().f4j_grouped_plan_create(
                int(complex_), _native.OP[opA], _native.OP[opB], n, p(M), p(N), p(K), p(A), p(la), p(B), p(lb),
                p(C), p(lc), _native.PREC[precision], ctypes.byref(h)))
        self.handle = h
        self.count = n
        self.flops_real = float(np.sum(2.0 * M * N * K)) * (4 if complex_ else 1)

    def run(self, stream=0):
        _native.check(_native.load().f4j_grouped_plan_run(self.handle, stream or None))

    def tune(self, reps=3):
        _native.check(_native.load().f4j_grouped_plan_tune(self.handle, reps))

    def info(self):
        fl = ctypes.c_double()
        cfg = (ctypes.c_int * 4)()
        tiles = (ctypes.c_int64 * 4)()
        _native.check(_native.load().f4j_grouped_plan_info(self.handle, ctypes.byref(fl), cfg, tiles, 4))
        return {"flops": fl.value, "tile_cfg": list(cfg), "tiles": list(tiles)}

    def split_report(self):
        """After a run of a split plan: (per-problem max |C - AB| / (|A||B|) over
        the checked rows, -1 where unchecked; real slice products per run)."""
        err = np.zeros(self.count, dtype=np.float64)
        prods = ctypes.c_int()
        _native.check(_native.load().f4j_grouped_plan_split_report(
            self.handle, ctypes.c_void_p(err.ctypes.data) if self.count else None, self.count, ctypes.byref(prods)))
        return err, prods.value

    def close(self):
        if self.handle:
            _native.check(_native.load().f4j_grouped_plan_destroy(self.handle))
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def mult_all(As, Bs, opA="N", opB="N", precision="fp64", slices=0, use3m=True, report=False):
    """[op(A_i) op(B_i) for i] for numpy arrays (float64 or complex128), in one
    grouped plan: operands are packed into one device buffer each way.  With a
    split precision and report=True, also returns the plan's split report."""
    cplx = any(np.iscomplexobj(a) for a in As) or any(np.iscomplexobj(b) for b in Bs)
    dt = np.complex128 if cplx else np.float64
    As = [np.ascontiguousarray(a, dtype=dt) for a in As]
    Bs = [np.ascontiguousarray(b, dtype=dt) for b in Bs]
    shapes = []
    for a, b in zip(As, Bs):
        M, K = a.shape if opA == "N" else a.shape[::-1]
        K2, N = b.shape if opB == "N" else b.shape[::-1]
        if K != K2:
            from .arrays import LinearAlgebraException
            raise LinearAlgebraException(f"Cannot multiply matrices with shapes {a.shape} and {b.shape}.")
        shapes.append((M, N, K))
    es = np.dtype(dt).itemsize
    offs_a = np.cumsum([0] + [a.size for a in As])
    offs_b = np.cumsum([0] + [b.size for b in Bs])
    offs_c = np.cumsum([0] + [m * n for m, n, _ in shapes])
    pa = np.concatenate([a.reshape(-1) for a in As]) if As else np.zeros(0, dt)
    pb = np.concatenate([b.reshape(-1) for b in Bs]) if Bs else np.zeros(0, dt)
    with DeviceBuffer.from_array(pa if pa.size else np.zeros(1, dt)) as dA, \
            DeviceBuffer.from_array(pb if pb.size else np.zeros(1, dt)) as dB, \
            DeviceBuffer(max(int(offs_c[-1]) * es, es)) as dC:
        plan = GroupedPlan(shapes, [dA.ptr + int(o) * es for o in offs_a[:-1]],
                           [dB.ptr + int(o) * es for o in offs_b[:-1]],
                           [dC.ptr + int(o) * es for o in offs_c[:-1]],
                           [max(a.shape[1], 1) for a in As], [max(b.shape[1], 1) for b in Bs],
                           [max(n, 1) for _, n, _ in shapes], opA, opB, cplx, precision, slices, use3m)
        plan.run()
        _native.check(_native.load().f4j_stream_sync(None))
        out = dC.download(dt, int(offs_c[-1])) if offs_c[-1] else np.zeros(0, dt)
        rep = plan.split_report() if (report and plan.split) else None
        plan.close()
    res = [out[offs_c[i]:offs_c[i + 1]].reshape(shapes[i][0], shapes[i][1]) for i in range(len(shapes))]
    return (res, rep) if report else res
