"""Seeded synthetic inputs with java.util.Random semantics, vectorised.

java.util.Random is the 48-bit LCG  s' = (0x5DEECE66D * s + 0xB) mod 2^48,
seeded with s0 = seed ^ 0x5DEECE66D, next(bits) = s >> (48 - bits), and
nextDouble() = ((next(26) << 27) + next(27)) * 2^-53 (Java SE API
specification of java.util.Random).  Because the step is affine, state n is
s_n = A_n s0 + C_n (mod 2^48) with A_n = a^n, C_n = c (a^n - 1)/(a - 1); this
module evaluates that in blocks with numpy uint64 arithmetic (wrapping mod
2^64 keeps the low 48 bits exact), so large inputs are generated without a
Python-level loop per element.  Flag4j draws its random matrices from
java.util.Random (M:rng/RandomComplex.java:69, M:rng/RandomDenseTensor.java:94-97),
so a Java harness seeded the same way produces the same matrices.
"""
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np

_A = 0x5DEECE66D
_C = 0xB
_MASK = (1 << 48) - 1
_BLOCK = 1 << 16
_STEP_TABLE = []


def _affine_pow(n):
    """(A_n, C_n) such that s_n = A_n * s_0 + C_n mod 2^48."""
    a, c = 1, 0
    pa, pc = _A, _C
    while n:
        if n & 1:
            a, c = (pa * a) & _MASK, (pa * c + pc) & _MASK
        pa, pc = (pa * pa) & _MASK, (pa * pc + pc) & _MASK
        n >>= 1
    return a, c


def _step_table():
    """(A_k, C_k) for k = 1.._BLOCK: the states of one block relative to its origin."""
    if not _STEP_TABLE:
        a1 = np.empty(_BLOCK, dtype=np.uint64)
        c1 = np.empty(_BLOCK, dtype=np.uint64)
        a, c = 1, 0
        for k in range(_BLOCK):
            a, c = (_A * a) & _MASK, (_A * c + _C) & _MASK
            a1[k], c1[k] = a, c
        _STEP_TABLE.extend([a1, c1])
    return _STEP_TABLE[0], _STEP_TABLE[1]


def _states(s0, start, count):
    """LCG states s_{start+1} .. s_{start+count} (the values next() shifts from)."""
    L = _BLOCK
    nblk = (count + L - 1) // L
    a1, c1 = _step_table()
    out = np.empty(nblk * L, dtype=np.uint64)
    mask = np.uint64(_MASK)
    ab, cb = _affine_pow(start)
    aL, cL = _affine_pow(L)
    origin = (ab * s0 + cb) & _MASK              # s_start
    for b in range(nblk):
        o = np.uint64(origin)
        with np.errstate(over="ignore"):
            out[b * L:(b + 1) * L] = (a1 * o + c1) & mask
        origin = (aL * origin + cL) & _MASK
    return out[:count]


def _next_doubles(seed, start_double, n):
    s0 = (seed ^ _A) & _MASK
    st = _states(s0, 2 * start_double, 2 * n)
    hi = (st[0::2] >> np.uint64(48 - 26)).astype(np.int64)
    lo = (st[1::2] >> np.uint64(48 - 27)).astype(np.int64)
    return ((hi << 27) + lo).astype(np.float64) * 2.0**-53


def _threads():
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def next_doubles(seed, n, chunk=1 << 22, start=0):
    """Values start .. start + n - 1 of new java.util.Random(seed).nextDouble().
    Chunks are independent (each jumps ahead to its own origin), so they are
    generated on a thread pool: numpy releases the interpreter lock inside its
    array loops.  The stream is the same whatever the thread count."""
    out = np.empty(n, dtype=np.float64)
    spans = [(s, min(n, s + chunk)) for s in range(0, n, chunk)]

    def fill(span):
        s, e = span
        out[s:e] = _next_doubles(seed, start + s, e - s)

    if len(spans) <= 1:
        for sp in spans:
            fill(sp)
    else:
        with ThreadPoolExecutor(max_workers=min(_threads(), len(spans))) as ex:
            list(ex.map(fill, spans))
    return out


def uniform(seed, rows, cols):
    """rows x cols matrix of 2r - 1 (uniform on [-1, 1)), row-major."""
    return (2.0 * next_doubles(seed, rows * cols) - 1.0).reshape(rows, cols)


def uniform_complex(seed, rows, cols):
    """Complex matrix whose real and imaginary parts are 2r - 1, drawn re then im per entry."""
    v = 2.0 * next_doubles(seed, 2 * rows * cols) - 1.0
    return v.view(np.complex128).reshape(rows, cols)


def scaled_uniform(seed, rows, cols, phi):
    """(2r - 1) * 10^(phi (2r' - 1)): entries spread over 2 phi decades (r, r' consecutive draws)."""
    v = next_doubles(seed, 2 * rows * cols)
    r, rp = v[0::2], v[1::2]
    return ((2.0 * r - 1.0) * 10.0 ** (phi * (2.0 * rp - 1.0))).reshape(rows, cols)


def _polar(u):
    """Accepted polar-method pairs of consecutive doubles u[0::2], u[1::2], in order."""
    v1, v2 = 2 * u[0::2] - 1, 2 * u[1::2] - 1
    s = v1 * v1 + v2 * v2
    ok = (s < 1) & (s != 0)
    mul = np.sqrt(-2 * np.log(s[ok]) / s[ok])
    g = np.empty(2 * int(ok.sum()))
    g[0::2], g[1::2] = v1[ok] * mul, v2[ok] * mul
    return g


def scaled_uniform_complex(seed, n, phi):
    """n complex entries: real part (2 r0 - 1) 10^(phi (2 r1 - 1)), imaginary part
    (2 r2 - 1) 10^(phi (2 r3 - 1)), from four consecutive draws r0 .. r3 per entry."""
    v = next_doubles(seed, 4 * n)
    out = np.empty(n, dtype=np.complex128)
    f = out.view(np.float64)
    f[0::2] = (2.0 * v[0::4] - 1.0) * 10.0 ** (phi * (2.0 * v[1::4] - 1.0))
    f[1::2] = (2.0 * v[2::4] - 1.0) * 10.0 ** (phi * (2.0 * v[3::4] - 1.0))
    return out


def next_gaussians(seed, n, chunk=1 << 22):
    """The first n values of new java.util.Random(seed).nextGaussian(): the
    polar method on consecutive nextDouble() pairs, rejected pairs skipped,
    each accepted pair giving v1*mul then v2*mul (log/sqrt from numpy).
    Chunks of pairs are transformed on a thread pool and joined in order."""
    parts, start, need = [], 0, n
    while need > 0:
        m = max(64, int(need * 0.7) + 64)            # pairs to draw (acceptance ~ pi/4)
        u = next_doubles(seed, 2 * m, start=start)
        start += 2 * m
        spans = [u[s:min(2 * m, s + chunk)] for s in range(0, 2 * m, chunk)]
        if len(spans) <= 1:
            got = [_polar(sp) for sp in spans]
        else:
            with ThreadPoolExecutor(max_workers=min(_threads(), len(spans))) as ex:
                got = list(ex.map(_polar, spans))
        parts.extend(got)
        need -= sum(g.size for g in got)
    out = np.concatenate(parts) if parts else np.empty(0)
    return out[:n]


def log_uniform_ints(seed, n, a, b):
    """n integers floor(exp(ln a + r (ln(b+1) - ln a))) in [a, b], r = nextDouble()."""
    r = next_doubles(seed, n)
    v = np.floor(np.exp(np.log(a) + r * (np.log(b + 1) - np.log(a)))).astype(np.int64)
    return np.clip(v, a, b)


def gaussian(seed, rows, cols):
    return next_gaussians(seed, rows * cols).reshape(rows, cols)
