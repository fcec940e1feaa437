"""Host emulation of the FP64 MFMA kernel's index math (tests/emu/gemm_emu.cpp):
tile mapping, LDS placement and per-lane fragment addressing for every operand
kind, with bounds checks, on exact integer inputs.  Runs without a GPU."""
import os
import subprocess

import pytest

from conftest import ROOT

EMU_SRC = os.path.join(ROOT, "tests", "emu", "gemm_emu.cpp")
EMU_BIN = os.path.join(ROOT, "_build", "gemm_emu")


def build_once(src, binary, deps):
    """Compile when a dependency is newer; concurrent test processes each compile
    into their own file and rename it into place, so none runs a half-written binary."""
    os.makedirs(os.path.dirname(binary), exist_ok=True)
    if os.path.exists(binary) and all(os.path.getmtime(p) <= os.path.getmtime(binary) for p in deps + [src]):
        return binary
    tmp = f"{binary}.{os.getpid()}"
    subprocess.run(["g++", "-O2", "-std=c++17", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", "-o", tmp, src],
                   check=True)
    os.replace(tmp, binary)
    return binary


@pytest.fixture(scope="module")
def emu():
    hdrs = [os.path.join(ROOT, "flag4j_amd", "csrc", h) for h in ("gemm_tile.h", "f4j_internal.h")]
    return build_once(EMU_SRC, EMU_BIN, hdrs)


KINDS = {"R_N": 0, "R_T": 1, "Z_N": 2, "Z_T": 3, "Z_C": 4}
PAIRS = [(a, b) for a in ("R_N", "R_T") for b in ("R_N", "R_T")] + \
        [(a, b) for a in ("Z_N", "Z_T", "Z_C") for b in ("Z_N", "Z_T", "Z_C")]


@pytest.mark.parametrize("ka,kb", PAIRS)
@pytest.mark.parametrize("shape,cfg", [((37, 45, 29), 1), ((130, 70, 35), 0), ((20, 129, 17), 2), ((1, 1, 1), 1),
                                       ((64, 64, 32), 1), ((130, 140, 36), 0), ((150, 130, 40), 4),
                                       ((130, 129, 70), 5), ((260, 140, 20), 6), ((70, 140, 18), 7),
                                       ((140, 70, 33), 8)])
@pytest.mark.parametrize("vec", [1, 2])
def test_kernel_index_math(emu, ka, kb, shape, cfg, vec):
    M, N, K = shape
    r = subprocess.run([emu, str(KINDS[ka]), str(KINDS[kb]), str(M), str(N), str(K), str(cfg), str(vec)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.strip().startswith("OK"), r.stdout


CRT_SRC = os.path.join(ROOT, "tests", "emu", "crt_emu.cpp")
CRT_BIN = os.path.join(ROOT, "_build", "crt_emu")


@pytest.fixture(scope="module")
def crt_emu():
    """Host emulation of the modular GEMM (fragment-major residue layout, LDS
    stages, epilogue) with the kernels' own offset functions."""
    deps = [os.path.join(ROOT, "flag4j_amd", "csrc", h) for h in ("gemm_tile.h", "crt_layout.h")]
    return build_once(CRT_SRC, CRT_BIN, deps)


@pytest.mark.parametrize("shape", [(1, 1, 1), (7, 300, 13), (130, 70, 257), (300, 200, 500), (257, 129, 384),
                                   (512, 1024, 128), (33, 260, 640)])
def test_crt_kernel_index_math(crt_emu, shape):
    r = subprocess.run([crt_emu, *map(str, shape)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and r.stdout.strip().startswith("OK"), r.stdout


SLICE_SRC = os.path.join(ROOT, "tests", "emu", "slice_emu.cpp")
SLICE_BIN = os.path.join(ROOT, "_build", "slice_emu")


@pytest.fixture(scope="module")
def slice_emu():
    """Host emulation of the fragment-major digit-slice product (digit split writes,
    LDS ring, fragment reads, FP32 flushes into FP64, epilogue) with the kernel's own
    offset functions."""
    deps = [os.path.join(ROOT, "flag4j_amd", "csrc", h) for h in ("gemm_tile.h", "slice_frag.h")]
    return build_once(SLICE_SRC, SLICE_BIN, deps)


@pytest.mark.parametrize("shape", [(1, 1, 1, 1, 4), (37, 45, 29, 2, 4), (300, 200, 500, 3, 4), (257, 129, 384, 4, 1),
                                   (260, 140, 300, 3, 1), (513, 257, 256, 2, 4)])
def test_digit_slice_kernel_index_math(slice_emu, shape):
    """shape = (M, N, K, digits, flush steps): BF16 flushes every 4 steps, FP16 every step."""
    r = subprocess.run([slice_emu, *map(str, shape)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and r.stdout.strip().startswith("OK"), r.stdout


RES_SRC = os.path.join(ROOT, "tests", "emu", "residue_emu.cpp")
RES_BIN = os.path.join(ROOT, "_build", "residue_emu")


def test_crt_residue_arithmetic():
    """The modular GEMM's residue arithmetic (crt_residue.h: operand residues for
    t = 1 .. 75, product residues of int32 accumulators) against exact integer
    residues, including the exact-quotient and extreme cases."""
    binary = build_once(RES_SRC, RES_BIN, [os.path.join(ROOT, "flag4j_amd", "csrc", "crt_residue.h")])
    r = subprocess.run([binary], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and r.stdout.strip().startswith("OK"), r.stdout


SCHED_SRC = os.path.join(ROOT, "tests", "emu", "sched_emu.cpp")
SCHED_BIN = os.path.join(ROOT, "_build", "sched_emu")


@pytest.fixture(scope="module")
def sched_emu():
    """The XCD-aware tile schedule of the persistent launches, on the host."""
    return build_once(SCHED_SRC, SCHED_BIN, [os.path.join(ROOT, "flag4j_amd", "csrc", "xcd_schedule.h")])


@pytest.mark.parametrize("nprod,workers,seed", [(1, 1, 1), (5, 3, 2), (200, 512, 3), (3000, 768, 4), (40, 8, 5),
                                                (1000, 100, 6)])
def test_xcd_schedule(sched_emu, nprod, workers, seed):
    r = subprocess.run([sched_emu, str(nprod), str(workers), str(seed)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.strip().startswith("OK"), r.stdout


GEMV_SRC = os.path.join(ROOT, "tests", "emu", "gemv_emu.cpp")
GEMV_BIN = os.path.join(ROOT, "_build", "gemv_emu")


@pytest.fixture(scope="module")
def gemv_emu():
    """Host emulation of the matrix-vector kernels (direct-to-LDS ring with its vmcnt
    waits, register-staged form, one-chain-per-row split kernel) with the kernels'
    own index functions."""
    return build_once(GEMV_SRC, GEMV_BIN, [os.path.join(ROOT, "flag4j_amd", "csrc", "gemv_layout.h")])


@pytest.mark.parametrize("mode,M,K,lda,incx", [
    ("dma2", 8, 512, 512, 1), ("dma2", 13, 1000, 1002, 1), ("dma2", 64, 64, 64, 1), ("dma2", 5, 2, 2, 1),
    ("dma2", 100, 4094, 4096, 1), ("dma2", 9, 514, 520, 1),
    ("dma1", 21, 600, 600, 3), ("dma1", 8, 130, 132, 1), ("dma1", 3, 2048, 2050, 2),
    ("reg", 7, 999, 1001, 2), ("reg", 17, 1, 1, 1), ("reg", 33, 513, 515, 1), ("reg", 8, 64, 64, 5),
    ("reg", 12, 1536, 1537, 1)])
def test_gemv_blocked_kernels(gemv_emu, mode, M, K, lda, incx):
    r = subprocess.run([gemv_emu, mode, str(M), str(K), str(lda), str(incx)], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0 and r.stdout.strip().endswith("ok"), r.stdout + r.stderr


@pytest.mark.parametrize("kind", ["chain", "cplx", "rx", "cx"])
@pytest.mark.parametrize("M,K", [(37, 200), (16, 64), (5, 1), (40, 129)])
def test_gemv_split_kernel(gemv_emu, kind, M, K):
    r = subprocess.run([gemv_emu, "split", kind, str(M), str(K)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and r.stdout.strip().endswith("ok"), r.stdout + r.stderr
