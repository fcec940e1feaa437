"""CPU builds of the host emulators and of the oracle under sanitizers.

The emulators run the kernels' own index math on the host (tests/emu/*.cpp); the
oracle's thread pool restates ThreadManager.  Here they are rebuilt with
AddressSanitizer + UndefinedBehaviorSanitizer (out-of-bounds accesses, use after
free, overflow of signed integers, misaligned or invalid shifts) and the oracle's
pool kernels with ThreadSanitizer (data races), and run on small cases; any
report fails the run (-fno-sanitize-recover, halt_on_error).  No GPU is used.
"""
import os
import shutil
import subprocess

import pytest

from conftest import ROOT

EMU = os.path.join(ROOT, "tests", "emu")
OUT = os.path.join(ROOT, "_build", "sanitize")
# -O0: the instrumented template code of the emulators compiles in seconds instead of a minute
ASAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-O0"]
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:halt_on_error=1:abort_on_error=0",
           UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1", TSAN_OPTIONS="halt_on_error=1")

pytestmark = pytest.mark.skipif(shutil.which("g++") is None or shutil.which("gcc") is None,
                                reason="needs gcc / g++")


def build(name, cmd):
    """Compile unless the binary is newer than every source under tests/emu, the
    oracle and the kernel headers it includes."""
    os.makedirs(OUT, exist_ok=True)
    binary = os.path.join(OUT, name)
    deps = [os.path.join(d, f) for d in (EMU, os.path.join(ROOT, "oracle"), os.path.join(ROOT, "flag4j_amd", "csrc"))
            for f in os.listdir(d) if f.endswith((".c", ".cpp", ".h"))]
    if os.path.exists(binary) and all(os.path.getmtime(p) <= os.path.getmtime(binary) for p in deps):
        return binary
    tmp = f"{binary}.{os.getpid()}"
    r = subprocess.run(cmd + ["-o", tmp], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    os.replace(tmp, binary)
    return binary


def run(binary, *args, timeout=300):
    r = subprocess.run([binary, *map(str, args)], capture_output=True, text=True, timeout=timeout, env=ENV)
    assert r.returncode == 0 and r.stdout.strip().splitlines()[-1].startswith("OK"), (r.stdout[-2000:],
                                                                                     r.stderr[-4000:])


EMU_CASES = {
    "gemm_emu": [(0, 0, 37, 45, 29, 1, 1), (2, 4, 130, 70, 35, 0, 2), (1, 1, 20, 129, 17, 2, 1), (3, 2, 150, 130, 40, 4, 2)],
    "crt_emu": [(7, 300, 13), (130, 70, 257), (257, 129, 384)],
    "sched_emu": [(5, 3, 2), (200, 512, 3), (1000, 100, 6)],
    "residue_emu": [()],
    "slice_emu": [(37, 45, 29, 2, 4), (257, 129, 300, 3, 1)],
}


@pytest.mark.parametrize("name", sorted(EMU_CASES))
def test_emulator_under_asan_ubsan(name):
    binary = build(name + "_asan", ["g++", "-std=c++17", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", *ASAN,
                                    os.path.join(EMU, name + ".cpp")])
    for args in EMU_CASES[name]:
        run(binary, *args)


def test_oracle_under_asan_ubsan():
    binary = build("oracle_asan", ["gcc", "-std=gnu11", "-ffp-contract=off", "-fno-fast-math", *ASAN,
                                   os.path.join(EMU, "oracle_sanitize.c"), os.path.join(ROOT, "oracle", "f4j_oracle.c"),
                                   "-lm", "-lpthread"])
    run(binary, 4)


def test_oracle_thread_pool_under_tsan():
    binary = build("oracle_tsan", ["gcc", "-std=gnu11", "-ffp-contract=off", "-fno-fast-math", "-fsanitize=thread",
                                   "-g", "-O1", os.path.join(EMU, "oracle_sanitize.c"),
                                   os.path.join(ROOT, "oracle", "f4j_oracle.c"), "-lm", "-lpthread"])
    run(binary, 6)
