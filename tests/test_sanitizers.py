"""Sanitizer runs of the host code, on the CPU (no GPU code is instrumented):

  * the whole CPU suite with the host library (cubiquity_amd/csrc/*.cpp) and the oracle built
    with AddressSanitizer + UndefinedBehaviorSanitizer (any report aborts the run);
  * tests/cpp/host_race_check.cpp under ThreadSanitizer and under ASan + UBSan: the parallel
    generator, the region-parallel edits with their lock-free interner and parallel merge,
    checkpoint/undo/redo and compaction, with 8 threads, and the store equal to the in-order one.

The sanitizer libraries are GCC's; the library's HIP objects are linked in uninstrumented.
"""
import concurrent.futures
import glob
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "cubiquity_amd", "csrc")
INC = os.path.join(ROOT, "include")
HIP_OBJS = sorted(glob.glob(os.path.join(ROOT, "cubiquity_amd", "_build", "*.hip.o")))
ASAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=undefined", "-fno-omit-frame-pointer", "-g", "-O1"]
TSAN = ["-fsanitize=thread", "-g", "-O1"]
COMMON = ["-pthread", "-ffp-contract=off", "-fno-fast-math", f"-I{INC}", f"-I{CSRC}"]


def _runtime(name):
    return subprocess.run(["gcc", f"-print-file-name={name}"], capture_output=True, text=True).stdout.strip()


def _run(cmd, **kw):
    r = subprocess.run(cmd, capture_output=True, text=True, **kw)
    assert r.returncode == 0, f"{' '.join(cmd)}\n{r.stdout[-3000:]}\n{r.stderr[-3000:]}"
    return r


@pytest.fixture(scope="module")
def asan_libs(tmp_path_factory):
    if not HIP_OBJS:
        pytest.skip("libcqx.so's HIP objects are not built (python -m cubiquity_amd.build)")
    d = str(tmp_path_factory.mktemp("asan"))
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.cpp")))
    objs = [os.path.join(d, os.path.basename(s) + ".o") for s in srcs]
    with concurrent.futures.ThreadPoolExecutor(8) as ex:
        list(ex.map(lambda so: _run(["g++", "-std=c++17", "-fPIC"] + ASAN + COMMON + ["-c", so[0], "-o", so[1]]),
                    zip(srcs, objs)))
    lib = os.path.join(d, "libcqx.so")
    _run(["g++", "-shared", "-fsanitize=address,undefined", "-o", lib] + objs + HIP_OBJS +
         ["-L/opt/rocm/lib", "-lamdhip64", "-lrocprofiler-sdk-roctx", "-Wl,-rpath,/opt/rocm/lib", "-pthread"])
    orc = os.path.join(d, "liboracle.so")
    _run(["gcc", "-std=c11", "-fPIC", "-shared"] + ASAN + COMMON + ["-o", orc] +
         sorted(glob.glob(os.path.join(ROOT, "oracle", "cpu_*.c"))) + ["-lm"])
    return lib, orc


@pytest.mark.slow
def test_cpu_suite_under_asan_and_ubsan(asan_libs):
    lib, orc = asan_libs
    env = dict(os.environ)
    env.update(CQX_LIBRARY=lib, CQX_ORACLE_LIBRARY=orc,
               LD_PRELOAD=f"{_runtime('libasan.so')} {_runtime('libubsan.so')}",
               ASAN_OPTIONS="detect_leaks=0:halt_on_error=1", UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    # the planted-slip builds and this file's own builds are not the instrumented libraries
    r = subprocess.run([sys.executable, "-m", "pytest", os.path.join(ROOT, "tests"), "-x", "-q", "-p", "no:cacheprovider",
                        "-m", "not gpu and not slow", "--ignore", os.path.join(ROOT, "tests", "test_sanitizers.py"),
                        "--ignore", os.path.join(ROOT, "tests", "test_oracle_mutations.py")],
                       capture_output=True, text=True, env=env, cwd=ROOT, timeout=1500)
    tail = r.stdout[-2500:] + r.stderr[-2500:]
    assert r.returncode == 0, tail
    assert " passed" in r.stdout and "ERROR: AddressSanitizer" not in tail and "runtime error" not in tail, tail


@pytest.mark.parametrize("flags", [TSAN, ASAN], ids=["tsan", "asan_ubsan"])
def test_host_threads_under_sanitizer(tmp_path, flags):
    exe = str(tmp_path / "host_race_check")
    _run(["g++", "-std=c++17"] + flags + COMMON + [os.path.join(ROOT, "tests", "cpp", "host_race_check.cpp")] +
         sorted(glob.glob(os.path.join(CSRC, "*.cpp"))) + ["-o", exe])
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=1", ASAN_OPTIONS="halt_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1")
    env.pop("LD_PRELOAD", None)
    r = _run([exe], env=env, timeout=600)
    assert "8 threads == 1 thread" in r.stdout, r.stdout + r.stderr
    assert "WARNING: ThreadSanitizer" not in r.stderr and "ERROR: AddressSanitizer" not in r.stderr
