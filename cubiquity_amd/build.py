"""Builds libcqx.so in-tree: host C++ with g++, HIP sources with hipcc for gfx950.

Run as ``python -m cubiquity_amd.build``. Objects go to ``cubiquity_amd/_build/``; the shared
library to ``cubiquity_amd/libcqx.so``. Every call recompiles (the build is a few seconds).
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libcqx.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")

COMMON = ["-O3", "-fPIC", "-ffp-contract=off", "-fno-fast-math", f"-I{INCLUDE}", f"-I{CSRC}"]
HOST_FLAGS = COMMON + ["-std=c++17", "-pthread", "-Wall", "-Wextra"]
# Exact IEEE FP32 for bit parity with the reference: no contraction, correctly rounded
# division/sqrt, denormals kept. Packed FP32 instructions are disabled for all HIP code.
HIP_FLAGS = COMMON + [
    "-std=c++17",
    "--offload-arch=gfx950",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines()
                    if "packed-fp32-ops" not in l and "not a recognized feature" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"build step failed ({r.returncode}): {' '.join(cmd)}\n{out}")
    if verbose and out.strip():
        print(out, flush=True)


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    jobs = []
    for name in sorted(os.listdir(CSRC)):
        src = os.path.join(CSRC, name)
        base, ext = os.path.splitext(name)
        if ext == ".cpp":
            obj = os.path.join(BUILD, base + ".o")
            jobs.append([CXX] + HOST_FLAGS + ["-c", src, "-o", obj])
        elif ext == ".hip":
            obj = os.path.join(BUILD, base + ".hip.o")
            jobs.append([HIPCC] + HIP_FLAGS + ["-c", src, "-o", obj])
        else:
            continue
        objs.append(obj)
    procs = []
    for cmd in jobs:
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
    errors = []
    for cmd, p in procs:
        out, _ = p.communicate()
        out = "\n".join(l for l in out.splitlines()
                        if "packed-fp32-ops" not in l and "not a recognized feature" not in l)
        if p.returncode != 0:
            errors.append(f"{' '.join(cmd)}\n{out}")
        elif verbose and out.strip():
            print(out, flush=True)
    if errors:
        raise RuntimeError("compile failed:\n" + "\n\n".join(errors))
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "--offload-arch=gfx950", "-o", tmp] + objs +
         ["-pthread", "-L/opt/rocm/lib", "-lrocprofiler-sdk-roctx", "-Wl,-rpath,/opt/rocm/lib"], verbose)
    os.replace(tmp, LIB)
    build_cpp_checks(verbose)
    return LIB


def build_cpp_checks(verbose=False):
    """Test programs: the C++ drop-in example run by the GPU tests (links libcqx.so through its
    rpath) and the host compilation of the device traversal core used by the CPU tests."""
    cpp = os.path.join(ROOT, "tests", "cpp")
    src = os.path.join(cpp, "dropin_check.cpp")
    if not os.path.exists(src):
        return None
    out = os.path.join(cpp, "dropin_check")
    _run([CXX, "-std=c++17", "-O2", f"-I{INCLUDE}", src, "-o", out, f"-L{HERE}", "-lcqx",
          "-Wl,-rpath,$ORIGIN/../../cubiquity_amd"], verbose)
    emu = os.path.join(cpp, "core_emulation.cpp")
    if os.path.exists(emu):
        _run([CXX, "-std=c++17", "-O2", "-ffp-contract=off", "-fno-fast-math", f"-I{os.path.join(cpp, 'host_shim')}",
              f"-I{INCLUDE}", f"-I{CSRC}", emu, "-o", os.path.join(cpp, "core_emulation")], verbose)
    return out


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
