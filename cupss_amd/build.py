"""Builds libcupss.so (HIP kernels, C-ABI and C++ API) in-tree for gfx950.

    python -m cupss_amd.build          # or: python cupss_amd/build.py

Every HIP source is compiled with hipcc --offload-arch=gfx950 and packed FP32
instructions disabled; the library lands next to this file.
"""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libcupss.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("CUPSS_ARCH", "gfx950")

# device code never gets packed FP32 instructions (v_pk_*_f32)
NO_PACKED_FP32 = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
HIP_FLAGS = ["--offload-arch=" + ARCH, "-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-function"] + \
    NO_PACKED_FP32 + ["-I" + os.path.join(ROOT, "include"), "-I" + CSRC]


def sources():
    out = []
    for d in (CSRC, os.path.join(CSRC, "host")):
        if os.path.isdir(d):
            for f in sorted(os.listdir(d)):
                if f.endswith((".hip", ".cpp")):
                    out.append(os.path.join(d, f))
    return out


def _compile(src):
    obj = os.path.join(BUILD, os.path.relpath(src, CSRC).replace(os.sep, "_") + ".o")
    cmd = [HIPCC] + HIP_FLAGS + ["-x", "hip", "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s%s" % (" ".join(cmd), r.stdout, r.stderr))
    return obj


# C++ programs linked against libcupss.so: the test drivers of tests/cpp and
# the example solvers of examples/ (binaries land in a bin/ folder beside them)
TOOLS = [os.path.join(ROOT, "tests", "cpp", "reference_cases.cpp"),
         os.path.join(ROOT, "tests", "cpp", "system_driver.cpp"),
         os.path.join(ROOT, "tests", "cpp", "hooks_driver.hip")]


def _tool_sources():
    out = list(TOOLS)
    ex = os.path.join(ROOT, "examples")
    if os.path.isdir(ex):
        for f in sorted(os.listdir(ex)):
            if f.endswith(".cpp"):
                out.append(os.path.join(ex, f))
    return out


def _build_tool(src):
    bindir = os.path.join(os.path.dirname(src), "bin")
    os.makedirs(bindir, exist_ok=True)
    exe = os.path.join(bindir, os.path.splitext(os.path.basename(src))[0])
    rpath = os.path.relpath(HERE, bindir)
    cmd = [HIPCC, "--offload-arch=" + ARCH, "-O2", "-std=c++17"] + NO_PACKED_FP32 + \
        ["-I" + os.path.join(ROOT, "include"), src,
           "-o", exe, "-L" + HERE, "-lcupss", "-Wl,-rpath,$ORIGIN/" + rpath]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("tool build failed: %s\n%s%s" % (" ".join(cmd), r.stdout, r.stderr))
    return exe


def build_tools():
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 4)) as ex:
        return list(ex.map(_build_tool, _tool_sources()))


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    srcs = sources()
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 4)) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-pthread", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s%s" % (" ".join(cmd), r.stdout, r.stderr))
    os.replace(tmp, LIB)
    tools = build_tools()
    if verbose:
        print("built", LIB)
        for t in tools:
            print("built", t)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
