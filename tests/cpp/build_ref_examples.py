"""Builds a few of the reference's example solver programs, unchanged, against
this library (include/ and cupss_amd/libcupss.so) for the GPU tests of
tests/test_gpu_ref_examples.py, which run them on an MI355X.

The sources are read from the reference tree (/root/reference/examples) and
compiled from a temporary copy outside this repository (its inc/ pointing at
include/; the unused <cuda.h>/<cuda_runtime.h> includes of example 04 are
dropped, as tests/test_examples_compile.py does); only the binaries land in
tests/cpp/bin/ref_examples/, which git ignores and the GPU runs receive with
the tree.  Without the reference tree this does nothing.

    python tests/cpp/build_ref_examples.py
"""
import os
import re
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = "/root/reference/examples"
OUT = os.path.join(HERE, "bin", "ref_examples")
LIBDIR = os.path.join(ROOT, "cupss_amd")
sys.path.insert(0, ROOT)
from cupss_amd.build import NO_PACKED_FP32  # noqa: E402

# (source under examples/, binary name)
PROGRAMS = [("01_diffusion/01_diffusion.cpp", "01_diffusion"),
            ("05_edwards_wilkinson/05_edwards_wilkinson.cpp", "05_edwards_wilkinson"),
            ("07_inhomogeneous_diffusion/07_inhomogeneous_diffusion.cu", "07_inhomogeneous_diffusion"),
            ("04_model_h/modelh_explicit_velocity.cpp", "modelh_explicit_velocity")]


def build(verbose=False):
    if not os.path.isdir(REF) or not os.path.exists(os.path.join(LIBDIR, "libcupss.so")):
        return []
    os.makedirs(OUT, exist_ok=True)
    built = []
    with tempfile.TemporaryDirectory() as tmp:
        os.symlink(os.path.join(ROOT, "include"), os.path.join(tmp, "inc"))
        os.makedirs(os.path.join(tmp, "examples"))
        os.symlink(os.path.join(ROOT, "include"), os.path.join(tmp, "examples", "inc"))
        for rel, name in PROGRAMS:
            src = open(os.path.join(REF, rel)).read()
            src = re.sub(r"#include\s*<cuda(_runtime)?\.h>\n", "", src)
            dst = os.path.join(tmp, "examples", rel)
            os.makedirs(os.path.dirname(dst), exist_ok=True)
            with open(dst, "w") as f:
                f.write(src)
            exe = os.path.join(OUT, name)
            cmd = ["/opt/rocm/bin/hipcc", "--offload-arch=gfx950", "-O2", "-std=c++17"] + NO_PACKED_FP32 + \
                ["-I" + os.path.join(ROOT, "include")]
            if rel.endswith(".cu"):
                cmd += ["-x", "hip"]
            cmd += [dst, "-o", exe, "-L" + LIBDIR, "-lcupss",
                    "-Wl,-rpath,$ORIGIN/" + os.path.relpath(LIBDIR, OUT)]
            r = subprocess.run(cmd, capture_output=True, text=True)
            if r.returncode != 0:
                raise RuntimeError("building %s failed:\n%s" % (rel, r.stderr[-3000:]))
            built.append(exe)
            if verbose:
                print("built", exe)
    return built


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
