"""Build recipes for the native parts of this package.

``build_library()`` compiles the HIP sources under ``csrc/`` for gfx950 into
``libdgpu.so`` next to this file.  It always recompiles: callers that only want
to load what is there use :func:`diablo3_test_amd.binding.load_library`.
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libdgpu.so")

HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # packed FP32 arithmetic is kept off: the compiler would otherwise form
    # v_pk_* FP32 instructions from scalar float code.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def hipcc():
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    cand = os.path.join(rocm, "bin", "hipcc")
    return cand if os.path.exists(cand) else "hipcc"


def build_library(out=LIB, verbose=False):
    """Compile csrc/dgpu_api.hip (which includes the kernel headers) to ``out``."""
    tmp = out + ".tmp"
    cmd = [hipcc()] + HIP_FLAGS + ["-o", tmp, os.path.join(CSRC, "dgpu_api.hip")]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + res.stdout + res.stderr)
    os.replace(tmp, out)
    return out
