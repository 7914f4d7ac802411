"""Build recipe of the in-tree HIP library ``matfree_amd/libslq.so`` (gfx950).

``python -m matfree_amd._build`` compiles every ``csrc/*.hip`` translation
unit with ``hipcc`` into one shared library next to this file.  The library is
loaded with ctypes by ``matfree_amd._lib``; nothing is JIT-compiled at import.
"""

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB_NAME = "libslq.so"
LIB_PATH = os.path.join(HERE, LIB_NAME)
ARCH = "gfx950"

# Packed FP32 VALU instructions are disabled for all device code: only scalar
# FP32 arithmetic is emitted for the vector units.
HIPCC_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-Wall",
    "-Wno-unused-function",
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
    # MFMA results and accumulators in ordinary VGPRs (no AGPR copies)
    "-mllvm",
    "-amdgpu-mfma-vgpr-form=1",
]


def sources():
    return sorted(
        os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip")
    )


def hipcc():
    path = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(path):
        raise RuntimeError("hipcc not found: the ROCm toolchain is required to build libslq")
    return path


def build(verbose=False):
    """Compile libslq.so for gfx950 (always recompiles)."""
    tmp = LIB_PATH + ".tmp"
    cmd = [hipcc(), *HIPCC_FLAGS, "-I", os.path.join(HERE, "..", "include"), *sources(), "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    proc = subprocess.run(cmd, capture_output=True, text=True)
    # The host pass does not know the packed-fp32 feature and says so; that
    # note is expected and not shown.
    noise = "is not a recognized feature for this target"
    msgs = [ln for ln in (proc.stdout + proc.stderr).splitlines() if noise not in ln]
    if proc.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + "\n".join(msgs))
    if verbose and msgs:
        print("\n".join(msgs), flush=True)
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
