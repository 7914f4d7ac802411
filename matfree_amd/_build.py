"""Build recipe of the in-tree HIP library ``matfree_amd/libslq.so`` (gfx950).

``python -m matfree_amd._build`` compiles every ``csrc/*.hip`` translation
unit with ``hipcc`` into one shared library next to this file.  The library is
loaded with ctypes by ``matfree_amd._lib``; nothing is JIT-compiled at import.
"""

import concurrent.futures
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


def _run(cmd):
    proc = subprocess.run(cmd, capture_output=True, text=True)
    # The host pass does not know the packed-fp32 feature and says so; that
    # note is expected and not shown.
    noise = "is not a recognized feature for this target"
    msgs = [ln for ln in (proc.stdout + proc.stderr).splitlines() if noise not in ln]
    return proc.returncode, msgs


def build(verbose=False):
    """Compile libslq.so for gfx950 (always recompiles).

    Every translation unit is compiled to an object in ``_build/`` (in
    parallel, one hipcc per file), then the objects are linked into the
    shared library next to this file.
    """
    objdir = os.path.join(HERE, "..", "_build")
    os.makedirs(objdir, exist_ok=True)
    inc = ["-I", os.path.join(HERE, "..", "include")]
    compile_flags = [f for f in HIPCC_FLAGS if f != "-shared"]
    jobs = []
    for src in sources():
        obj = os.path.join(objdir, os.path.basename(src)[:-4] + ".o")
        jobs.append((src, obj, [hipcc(), *compile_flags, *inc, "-c", src, "-o", obj]))
    workers = max(1, min(len(jobs), os.cpu_count() or 1, 8))
    with concurrent.futures.ThreadPoolExecutor(workers) as pool:
        results = list(pool.map(lambda job: (job, _run(job[2])), jobs))
    for (src, _, cmd), (rc, msgs) in results:
        if verbose:
            print(" ".join(cmd), flush=True)
        if rc != 0:
            raise RuntimeError(f"hipcc failed on {os.path.basename(src)}:\n" + "\n".join(msgs))
        if verbose and msgs:
            print("\n".join(msgs), flush=True)
    tmp = LIB_PATH + ".tmp"
    link = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", *[obj for _, obj, _ in jobs],
            "-o", tmp]
    if verbose:
        print(" ".join(link), flush=True)
    rc, msgs = _run(link)
    if rc != 0:
        raise RuntimeError("hipcc link failed:\n" + "\n".join(msgs))
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
