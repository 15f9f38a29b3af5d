"""Build libcbk_amd.so (the HIP library behind include/cbk.h) in-tree for gfx950.

    python -m cbk_amd.build        # or cbk_amd.build.build()

Every source under csrc/ is compiled by hipcc for gfx950 only, with FMA
contraction off (the kernels spell out each fused multiply-add, which keeps
their bits equal to the reference's) and without packed FP32 instructions,
then linked into cbk_amd/libcbk_amd.so.
"""

import concurrent.futures
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libcbk_amd.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

CFLAGS = [
    "-std=c++17", "-O3", "-fPIC", f"--offload-arch={ARCH}",
    "-ffp-contract=off", "-fno-math-errno",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"), "-I", CSRC,
    "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument",
]


def _up_to_date(src, obj):
    """obj is newer than src and every header its dependency file lists"""
    dep = obj[:-2] + ".d"
    if not (os.path.exists(obj) and os.path.exists(dep)):
        return False
    text = open(dep).read().replace("\\\n", " ")
    files = text.split(":", 1)[1].split() if ":" in text else [src]
    t = os.path.getmtime(obj)
    return all(os.path.exists(f) and os.path.getmtime(f) < t for f in files + [src])


def _compile(src, incremental=False):
    obj = os.path.join(OBJ, os.path.basename(src) + ".o")
    if incremental and _up_to_date(src, obj):
        return obj
    lang = ["-x", "hip"] if src.endswith((".hip", ".cpp")) else []
    cmd = [HIPCC] + CFLAGS + lang + ["-MMD", "-MF", obj[:-2] + ".d", "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose=True, incremental=False):
    """Compile every source (incremental: only those older than their object
    or one of its headers) and link the library."""
    os.makedirs(OBJ, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")) + glob.glob(os.path.join(CSRC, "*.cpp")))
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(lambda s: _compile(s, incremental), srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}",
           "-Wl,--version-script=" + os.path.join(CSRC, "exports.map"), "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB} from {len(srcs)} sources")
    return LIB


if __name__ == "__main__":
    build(incremental="--incremental" in sys.argv)
    sys.exit(0)
