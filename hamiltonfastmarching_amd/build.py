"""Build recipe of libhfm_amd.so (hipcc, gfx950), compiled in-tree.

    python -m hamiltonfastmarching_amd.build

Every HIP translation unit is compiled with
  * --offload-arch=gfx950 (MI355X only),
  * -ffp-contract=off: the Selling sign tests and the FP64 update must see the same FP64
    roundings as the reference, which is compiled without contraction,
  * -Xclang -target-feature -Xclang -packed-fp32-ops: no packed FP32 instructions.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "_lib")
LIB = os.path.join(OUT_DIR, "libhfm_amd.so")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")

SOURCES = ["hfm_kernels.hip", "hfm_plan.hip", "hfm_adjoint.hip", "hfm_server.hip", "hfm_geodesic.hip", "hfm_walls.hip",
           "hfm_exact.hip", "hfm_io.cpp"]
HEADERS = ["hfm_device.h", "hfm_kernels.h", "hfm_internal.h", "hfm_recompute.h", "selling.h"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wall", "-Wno-unused-function", "-Wno-unused-result", f"-I{INCLUDE}"]


def _compile(src: str) -> str:
    obj = os.path.join(OUT_DIR, os.path.basename(src) + ".o")
    lang = [] if src.endswith(".hip") else ["-x", "hip"]
    cmd = [HIPCC, *FLAGS, *lang, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    msg = "\n".join(l for l in (r.stdout + r.stderr).splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{msg}")
    if msg.strip():
        print(msg, file=sys.stderr)
    return obj


def build(verbose: bool = True) -> str:
    """Compile every source and link libhfm_amd.so.  Always recompiles."""
    os.makedirs(OUT_DIR, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with ThreadPoolExecutor(max_workers=min(len(srcs), 4)) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, LIB)
    for o in objs:
        os.remove(o)
    build_file_executables()
    if verbose:
        print(f"built {LIB}")
    return LIB


FILE_MODELS = ["Riemann2", "Riemann3", "ReedsSheppForward2", "Elastica2"]


def build_file_executables() -> str:
    """The FileHFM-compatible executable (host code over libhfm_amd.so): _lib/FileHFM_AMD, plus
    one copy per model named FileHFM_<Model> (the name fixes the model, as the reference's
    per-model executables do)."""
    exe = os.path.join(OUT_DIR, "FileHFM_AMD")
    tmp = exe + ".tmp"
    cmd = ["g++", "-std=c++17", "-O2", f"-I{INCLUDE}", os.path.join(CSRC, "hfm_file_main.cpp"), "-o", tmp,
           f"-L{OUT_DIR}", "-lhfm_amd", "-Wl,-rpath,$ORIGIN", "-Wl,-rpath,/opt/rocm/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("FileHFM executable failed to build:\n" + r.stdout + r.stderr)
    os.replace(tmp, exe)
    for m in FILE_MODELS:
        dst = os.path.join(OUT_DIR, f"FileHFM_{m}")
        shutil.copyfile(exe, dst + ".tmp")
        os.chmod(dst + ".tmp", 0o755)
        os.replace(dst + ".tmp", dst)
    return exe


if __name__ == "__main__":
    build()
