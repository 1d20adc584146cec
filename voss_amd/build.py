"""Build libvoss.so (host C++ engine + gfx950 HIP kernels) in-tree.

    python voss_amd/build.py [--sanitize]

(run as a script: importing the package itself needs the built library)

Every translation unit is compiled by hipcc for gfx950 into _build/, then
linked into voss_amd/libvoss.so next to this file (where the Python
package's loader looks first).
"""

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
OUT = PKG / "libvoss.so"
BUILD = ROOT / "_build"
SOURCES = ["kernels.hip", "misc_kernels.hip", "engine.hip", "engine_callers.hip",
           "serve.cpp", "capi.cpp", "host_math.cpp"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
         # packed FP32 ops off for every device pass (they have produced wrong
         # results on these boxes; the host pass warns that it does not know it)
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wall", "-Wno-unused-function", "-Wno-unknown-warning-option",
         f"-I{ROOT / 'include'}", f"-I{CSRC}"]


# host code only (never the device code): AddressSanitizer and UBSan
SANITIZE = ["-Xarch_host", "-fsanitize=address,undefined", "-Xarch_host", "-fno-omit-frame-pointer"]


def _compile(src: str, bdir: Path = BUILD, extra=()) -> Path:
    obj = bdir / (src + ".o")
    cmd = [HIPCC, *FLAGS, *extra, "-c", str(CSRC / src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr[-6000:]}")
    return obj


def build(verbose: bool = False, sanitize: bool = False) -> Path:
    """Build voss_amd/libvoss.so; sanitize=True instead builds
    _build/san/libvoss.so with the host code under AddressSanitizer and
    UBSan (load it with VOSS_LIBRARY_PATH and libasan preloaded)."""
    bdir = BUILD / "san" if sanitize else BUILD
    out = bdir / "libvoss.so" if sanitize else OUT
    extra = SANITIZE if sanitize else []
    bdir.mkdir(parents=True, exist_ok=True)
    srcs = [s for s in SOURCES if (CSRC / s).exists()]
    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(lambda src: _compile(src, bdir, extra), srcs))
    tmp = out.with_suffix(".so.tmp")
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *extra, *map(str, objs), "-o", str(tmp)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-6000:]}")
    os.replace(tmp, out)
    if verbose:
        print(f"built {out}")
    return out


if __name__ == "__main__":
    build(verbose=True, sanitize="--sanitize" in sys.argv[1:])
    sys.exit(0)
