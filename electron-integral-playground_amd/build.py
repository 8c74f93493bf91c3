"""Build libmolecular_integral.so for gfx950: run the kernel generator, compile each source with hipcc
in parallel, link.  ``build()`` always recompiles; ``ensure_built()`` only builds when the library is
missing.
"""

import concurrent.futures as cf
import os
import pathlib
import shutil
import subprocess
import sys

PKG = pathlib.Path(__file__).resolve().parent
CSRC = PKG / "csrc"
GEN = CSRC / "generated"
LIBDIR = PKG / "lib"
LIB = LIBDIR / "libmolecular_integral.so"
OBJ = PKG / "lib" / "obj"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# Packed FP32 instructions are switched off for all device code: hipcc would otherwise form them from
# scalar float code (float32 kernels of the mixed-precision path).
FLAGS = ["--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-Wno-unused-variable",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-I", str(CSRC)]


def _sources():
    srcs = sorted(GEN.glob("*.hip")) + [GEN / "registry.cpp", CSRC / "generic_kernels.hip", CSRC / "library.hip"]
    # largest first, so the long compiles start early
    return sorted(srcs, key=lambda p: -p.stat().st_size)


def _compile(src):
    obj = OBJ / (src.stem + ".o")
    cmd = [HIPCC] + FLAGS + ["-c", str(src), "-o", str(obj)]
    if src.suffix == ".cpp":
        cmd = [HIPCC, "-x", "hip"] + FLAGS + ["-c", str(src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src.name}:\n{r.stderr[-4000:]}")
    return obj


def build(jobs=None, verbose=False):
    sys.path.insert(0, str(PKG))
    try:
        from codegen.generate import generate
    finally:
        sys.path.pop(0)
    generate(str(GEN))
    OBJ.mkdir(parents=True, exist_ok=True)
    jobs = jobs or max(1, min(16, os.cpu_count() or 1))
    srcs = _sources()
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIBDIR / "libmolecular_integral.so.tmp"
    cmd = [HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", str(tmp)] + [str(o) for o in objs] + ["-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stderr[-4000:])
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB} from {len(objs)} objects")
    return LIB


def ensure_built():
    if not LIB.exists():
        build()
    return LIB


if __name__ == "__main__":
    build(verbose=True)
