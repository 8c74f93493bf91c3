"""Build libmolecular_integral.so for gfx950: run the kernel generator, compile each source with hipcc
in parallel, link.  ``build()`` always recompiles; ``ensure_built()`` only builds when the library is
missing.
"""

import concurrent.futures as cf
import json
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
    # the persistent queue kernels (j3cq_*) are generated and compiled in a second pass
    srcs = sorted(p for p in GEN.glob("*.hip") if not p.name.startswith("j3cq_")) + [
        GEN / "registry.cpp", CSRC / "generic_kernels.hip", CSRC / "library.hip", CSRC / "mixed_estimate.hip",
        CSRC / "pair_screening.cpp"]
    # largest first, so the long compiles start early
    return sorted(srcs, key=lambda p: -p.stat().st_size)


def _compile(src):
    """Compile one source; returns (object, {kernel: scratch bytes per lane})."""
    obj = OBJ / (src.stem + ".o")
    cmd = [HIPCC] + FLAGS + ["-Rpass-analysis=kernel-resource-usage", "-c", str(src), "-o", str(obj)]
    if src.suffix == ".cpp":
        cmd = [HIPCC, "-x", "hip"] + FLAGS + ["-c", str(src), "-o", str(obj)]
    elif src.name.startswith("j3cq_"):
        # Persistent queue kernels: machine LICM would hoist every body's constants (float64 literals in
        # scalar registers) out of the persistent loop, where all of them stay live at once (hundreds of
        # scalar-register spills); without it each body keeps the register footprint of its own kernel.
        cmd = cmd[:1] + ["-mllvm", "-disable-machine-licm"] + cmd[1:]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src.name}:\n{r.stderr[-4000:]}")
    res, name = {}, None
    keys = {"VGPRs:": "vgpr", "AGPRs:": "agpr", "ScratchSize [bytes/lane]:": "scratch",
            "Occupancy [waves/SIMD]:": "occupancy", "VGPRs Spill:": "vgpr_spill", "SGPRs Spill:": "sgpr_spill"}
    for line in r.stderr.splitlines():
        if "Function Name:" in line:
            name = line.split("Function Name:")[1].split()[0]
            res[name] = {}
            continue
        for k, v in keys.items():
            if k in line and name:
                res[name][v] = int(line.split(k)[1].split()[0])
    return obj, res


def spills(res):
    """A kernel spills when it moves registers to scratch memory.  (VGPRs spilled to AGPRs and SGPRs
    spilled to VGPR lanes stay in registers; a few dozen bytes of private segment hold a few SGPR spills
    or the backend's reserved scavenging slots.)"""
    return res.get("scratch", 0) > 128


def _spilling_classes(scratch, shrink):
    """Three-centre classes (la, lb, lc) whose default chunk plan has a kernel spilling to scratch memory."""
    sys.path.insert(0, str(PKG))
    try:
        from codegen.gen_kernels import m_chunks
    finally:
        sys.path.pop(0)
    out = set()
    for name, nbytes in scratch.items():
        # the job-table twins ("_jobs") share their body with the kernarg kernels; the kernarg kernels decide
        if nbytes > 0 and name.startswith("j3c_") and not name.endswith("_jobs"):
            parts = name.split("_")
            cls = parts[1]
            la, lb, lc = int(cls[0]), int(cls[1]), int(cls[2])
            if (int(parts[2]), int(parts[3])) in m_chunks(la, lb, lc, shrink.get((la, lb, lc), 0)):
                out.add((la, lb, lc))
    return out


def build(jobs=None, verbose=False):
    sys.path.insert(0, str(PKG))
    try:
        from codegen.generate import generate
    finally:
        sys.path.pop(0)
    # Kernels must not spill to scratch memory: classes whose kernels still do are regenerated with
    # smaller auxiliary-component chunks (up to 3 times); the final table is kept for the next build.
    shrink_file = PKG / "codegen" / "chunk_shrink.json"
    shrink = {}
    if shrink_file.exists():
        shrink = {tuple(json.loads(k)): v for k, v in json.loads(shrink_file.read_text()).items()}
    OBJ.mkdir(parents=True, exist_ok=True)
    jobs = jobs or max(1, min(16, os.cpu_count() or 1))
    results = {}
    for attempt in range(4):
        before = {p: p.stat().st_mtime_ns for p in GEN.glob("*")}
        generate(str(GEN), shrink)
        srcs = [p for p in _sources() if attempt == 0 or before.get(p) != p.stat().st_mtime_ns]
        with cf.ThreadPoolExecutor(jobs) as ex:
            for src, res in zip(srcs, ex.map(_compile, srcs)):
                results[src] = res
        objs = [o for o, _ in results.values()]
        resources = {k: v for _, sc in results.values() for k, v in sc.items()}
        scratch = {k: (v.get("scratch", 0) if spills(v) else 0) for k, v in resources.items()}
        spilling = {c for c in _spilling_classes(scratch, shrink) if shrink.get(c, 0) < 3}
        if not spilling or attempt == 3:
            break
        for c in spilling:
            shrink[c] = shrink.get(c, 0) + 1
        if verbose:
            print("regenerating with smaller chunks:", sorted(spilling))
    shrink_file.write_text(json.dumps({json.dumps(list(k)): v for k, v in shrink.items()}))
    # Second pass: the persistent queue kernels, one per register tier (tiers from the occupancy the class
    # kernels just compiled to).
    sys.path.insert(0, str(PKG))
    try:
        from codegen.generate import generate_queue
    finally:
        sys.path.pop(0)
    qsrcs = [pathlib.Path(p) for p in generate_queue(str(GEN), {k: v for _, sc in results.values() for k, v in sc.items()},
                                                     shrink)]
    qsrcs.sort(key=lambda p: -p.stat().st_size)
    with cf.ThreadPoolExecutor(jobs) as ex:
        for src, res in zip(qsrcs, ex.map(_compile, qsrcs)):
            results[src] = res
    objs = [o for o, _ in results.values()]
    resources = {k: v for _, sc in results.values() for k, v in sc.items()}
    # per-kernel registers / scratch / occupancy, read by the autotuner (variants that spill are skipped)
    LIBDIR.mkdir(parents=True, exist_ok=True)
    (LIBDIR / "kernel_resources.json").write_text(json.dumps(resources, indent=0, sort_keys=True))
    still = sorted(k for k, v in scratch.items() if v > 0)
    if verbose and still:
        print("kernels still using scratch:", still)
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
