"""Write every generated source of the HIP library into csrc/generated/.

Usage: python -m codegen.generate <out_dir>   (run from the package directory), or generate(out_dir).

Specialised straight-line kernels cover the classes the workloads use (orbital l <= 3, auxiliary
l <= 4, spherical output); every other class the reference's C ABI accepts (orbital l <= 6, auxiliary
l <= 9, Cartesian output) goes to the hand-written generic kernels in csrc/generic_kernels.hip.
"""

import json
import os
import sys

from . import boys_table
from .gen_kernels import (bf16_chunks, bf16_dense_width, emit_j2c, emit_j3c, emit_j3c_bf16, emit_j3c_mfma, emit_nuc,
                          emit_pair_bf16, emit_pair_dense, emit_pairE, m_chunks, mfma_chunks, mfma_dense_width,
                          pair_layout)

SPEC_MAX_L = 3
SPEC_MAX_AUX_L = 4
# Register budgets (doubles per lane) of the chunk-plan variants; variant 0 is the default.  The build's
# spill guard lowers only variant 0's budget (chunk_shrink.json); the others keep theirs, so classes whose
# default is cut to one auxiliary component per kernel still offer wider chunks to the autotuner (which
# skips variants that spill to scratch memory).
VARIANT_BUDGETS = (200, 240, 160, 120, 320, 280)
# Float64 matrix-core variants (appended after the vector ones) for pair classes with at least 9 output
# rows: budgets of live Hermite integrals per chunk (the accumulators sit in the matrix-core registers).
MFMA_PAIR_CLASSES = ((1, 1), (1, 2), (2, 2), (1, 3), (2, 3), (3, 3))
MFMA_R_BUDGETS = (84, 112, 160, 220)
# Float32-tier variants on the bf16 matrix cores (split operands, appended after the float64 matrix-core
# ones; selectable for the float32 tier only): budgets of live vector registers per chunk.
BF16_PAIR_CLASSES = ((1, 1), (0, 2), (1, 2), (0, 3), (2, 2), (1, 3), (2, 3), (3, 3))
BF16_BUDGETS = (256, 384, 512)

HEADER = """// Generated by codegen/generate.py -- do not edit.
#include "../md_device.hpp"
#include "../md_args.hpp"
"""


def _write(path, text):
    old = None
    if os.path.exists(path):
        with open(path) as f:
            old = f.read()
    if old != text:                       # keep mtimes stable for incremental builds
        with open(path, "w") as f:
            f.write(text)


STAGE_CHOICE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "stage_choice.json")


def stage_table(path=STAGE_CHOICE):
    """{"<la><lb><lc>_<prec>": "lds" | "dpp" | "ldsdpp"}: the measured staging mode per class (choose_stage.py)."""
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        return json.load(f)


def vector_plans(la, lb, lc, shrink):
    """The distinct chunk plans of the vector three-centre kernels of a class, in variant order."""
    cls_plans = []
    for vi, budget in enumerate(VARIANT_BUDGETS):
        plan = m_chunks(la, lb, lc, shrink.get((la, lb, lc), 0) if vi == 0 else 0, budget)
        if plan not in cls_plans:
            cls_plans.append(plan)
    return cls_plans


def generate(out_dir, shrink=None):
    """shrink: {(la, lb, lc): k} lowers the register budget of a class by 2^k (see m_chunks)."""
    shrink = shrink or {}
    stages = stage_table()
    os.makedirs(out_dir, exist_ok=True)
    boys_path = os.path.join(out_dir, "boys_table.inc")
    # (re)written when missing or from an older layout without the coarse LDS slices
    if not os.path.exists(boys_path) or "boys_lds64" not in open(boys_path).read():
        boys_table.emit(boys_path)

    pair_classes = [(la, lb) for lb in range(SPEC_MAX_L + 1) for la in range(lb + 1)]
    reg = ['// Generated by codegen/generate.py -- host-side dispatch to the specialised kernels.',
           '#include "../md_args.hpp"',
           f'#define EIP_SPEC_MAX_L {SPEC_MAX_L}',
           f'#define EIP_SPEC_MAX_AUX_L {SPEC_MAX_AUX_L}']
    disp = ['// Generated by codegen/generate.py -- host-side dispatch to the specialised kernels.',
            '#include "registry.hpp"', '']

    # ---- pair preparation ----
    src = [HEADER]
    nnz_table = {}
    cases = []
    for la, lb in pair_classes:
        name, nnz, code = emit_pairE(la, lb, True)
        nnz_table[(la, lb)] = nnz
        src.append(code)
        cases.append(f"        case {la * 10 + lb}: hipLaunchKernelGGL({name}, dim3(nblocks), dim3(256), 0, s, a); return 0;")
    src.append("int eip_launch_pairE(int la, int lb, const PairPrepArgs& a, int nblocks, hipStream_t s)\n{\n"
               "    switch (la * 10 + lb) {\n" + "\n".join(cases) + "\n    }\n    return -1;\n}\n")
    _write(os.path.join(out_dir, "pairE.hip"), "\n".join(src))
    reg.append("int eip_launch_pairE(int la, int lb, const PairPrepArgs& a, int nblocks, hipStream_t s);")
    disp.append("int eip_pair_nnz(int la, int lb)\n{\n    switch (la * 10 + lb) {")
    for (la, lb), nnz in nnz_table.items():
        disp.append(f"        case {la * 10 + lb}: return {nnz};")
    disp.append("    }\n    return -1;\n}\n")
    reg.append("int eip_pair_nnz(int la, int lb);")

    # ---- dense pair matrices for the matrix-core kernels ----
    src = [HEADER]
    cases = []
    for la, lb in MFMA_PAIR_CLASSES:
        name, code = emit_pair_dense(la, lb)
        src.append(code)
        cases.append(f"        case {la * 10 + lb}: hipLaunchKernelGGL({name}, dim3((n_pp + 255) / 256), dim3(256), 0, s, E, D, n_pp); return 0;")
    src.append("int eip_launch_pairD(int la, int lb, const double* E, double* D, int n_pp, hipStream_t s)\n{\n"
               "    switch (la * 10 + lb) {\n" + "\n".join(cases) + "\n    }\n    return -1;\n}\n")
    _write(os.path.join(out_dir, "pairD.hip"), "\n".join(src))
    reg.append("int eip_launch_pairD(int la, int lb, const double* E, double* D, int n_pp, hipStream_t s);")
    disp.append("int eip_pair_dense_width(int la, int lb)\n{\n    switch (la * 10 + lb) {")
    for la, lb in MFMA_PAIR_CLASSES:
        disp.append(f"        case {la * 10 + lb}: return {mfma_dense_width(la, lb)};")
    disp.append("    }\n    return -1;\n}\n")
    reg.append("int eip_pair_dense_width(int la, int lb);")

    # ---- split bfloat16 pair matrices for the float32-tier matrix-core kernels ----
    src = [HEADER]
    cases = []
    for la, lb in BF16_PAIR_CLASSES:
        name, code = emit_pair_bf16(la, lb)
        src.append(code)
        cases.append(f"        case {la * 10 + lb}: hipLaunchKernelGGL({name}, dim3((n_pp + 255) / 256), dim3(256), 0, s, E, (eip::bf8*) D, n_pp); return 0;")
    src.append("int eip_launch_pairB(int la, int lb, const double* E, void* D, int n_pp, hipStream_t s)\n{\n"
               "    switch (la * 10 + lb) {\n" + "\n".join(cases) + "\n    }\n    return -1;\n}\n")
    _write(os.path.join(out_dir, "pairB.hip"), "\n".join(src))
    reg.append("int eip_launch_pairB(int la, int lb, const double* E, void* D, int n_pp, hipStream_t s);")
    disp.append("int eip_pair_bf16_width(int la, int lb)\n{\n    switch (la * 10 + lb) {")
    for la, lb in BF16_PAIR_CLASSES:
        disp.append(f"        case {la * 10 + lb}: return {bf16_dense_width(la, lb)};")
    disp.append("    }\n    return -1;\n}\n")
    reg.append("int eip_pair_bf16_width(int la, int lb);")

    # ---- three-centre ----
    # Each class gets up to len(VARIANT_BUDGETS) chunk plans (register budgets); variant 0 is the
    # default, the others are candidates for the autotuner.  Identical chunks are emitted once.
    plan_rows = []          # (la, lb, lc, variant, chunk, m0, m1, kind: 0 vector, 1 f64 matrix cores, 2 bf16 matrix cores)
    for la, lb in pair_classes:
        plans = {}
        mplans = {}
        bplans = {}
        for lc in range(SPEC_MAX_AUX_L + 1):
            cls_plans = vector_plans(la, lb, lc, shrink)
            plans[lc] = cls_plans
            for v, plan in enumerate(cls_plans):
                for ci, (m0, m1) in enumerate(plan):
                    plan_rows.append((la, lb, lc, v, ci, m0, m1, 0))
            mplans[lc] = []
            if (la, lb) in MFMA_PAIR_CLASSES:
                for budget in MFMA_R_BUDGETS:
                    plan = mfma_chunks(la, lb, lc, budget)
                    if plan not in mplans[lc]:
                        mplans[lc].append(plan)
                for k, plan in enumerate(mplans[lc]):
                    for ci, (m0, m1) in enumerate(plan):
                        plan_rows.append((la, lb, lc, len(cls_plans) + k, ci, m0, m1, 1))
            bplans[lc] = []
            if (la, lb) in BF16_PAIR_CLASSES:
                for budget in BF16_BUDGETS:
                    plan = bf16_chunks(la, lb, lc, budget)
                    if plan not in bplans[lc]:
                        bplans[lc].append(plan)
                for k, plan in enumerate(bplans[lc]):
                    for ci, (m0, m1) in enumerate(plan):
                        plan_rows.append((la, lb, lc, len(cls_plans) + len(mplans[lc]) + k, ci, m0, m1, 2))
        if (la, lb) in MFMA_PAIR_CLASSES:
            src = [HEADER]
            cases = []
            for lc, cls_plans in mplans.items():
                for m0, m1 in sorted({c for plan in cls_plans for c in plan}):
                    name, code = emit_j3c_mfma(la, lb, lc, m0, m1)
                    src.append(code)
                    cases.append(f"        case {lc * 10000 + m0 * 100 + m1}: if (a.jobs) hipLaunchKernelGGL({name}_jobs, dim3(nblocks), dim3(256), 0, s, a);"
                                 f" else hipLaunchKernelGGL({name}, dim3(nblocks), dim3(256), 0, s, a); return 0;")
            fn = f"eip_launch_j3cm_{la}{lb}"
            src.append(f"int {fn}(int lc, int m0, int m1, const J3cArgs& a, int nblocks, hipStream_t s)\n{{\n"
                       "    switch (lc * 10000 + m0 * 100 + m1) {\n" + "\n".join(cases) + "\n    }\n    return -1;\n}\n")
            _write(os.path.join(out_dir, f"j3cm_{la}{lb}.hip"), "\n".join(src))
            reg.append(f"int {fn}(int lc, int m0, int m1, const J3cArgs& a, int nblocks, hipStream_t s);")
        if (la, lb) in BF16_PAIR_CLASSES:
            src = [HEADER]
            cases = []
            for lc, cls_plans in bplans.items():
                for m0, m1 in sorted({c for plan in cls_plans for c in plan}):
                    name, code = emit_j3c_bf16(la, lb, lc, m0, m1)
                    src.append(code)
                    cases.append(f"        case {lc * 10000 + m0 * 100 + m1}: if (a.jobs) hipLaunchKernelGGL({name}_jobs, dim3(nblocks), dim3(256), 0, s, a);"
                                 f" else hipLaunchKernelGGL({name}, dim3(nblocks), dim3(256), 0, s, a); return 0;")
            fn = f"eip_launch_j3cb_{la}{lb}"
            src.append(f"int {fn}(int lc, int m0, int m1, const J3cArgs& a, int nblocks, hipStream_t s)\n{{\n"
                       "    switch (lc * 10000 + m0 * 100 + m1) {\n" + "\n".join(cases) + "\n    }\n    return -1;\n}\n")
            _write(os.path.join(out_dir, f"j3cb_{la}{lb}.hip"), "\n".join(src))
            reg.append(f"int {fn}(int lc, int m0, int m1, const J3cArgs& a, int nblocks, hipStream_t s);")
        for prec in ("f64", "f32"):
            src = [HEADER]
            cases = []
            for lc, cls_plans in plans.items():
                for m0, m1 in sorted({c for plan in cls_plans for c in plan}):
                    name, code = emit_j3c(la, lb, lc, m0, m1, prec, stage=stages.get(f"{la}{lb}{lc}_{prec}"))
                    src.append(code)
                    cases.append(f"        case {lc * 10000 + m0 * 100 + m1}: if (a.jobs) hipLaunchKernelGGL({name}_jobs, dim3(nblocks), dim3(64), 0, s, a);"
                                 f" else hipLaunchKernelGGL({name}, dim3(nblocks), dim3(64), 0, s, a); return 0;")
            fn = f"eip_launch_j3c_{la}{lb}_{prec}"
            src.append("#ifndef EIP_BODIES_ONLY")
            src.append(f"int {fn}(int lc, int m0, int m1, const J3cArgs& a, int nblocks, hipStream_t s)\n{{\n"
                       "    switch (lc * 10000 + m0 * 100 + m1) {\n" + "\n".join(cases) + "\n    }\n    return -1;\n}\n")
            src.append("#endif")
            _write(os.path.join(out_dir, f"j3c_{la}{lb}_{prec}.hip"), "\n".join(src))
            reg.append(f"int {fn}(int lc, int m0, int m1, const J3cArgs& a, int nblocks, hipStream_t s);")
    reg.append("int eip_launch_j3c(int la, int lb, int lc, int variant, int chunk, int f32, const J3cArgs& a, int nblocks, hipStream_t s);")
    reg.append("int eip_j3c_nvariants(int la, int lb, int lc);")
    reg.append("int eip_j3c_nchunks(int la, int lb, int lc, int variant);")
    reg.append("int eip_j3c_chunk(int la, int lb, int lc, int variant, int chunk, int* m0, int* m1);")
    disp.append("static const short kChunks[][8] = {")
    for row in plan_rows:
        disp.append("    {%d, %d, %d, %d, %d, %d, %d, %d}," % row)
    disp.append("};")
    disp.append("int eip_j3c_chunk(int la, int lb, int lc, int variant, int chunk, int* m0, int* m1)\n{\n"
                "    for (const auto& r : kChunks)\n"
                "        if (r[0] == la && r[1] == lb && r[2] == lc && r[3] == variant && r[4] == chunk) { *m0 = r[5]; *m1 = r[6]; return 0; }\n"
                "    return -1;\n}\n")
    disp.append("int eip_j3c_nchunks(int la, int lb, int lc, int variant)\n{\n    int n = 0;\n"
                "    for (const auto& r : kChunks) n += (r[0] == la && r[1] == lb && r[2] == lc && r[3] == variant);\n    return n;\n}\n")
    disp.append("int eip_j3c_nvariants(int la, int lb, int lc)\n{\n    int n = 0;\n"
                "    for (const auto& r : kChunks) if (r[0] == la && r[1] == lb && r[2] == lc && r[3] + 1 > n) n = r[3] + 1;\n    return n;\n}\n")
    disp.append("int eip_j3c_variant_mfma(int la, int lb, int lc, int variant)\n{\n"
                "    for (const auto& r : kChunks)\n"
                "        if (r[0] == la && r[1] == lb && r[2] == lc && r[3] == variant) return r[7];\n    return -1;\n}\n")
    reg.append("int eip_j3c_variant_mfma(int la, int lb, int lc, int variant);")
    # persistent queue kernels (second generation pass, generate_queue: needs the compiled registers)
    reg.append("int eip_j3cq_ntiers(int f32);")
    reg.append("int eip_j3cq_chunk(int la, int lb, int lc, int f32, int chunk, int* m0, int* m1, int* tier, int* kind);")
    reg.append("int eip_launch_j3cq(int f32, int tier, const J3cQueueArgs& q, int nblocks, hipStream_t s);")
    reg.append("const void* eip_j3cq_kernel(int f32, int tier);")
    disp.append("int eip_launch_j3c(int la, int lb, int lc, int variant, int chunk, int f32, const J3cArgs& a, int nblocks, hipStream_t s)\n{")
    disp.append("    int m0, m1;")
    disp.append("    if (eip_j3c_chunk(la, lb, lc, variant, chunk, &m0, &m1) != 0) return -1;")
    disp.append("    if (eip_j3c_variant_mfma(la, lb, lc, variant) == 2) {")
    disp.append("        if (!f32) return -1;                   // bf16 matrix-core kernels are the float32 tier only")
    disp.append("        switch (la * 10 + lb) {")
    for la, lb in BF16_PAIR_CLASSES:
        disp.append(f"            case {la * 10 + lb}: return eip_launch_j3cb_{la}{lb}(lc, m0, m1, a, nblocks, s);")
    disp.append("        }")
    disp.append("        return -1;")
    disp.append("    }")
    disp.append("    if (eip_j3c_variant_mfma(la, lb, lc, variant) == 1) {")
    disp.append("        if (f32) return -1;                    // float64 matrix-core kernels")
    disp.append("        switch (la * 10 + lb) {")
    for la, lb in MFMA_PAIR_CLASSES:
        disp.append(f"            case {la * 10 + lb}: return eip_launch_j3cm_{la}{lb}(lc, m0, m1, a, nblocks, s);")
    disp.append("        }")
    disp.append("        return -1;")
    disp.append("    }")
    disp.append("    switch (la * 10 + lb) {")
    for la, lb in pair_classes:
        disp.append(f"        case {la * 10 + lb}: return f32 ? eip_launch_j3c_{la}{lb}_f32(lc, m0, m1, a, nblocks, s)"
                    f" : eip_launch_j3c_{la}{lb}_f64(lc, m0, m1, a, nblocks, s);")
    disp.append("    }\n    return -1;\n}\n")

    # ---- nuclear attraction ----
    src = [HEADER]
    cases = []
    for la, lb in pair_classes:
        name, code = emit_nuc(la, lb)
        src.append(code)
        cases.append(f"        case {la * 10 + lb}: hipLaunchKernelGGL({name}, dim3(nblocks), dim3(64), 0, s, a); return 0;")
    src.append("int eip_launch_nuc(int la, int lb, const NucArgs& a, int nblocks, hipStream_t s)\n{\n"
               "    switch (la * 10 + lb) {\n" + "\n".join(cases) + "\n    }\n    return -1;\n}\n")
    _write(os.path.join(out_dir, "nuc.hip"), "\n".join(src))
    reg.append("int eip_launch_nuc(int la, int lb, const NucArgs& a, int nblocks, hipStream_t s);")

    # ---- two-centre ----
    src = [HEADER]
    cases = []
    for lj in range(SPEC_MAX_AUX_L + 1):
        for li in range(lj + 1):
            name, code = emit_j2c(li, lj)
            src.append(code)
            cases.append(f"        case {li * 10 + lj}: hipLaunchKernelGGL({name}, dim3(nblocks), dim3(64), 0, s, a); return 0;")
    src.append("int eip_launch_j2c(int li, int lj, const J2cArgs& a, int nblocks, hipStream_t s)\n{\n"
               "    switch (li * 10 + lj) {\n" + "\n".join(cases) + "\n    }\n    return -1;\n}\n")
    _write(os.path.join(out_dir, "j2c.hip"), "\n".join(src))
    reg.append("int eip_launch_j2c(int li, int lj, const J2cArgs& a, int nblocks, hipStream_t s);")

    _write(os.path.join(out_dir, "registry.hpp"), "\n".join(reg) + "\n")
    _write(os.path.join(out_dir, "registry.cpp"), "\n".join(disp) + "\n")
    emit_generic_tables(out_dir)
    return sorted(f for f in os.listdir(out_dir) if f.endswith((".hip", ".cpp")))


def emit_generic_tables(out_dir):
    """Tables for the runtime-generic kernels (every l <= 9, spherical and Cartesian).

    * basis rows: for (l, spherical) the output rows as lists of (cartesian index, coefficient);
    * auxiliary rows: for (l, spherical) the Hermite-to-output transform rows as lists of
      (sx, sy, sz, j, coefficient), see tables.aux_transform;
    * reverse Hermite coefficients H(t, m) = t! / ((2m-t)! (t-m)! 2^(t-m)) for the closed form of R.
    """
    from math import factorial
    from .tables import SqrtRational, aux_transform, out_basis_rows

    lmax = 9
    brow_off, bterm_c, bterm_v = [], [], []
    brow_base = []          # per (l, sph): first row index into brow_off
    arow_off, aterm, aterm_v = [], [], []
    arow_base = []
    for sph in (0, 1):
        for l in range(lmax + 1):
            brow_base.append(len(brow_off))
            for r, terms in out_basis_rows(l, bool(sph)):
                brow_off.append(len(bterm_c))
                for c, q in sorted(terms.items()):
                    bterm_c.append(c)
                    bterm_v.append(SqrtRational(r, q).to_float())
            arow_base.append(len(arow_off))
            for r, terms in aux_transform(l, bool(sph)):
                arow_off.append(len(aterm))
                for s, (q, j) in sorted(terms.items()):
                    aterm.append(s[0] | (s[1] << 5) | (s[2] << 10) | (j << 15))
                    aterm_v.append(SqrtRational(r, q).to_float())
    brow_off.append(len(bterm_c))
    arow_off.append(len(aterm))
    H = []
    for t in range(25):
        row = []
        for m in range(25):
            if m < (t + 1) // 2 or m > t:
                row.append(0.0)
            else:
                row.append(factorial(t) / (factorial(2 * m - t) * factorial(t - m) * 2 ** (t - m)))
        H.append(row)

    def arr(ctype, name, vals, f=str):
        return f"__device__ const {ctype} {name}[{len(vals)}] = {{" + ", ".join(f(v) for v in vals) + "};"

    from .tables import fmt
    lines = ["// Generated by codegen/generate.py -- tables for the generic kernels.", "#pragma once",
             "#define GEN_LMAX 9",
             arr("int", "gen_brow_base", brow_base), arr("int", "gen_brow_off", brow_off),
             arr("int", "gen_bterm_c", bterm_c), arr("double", "gen_bterm_v", bterm_v, fmt),
             arr("int", "gen_arow_base", arow_base), arr("int", "gen_arow_off", arow_off),
             arr("int", "gen_aterm", aterm), arr("double", "gen_aterm_v", aterm_v, fmt),
             arr("double", "gen_herm", [v for row in H for v in row], fmt)]
    _write(os.path.join(out_dir, "generic_tables.inc"), "\n".join(lines) + "\n")


if __name__ == "__main__":
    print(generate(sys.argv[1]))


# ---------------------------------------------------------------------------------------------
# Persistent queue kernels (second pass: tiers come from the compiled kernels' register use)
# ---------------------------------------------------------------------------------------------

TUNING_FILE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "tuning",
                           "three_center_variants.json")
# Register tiers of the persistent kernels, by the occupancy (waves per SIMD) of the class chunk's own
# kernel: a tier kernel runs at the occupancy of its largest body, so bodies are grouped with bodies of
# the same occupancy.
QUEUE_TIERS = ((4, 99), (3, 3), (2, 2), (1, 1))


# Largest register footprint (VGPRs + AGPRs per lane, of the chunk's own kernel) of a body in a queue
# kernel: the persistent loop around the bodies needs a few registers of its own, so a body at the
# 512-register limit would spill there.
QUEUE_MAX_REGS = 448


def _regs(resources, la, lb, lc, m0, m1, prec):
    r = resources.get(f"j3c_{la}{lb}{lc}_{m0}_{m1}" + ("_f32" if prec == "f32" else "") + "_jobs", {})
    return r.get("vgpr", 512) + r.get("agpr", 0)


def queue_plan(la, lb, lc, prec, shrink, variants, resources=None):
    """Chunk plan of a class in the queue kernels: the autotuned vector variant (tuning table), or the
    default plan where the table picks a matrix-core variant (those run four waves per workgroup); when a
    chunk of that plan needs more than QUEUE_MAX_REGS registers, the vector plan with the fewest chunks
    among those that stay within it."""
    plans = vector_plans(la, lb, lc, shrink)
    v = variants.get(f"{la}{lb}{lc}_{prec}", 0)
    plan = plans[v] if v < len(plans) else plans[0]
    if resources is None:
        return plan
    fits = lambda p: all(_regs(resources, la, lb, lc, m0, m1, prec) <= QUEUE_MAX_REGS for m0, m1 in p)
    if fits(plan):
        return plan
    ok = [p for p in plans if fits(p)]
    # None: no plan fits -- the class is left to its own kernels
    return min(ok, key=len) if ok else None


def generate_queue(out_dir, resources, shrink=None):
    """Write j3cq_<prec>_t<tier>.hip (one persistent kernel each) and j3cq_registry.cpp; returns the
    source paths.  resources: {kernel name: {"occupancy": ...}} of the compiled class kernels."""
    from .gen_kernels import j3c_smem_bytes
    shrink = shrink or {}
    stages = stage_table()
    variants = {}
    if os.path.exists(TUNING_FILE):
        with open(TUNING_FILE) as f:
            variants = json.load(f).get("variants", {})
    pair_classes = [(la, lb) for lb in range(SPEC_MAX_L + 1) for la in range(lb + 1)]
    rows = []
    bodies = {(f32, t): [] for f32 in (0, 1) for t in range(len(QUEUE_TIERS))}
    for f32, prec in ((0, "f64"), (1, "f32")):
        for la, lb in pair_classes:
            for lc in range(SPEC_MAX_AUX_L + 1):
                for ci, (m0, m1) in enumerate(queue_plan(la, lb, lc, prec, shrink, variants, resources) or []):
                    name = f"j3c_{la}{lb}{lc}_{m0}_{m1}" + ("_f32" if f32 else "")
                    occ = resources.get(name + "_jobs", {}).get("occupancy", 1)
                    tier = next((t for t, (lo, hi) in enumerate(QUEUE_TIERS) if lo <= occ <= hi), 0)
                    smem = j3c_smem_bytes(la, lb, lc, m0, m1, prec, stages.get(f"{la}{lb}{lc}_{prec}"))
                    kind = len(bodies[(f32, tier)])
                    bodies[(f32, tier)].append((name, smem, (la, lb)))
                    rows.append((la, lb, lc, f32, ci, m0, m1, tier, kind))
    paths = []
    reg = ['// Generated by codegen/generate.py -- dispatch to the persistent queue kernels.',
           '#include "registry.hpp"', '']
    for (f32, tier), bl in bodies.items():
        prec = "f32" if f32 else "f64"
        kname = f"j3cq_{prec}_t{tier}"
        smem = max([16] + [b[1] for b in bl])
        src = [HEADER, '#include "../queue_device.hpp"', "#define EIP_BODIES_ONLY"]
        for la, lb in sorted({b[2] for b in bl}):
            src.append(f'#include "j3c_{la}{lb}_{prec}.hip"')
        src.append("#undef EIP_BODIES_ONLY")
        src.append(f"// {len(bl)} class-chunk bodies of occupancy {QUEUE_TIERS[tier][0]}-{QUEUE_TIERS[tier][1]}")
        # the tier's occupancy as a register target (the bodies fit it alone; the loop around them adds a few)
        src.append(f'extern "C" __global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu({QUEUE_TIERS[tier][0]})))'
                   f' {kname}(const J3cQueueArgs q)')
        src.append("{")
        src.append(f"    __shared__ __attribute__((aligned(16))) char smem[{smem}];")
        src.append("    int part = (int) (blockIdx.x % kQueueParts);")
        src.append("    for (;;) {")
        src.append("        const int e = eip::queue_next(q.ctl, part);")
        src.append("        if (e < 0) break;")
        src.append("        const eip::cptr<J3cQueueEntry> en = eip::ro(q.entries) + e;")
        src.append("        const int slot = __builtin_amdgcn_readfirstlane(en->slot);")
        src.append("        const int item0 = __builtin_amdgcn_readfirstlane(en->item);")
        src.append("        const int count = __builtin_amdgcn_readfirstlane(en->count);")
        src.append("        const eip::cptr<J3cQueueSlot> sl = eip::ro(q.slots) + slot;")
        # a fresh, opaque copy of the lane index and the LDS buffer per item: the bodies' lane- and
        # buffer-derived values are then computed inside the item, not hoisted out of this loop (where
        # every body's values would stay live at once)
        src.append("        int tid = (int) threadIdx.x, smo = 0;")
        src.append('        asm volatile("" : "+v"(tid), "+s"(smo));')
        src.append("        char* const sm = smem + smo;")
        src.append("        switch (sl->kind) {")
        for kind, (name, _, _) in enumerate(bl):
            src.append(f"            case {kind}: for (int k = 0; k < count; ++k) {name}_body<true>(sl->a, item0 + k, sm, tid); break;")
        src.append("            default: break;")
        src.append("        }")
        src.append("    }")
        src.append("    eip::queue_finish(q.ctl);")
        src.append("}")
        src.append(f"int eip_launch_{kname}(const J3cQueueArgs& q, int nblocks, hipStream_t s)\n{{\n"
                   f"    hipLaunchKernelGGL({kname}, dim3(nblocks), dim3(64), 0, s, q);\n    return 0;\n}}")
        src.append(f"const void* eip_{kname}_kernel() {{ return (const void*) {kname}; }}")
        path = os.path.join(out_dir, kname + ".hip")
        _write(path, "\n".join(src) + "\n")
        paths.append(path)
        reg.append(f"int eip_launch_{kname}(const J3cQueueArgs& q, int nblocks, hipStream_t s);")
        reg.append(f"const void* eip_{kname}_kernel();")
    reg.append("static const short kQueueChunks[][9] = {")
    for row in rows:
        reg.append("    {%d, %d, %d, %d, %d, %d, %d, %d, %d}," % row)
    reg.append("};")
    reg.append(f"int eip_j3cq_ntiers(int f32) {{ (void) f32; return {len(QUEUE_TIERS)}; }}")
    reg.append("int eip_j3cq_chunk(int la, int lb, int lc, int f32, int chunk, int* m0, int* m1, int* tier, int* kind)\n{\n"
               "    for (const auto& r : kQueueChunks)\n"
               "        if (r[0] == la && r[1] == lb && r[2] == lc && r[3] == f32 && r[4] == chunk) {\n"
               "            *m0 = r[5]; *m1 = r[6]; *tier = r[7]; *kind = r[8]; return 0;\n        }\n"
               "    return -1;\n}")
    reg.append("int eip_launch_j3cq(int f32, int tier, const J3cQueueArgs& q, int nblocks, hipStream_t s)\n{\n"
               "    switch (f32 * 10 + tier) {")
    for f32, tier in bodies:
        prec = "f32" if f32 else "f64"
        reg.append(f"        case {f32 * 10 + tier}: return eip_launch_j3cq_{prec}_t{tier}(q, nblocks, s);")
    reg.append("    }\n    return -1;\n}")
    reg.append("const void* eip_j3cq_kernel(int f32, int tier)\n{\n    switch (f32 * 10 + tier) {")
    for f32, tier in bodies:
        prec = "f32" if f32 else "f64"
        reg.append(f"        case {f32 * 10 + tier}: return eip_j3cq_{prec}_t{tier}_kernel();")
    reg.append("    }\n    return nullptr;\n}")
    path = os.path.join(out_dir, "j3cq_registry.cpp")
    _write(path, "\n".join(reg) + "\n")
    paths.append(path)
    return paths
