"""Kernel generator checks that need no GPU: the register layout of pair-matrix rows in the DPP staging
mode, the two-stage two-centre contraction, and the per-class staging choice."""

import math
import re
import sys

import pytest

from electron_integral_playground_amd.codegen import choose_stage, gen_kernels
from electron_integral_playground_amd.codegen.gen_kernels import emit_j2c, emit_j3c, pair_nnz


@pytest.mark.parametrize("la,lb,lc,m0,m1,prec", [(0, 0, 0, 0, 1, "f64"), (0, 1, 3, 0, 7, "f64"),
                                                 (1, 2, 3, 0, 7, "f64"), (2, 2, 3, 4, 7, "f32")])
def test_dpp_rows_cover_every_entry_once(la, lb, lc, m0, m1, prec):
    """Entry k of a row is broadcast from register k // 16, lane k % 16, exactly once per contraction,
    each register is reloaded (next primitive pair) only after its last use, and the per-lane load
    index stays inside the row."""
    _, code = emit_j3c(la, lb, lc, m0, m1, prec, stage="dpp")
    nnz = pair_nnz(la, lb, True)
    nv = (nnz + 15) // 16
    uses = [(int(a), int(b)) for a, b in re.findall(r"eip::bcast16<(\d+)>\(Er_(\d+)\)", code)]
    assert uses == [(k % 16, k // 16) for k in range(nnz)]
    body = code.split("for (int pp = pp0; pp < pp1; ++pp) {", 1)[1]
    for v in range(nv):
        last_use = body.rfind(f"(Er_{v})")
        reload = body.find(f"Er_{v} = Erow[(size_t) ppn")
        assert 0 <= last_use < reload
    idx = dict(re.findall(r"const int ei_(\d+) = (.*?);", code))
    assert len(idx) == nv
    for v, e in idx.items():
        v = int(v)
        hi = 16 * v + 15
        assert (f"min({16 * v} + l16, {nnz - 1})" == e) if hi >= nnz else (e == f"{16 * v} + l16")
    assert "__shared__" not in code and "__syncthreads" not in code


@pytest.mark.parametrize("la,lb,lc,m0,m1,prec", [(0, 0, 0, 0, 1, "f64"), (0, 2, 1, 0, 3, "f64"),
                                                 (2, 3, 4, 0, 2, "f64"), (3, 3, 3, 4, 6, "f32")])
def test_ldsdpp_groups_are_loaded_before_use_and_kept_until_done(la, lb, lc, m0, m1, prec):
    """Rows staged in LDS and read 16 entries per load: replaying the generated statements in order,
    every broadcast of entry k reads register Eg_{(k//16) % 3} while it holds group k // 16 (loaded from
    LDS offset 16 * (k // 16) + lane, clamped to the row), entries are used in order exactly once per
    contraction, and no group is loaded twice."""
    _, code = emit_j3c(la, lb, lc, m0, m1, prec, stage="ldsdpp")
    nnz = pair_nnz(la, lb, True)
    assert "__shared__" in code
    held = {}
    loads = []
    uses = []
    for line in code.splitlines():
        m = re.search(r"\bEg_(\d) = E\[min\((?:(\d+) \+ )?l16s, (\d+)\)\]", line)
        if m:
            g = int(m.group(2) or 0) // 16
            assert int(m.group(3)) == nnz - 1 and int(m.group(2) or 0) == 16 * g
            held[int(m.group(1))] = g
            loads.append(g)
        m = re.search(r"eip::bcast16<(\d+)>\(Eg_(\d)\)", line)
        if m:
            j, r = int(m.group(1)), int(m.group(2))
            k = len(uses)
            assert (j, r) == (k % 16, (k // 16) % 3), f"entry {k}"
            assert held.get(r) == k // 16, f"entry {k}: register {r} holds group {held.get(r)}"
            uses.append(k)
    assert len(uses) == nnz and sorted(loads) == loads == list(range((nnz + 15) // 16))


def test_large_rows_stay_in_lds():
    la, lb = 3, 3
    assert pair_nnz(la, lb, True) > gen_kernels.DPP_MAX_ENTRIES
    _, code = emit_j3c(la, lb, 0, 0, 1, "f64", stage="dpp")
    assert "bcast16" not in code and "__shared__" in code


@pytest.mark.parametrize("li,lj", [(0, 0), (2, 4), (4, 4)])
def test_two_center_contraction_uses_one_sided_coefficients(li, lj):
    """The (P|Q) kernel contracts in two stages: W_mj(t) from the j side's transform, then the i side's;
    every accumulator is updated once per primitive pair."""
    _, code = emit_j2c(li, lj)
    n_i, n_j = 2 * li + 1, 2 * lj + 1
    for mi in range(n_i):
        for mj in range(n_j):
            assert code.count(f"acc_{mi}_{mj} += ") == 1
    assert len(re.findall(r"const double W_(\d+)_", code)) >= n_j


def test_choose_stage_needs_a_clear_win(tmp_path):
    lds = tmp_path / "lds.log"
    dpp = tmp_path / "dpp.log"
    lds.write_text("  autotune 123_f64 variant 0: 2.000 ms\n  autotune 123_f64 variant 1: 1.500 ms\n"
                   "  autotune 011_f64 variant 0: 1.000 ms\n  autotune 333_f64 variant 0: 5.000 ms\n")
    dpp.write_text("  autotune 123_f64 variant 0: 1.400 ms\n"
                   "  autotune 011_f64 variant 0: 0.990 ms\n  autotune 333_f64 variant 0: 1.000 ms\n")
    table = choose_stage.choose([str(lds)], [str(dpp)])
    # 1.4 < 0.98 * 1.5: dpp; 0.99 is not 2 % faster than 1.0: lds; (ff) rows never leave LDS
    assert table == {"011_f64": "lds", "123_f64": "dpp", "333_f64": "lds"}
    # the LDS-read + DPP-broadcast mode replaces "lds" only, and only on a 2 % win
    alt = tmp_path / "ldsdpp.log"
    alt.write_text("  autotune 011_f64 variant 0: 0.900 ms\n  autotune 123_f64 variant 0: 1.000 ms\n"
                   "  autotune 333_f64 variant 0: 4.950 ms\n")
    assert choose_stage.upgrade_lds(table, [str(lds)], [str(alt)]) == {"011_f64": "ldsdpp", "123_f64": "dpp",
                                                                       "333_f64": "lds"}


def _eval_generated(lines, env):
    """Evaluate generated ``const double name = expr;`` lines in exact rational arithmetic."""
    from fractions import Fraction
    for line in lines:
        m = re.match(r"\s*const double (\w+) = (.*);", line)
        expr = re.sub(r"(\d+\.\d+)", lambda mm: f"Fraction('{mm.group(1)}')", m.group(2))
        env[m.group(1)] = eval(expr, {"Fraction": Fraction}, env)
    return env


def test_pair_hermite_recurrence_equals_reference_transfer_exactly():
    """The generated two-index recurrence gives, in exact arithmetic, the coefficients of the
    reference's vertical recurrence plus binomial transfer E^{ij}_t = sum_m C(j, m) AB^{j-m} E^{i+m,0}_t
    (mcmurchie_davidson_E_term.hpp:22-45, 166-199) for every (i, j, t) up to l = 3."""
    from fractions import Fraction as F
    from math import comb
    import random
    rng = random.Random(5)
    for la in range(4):
        for lb in range(4):
            PA, AB, o2p = F(rng.randint(-60, 60), 7), F(rng.randint(-60, 60), 3), F(1, rng.randint(1, 50))
            lines = []
            gen_kernels._emit_E_axis(lines, "x", la, lb)
            env = _eval_generated(lines, {"PAx": PA, "PBx": PA + AB, "o2p": o2p})
            E0 = {(0, 0): F(1)}
            for i in range(la + lb):
                for t in range(i + 2):
                    E0[(i + 1, t)] = ((o2p * E0[(i, t - 1)] if t >= 1 else 0) + (PA * E0[(i, t)] if t <= i else 0)
                                      + ((t + 1) * E0[(i, t + 1)] if t + 1 <= i else 0))
            for i in range(la + 1):
                for j in range(lb + 1):
                    for t in range(i + j + 1):
                        ref = sum((comb(j, j - m) * AB ** (j - m) * E0[(i + m, t)] for m in range(j + 1) if t <= i + m),
                                  F(0))
                        assert env[f"Ex_{i}_{j}_{t}"] == ref, (la, lb, i, j, t)


def test_pair_hermite_recurrence_is_stable_for_a_tight_second_centre():
    """Diffuse A (a = 0.087), tight B (b = 4790), |AB| = 5.3 Bohr, (d f): in float64 the transfer
    formula loses most digits of E^{2,3}_0 while the generated recurrence keeps it to 1e-12 relative."""
    from fractions import Fraction as F
    from math import comb
    a, b, AB = 0.0866, 4790.0, 5.28
    p = a + b
    PA, PB, o2p = b / p * AB, -a / p * AB, 0.5 / p         # P - A = b/p (B - A) up to sign convention
    PA, PB = -PA, -PB
    lines = []
    gen_kernels._emit_E_axis(lines, "x", 2, 3)
    exact = _eval_generated(lines, {"PAx": F(PA), "PBx": F(PB), "o2p": F(o2p)})["Ex_2_3_0"]
    env = {"PAx": PA, "PBx": PB, "o2p": o2p}
    for line in lines:
        m = re.match(r"\s*const double (\w+) = (.*);", line)
        env[m.group(1)] = eval(m.group(2), {}, env)
    rec = env["Ex_2_3_0"]
    E0 = {(0, 0): 1.0}
    for i in range(5):
        for t in range(i + 2):
            E0[(i + 1, t)] = ((o2p * E0[(i, t - 1)] if t >= 1 else 0.0) + (PA * E0[(i, t)] if t <= i else 0.0)
                              + ((t + 1) * E0[(i, t + 1)] if t + 1 <= i else 0.0))
    ab = PB - PA                                         # A - B
    hrr = sum(comb(3, 3 - m) * ab ** (3 - m) * E0[(2 + m, 0)] for m in range(4))
    assert abs(rec - float(exact)) <= 1e-12 * abs(float(exact))
    assert abs(hrr - float(exact)) > 1e3 * abs(rec - float(exact))


def test_coarse_boys_slices_reproduce_mpmath():
    """The coarse grid the LDS-staged Boys function reads (h = 1/2, BOYS_KC + 1 Taylor terms about the
    nearest grid point, then the downward recursion) reproduces F_m(T) to 1e-15 relative for every order
    with a staged slice; the slice size the kernels reserve matches the table layout."""
    import mpmath
    import numpy as np
    from electron_integral_playground_amd.codegen import boys_table as bt
    rows = bt.grid_values(bt.BOYS_C_H_INV)
    nrow = len(rows)
    assert nrow == int(bt.BOYS_T_SWITCH * bt.BOYS_C_H_INV) + 1
    assert gen_kernels.BOYS_LDS_BYTES == 8 * (((bt.BOYS_KC + 1) * nrow + 1) // 2 * 2)
    assert gen_kernels.BOYS_LDS_LMAX == bt.BOYS_LDS_LMAX and bt.BOYS_LDS_LMAX + bt.BOYS_KC <= bt.BOYS_NMAX
    rng = np.random.default_rng(4)
    mpmath.mp.dps = 40
    for T in list(rng.uniform(0, bt.BOYS_T_SWITCH, 40)) + [0.0, 0.2499, 0.25, 35.9]:
        k = int(round(T * bt.BOYS_C_H_INV))
        x = k / bt.BOYS_C_H_INV - T
        eT = math.exp(-T)
        for L in (0, 3, 9, bt.BOYS_LDS_LMAX):
            acc = rows[k][L + bt.BOYS_KC]
            for j in range(bt.BOYS_KC - 1, -1, -1):
                acc = x * (1.0 / (j + 1)) * acc + rows[k][L + j]
            F = [0.0] * (L + 1)
            F[L] = acc
            for m in range(L - 1, -1, -1):
                F[m] = (2 * T * F[m + 1] + eT) / (2 * m + 1)
            for m in range(L + 1):
                ref = mpmath.mpf(1) / (2 * m + 1) if T == 0 else \
                    mpmath.gammainc(m + 0.5, 0, T) / (2 * mpmath.mpf(T) ** (m + 0.5))
                assert abs(F[m] - float(ref)) <= 1e-15 * float(ref), (T, L, m)


@pytest.mark.parametrize("la,lb,lc,m0,m1,prec", [(0, 0, 0, 0, 1, "f64"), (1, 2, 3, 0, 7, "f64"),
                                                 (2, 3, 4, 0, 2, "f64"), (3, 3, 4, 6, 8, "f32")])
def test_demand_ordered_hermite_integrals_define_before_use(la, lb, lc, m0, m1, prec):
    """Emitted on demand, the kernel defines the same Hermite integrals with the same expressions as the
    level-by-level order, each once, every one before its first use; and the first auxiliary-transform
    term (V) of every Hermite index comes after all the integrals it reads."""
    _, lev = emit_j3c(la, lb, lc, m0, m1, prec, r_order="level")
    _, dem = emit_j3c(la, lb, lc, m0, m1, prec, r_order="demand")
    ty = "float" if prec == "f32" else "double"
    pat = re.compile(rf"\s*const {ty} (r\d+_\d+_\d+_\d+) = (.*);")
    defs = lambda code: [m.groups() for m in map(pat.match, code.splitlines()) if m]
    assert sorted(defs(lev)) == sorted(defs(dem)) and len(set(n for n, _ in defs(dem))) == len(defs(dem))
    seen = set()
    for line in dem.splitlines():
        for ref in re.findall(r"\b(r\d+_\d+_\d+_\d+)\b", line.split("=", 1)[1] if "=" in line else ""):
            assert ref in seen, f"{ref} read before it is defined: {line.strip()[:80]}"
        m = pat.match(line)
        if m:
            seen.add(m.group(1))
    # the demand order interleaves: some V is computed before the last Hermite integral
    first_v = next(i for i, l in enumerate(dem.splitlines()) if re.match(rf"\s*const {ty} V_", l))
    last_r = max(i for i, l in enumerate(dem.splitlines()) if pat.match(l))
    assert first_v < last_r or len(defs(dem)) <= 1


@pytest.mark.parametrize("emitter,ty", [("emit_j3c_bf16", "float")])
@pytest.mark.parametrize("cls", [(2, 3, 2, 0, 5), (3, 3, 4, 0, 1)])
def test_matrix_core_kernels_demand_order(emitter, ty, cls, monkeypatch):
    """The bf16 matrix-core emitter under the demand order: the same Hermite-integral definitions as the
    level order, each defined before it is read.  (The float64 matrix-core emitter keeps the level order.)"""
    emit = getattr(gen_kernels, emitter)
    monkeypatch.setattr(gen_kernels, "R_ORDER", "level")
    lev = emit(*cls)[1]
    monkeypatch.setattr(gen_kernels, "R_ORDER", "demand")
    dem = emit(*cls)[1]
    pat = re.compile(rf"\s*const {ty} (r\d+_\d+_\d+_\d+) = (.*);")
    defs = lambda code: sorted(m.groups() for m in map(pat.match, code.splitlines()) if m)
    assert defs(lev) == defs(dem) and len(defs(dem)) > 0
    seen = set()
    for line in dem.splitlines():
        rhs = line.split("=", 1)[1] if "=" in line else line
        for ref in re.findall(r"\b(r\d+_\d+_\d+_\d+)\b", rhs):
            assert ref in seen, f"{ref} read before it is defined"
        m = pat.match(line)
        if m:
            seen.add(m.group(1))
