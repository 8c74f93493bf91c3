"""Kernel generator checks that need no GPU: the register layout of pair-matrix rows in the DPP staging
mode, the two-stage two-centre contraction, and the per-class staging choice."""

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
