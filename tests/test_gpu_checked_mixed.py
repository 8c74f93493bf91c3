"""Float32 tier on the bf16 matrix cores (split operands) and checked mixed precision (float32 first,
float64 for the items whose term scale reaches the threshold)."""

import numpy as np
import pytest

from tests.helpers import random_molecule
from electron_integral_playground_amd import generators, pair_utility

pytestmark = pytest.mark.gpu


def _class_index(plan, pi, ai):
    """(packed rows [n_sp * nij], output columns [n_ks * nk]) of class (pi, ai)."""
    pc, ac = plan.pairs[pi], plan.aux[ai]
    sp = np.arange(pc.n_sp)
    I = (pc.sp_i[sp][:, None, None] + np.arange(2 * pc.la + 1)[None, :, None]).astype(np.int64)
    J = (pc.sp_j[sp][:, None, None] + np.arange(2 * pc.lb + 1)[None, None, :]).astype(np.int64)
    hi, lo = np.maximum(I, J), np.minimum(I, J)
    rows = (hi * (hi + 1) // 2 + lo).ravel()
    cols = (ac.ks_start[:, None] + np.arange(2 * ac.lc + 1)[None, :]).ravel()
    return rows, cols


def test_bf16_variants_match_float64():
    """Every bf16 matrix-core variant of the float32 tier gives the float64 kernels' (ij|P) to
    float32-level accuracy (2e-6 of the class's largest value; the vector float32 kernels' error is
    printed beside it), through the packed whole-class launch and the compact block launch, and is
    rejected for the float64 tier."""
    import torch
    from electron_integral_playground_amd.engine import ThreeCenterPlan, apply_tuning
    mol = random_molecule(n_atom=4, seed=23, orbital_l=(0, 1, 2, 3), aux_l=(0, 1, 2, 3, 4), nprim=(1, 4))
    pairs = pair_utility.form_pair_data(mol, 1e-14)
    aux = pair_utility.form_auxiliary_pair_data(mol)
    plan = ThreeCenterPlan(mol, pairs, aux)
    lib = plan.lib
    classes = [(pi, ai) for pi, pc in enumerate(plan.pairs) for ai, ac in enumerate(plan.aux)
               if any(lib.eip_three_center_variant_mfma(pc.la, pc.lb, ac.lc, v) == 2
                      for v in range(lib.eip_three_center_nvariants(pc.la, pc.lb, ac.lc)))]
    assert len(classes) >= 30
    try:
        ref = plan.new_output()
        ref32 = plan.new_output()
        for pi, ai in classes:
            pc, ac = plan.pairs[pi], plan.aux[ai]
            lib.eip_set_three_center_variant(pc.la, pc.lb, ac.lc, 0, 0)
            lib.eip_set_three_center_variant(pc.la, pc.lb, ac.lc, 1, 0)
            plan.launch_class(pi, ai, ref)
            plan.launch_class(pi, ai, ref32, use_f32=True)
        torch.cuda.synchronize()
        checked = 0
        for pi, ai in classes:
            pc, ac = plan.pairs[pi], plan.aux[ai]
            nij = (2 * pc.la + 1) * (2 * pc.lb + 1)
            rows, cols = _class_index(plan, pi, ai)
            rows_t, cols_t = torch.as_tensor(rows, device="cuda"), torch.as_tensor(cols, device="cuda")
            r = ref[rows_t].index_select(1, cols_t)
            scale = r.abs().max().item()
            err32 = (ref32[rows_t].index_select(1, cols_t) - r).abs().max().item()
            for v in range(lib.eip_three_center_nvariants(pc.la, pc.lb, ac.lc)):
                if lib.eip_three_center_variant_mfma(pc.la, pc.lb, ac.lc, v) != 2:
                    continue
                print(f"case class ({pc.la}{pc.lb}|{ac.lc}) bf16 variant {v}", flush=True)
                assert lib.eip_set_three_center_variant(pc.la, pc.lb, ac.lc, 1, v) == 0
                assert lib.eip_set_three_center_variant(pc.la, pc.lb, ac.lc, 0, v) == 1   # float32 tier only
                assert lib.eip_three_center_f32_uses_bf16(pc.la, pc.lb, ac.lc) == 1
                got = torch.full_like(ref, float("nan"))
                plan.launch_class(pi, ai, got, use_f32=True)
                blk = torch.full((pc.n_sp * nij, mol.n_aux), float("nan"), dtype=torch.float64, device="cuda")
                # the compact block launch of the float32 tier: every item through the job table
                base = blk.data_ptr()
                jobs = plan.job_table([plan.job_words(pi, ai, base, plan.n_aux)])
                plan.launch_work(pi, ai, plan.items[(pi, ai)], jobs, layout=1, use_f32=True)
                torch.cuda.synchronize()
                e_pk = (got[rows_t].index_select(1, cols_t) - r).abs().max().item()
                e_blk = (blk.index_select(1, cols_t) - r).abs().max().item()
                print(f"  max |bf16 - f64| packed {e_pk:.3e} block {e_blk:.3e}; vector f32 {err32:.3e}; "
                      f"max |x| {scale:.3e}", flush=True)
                assert e_pk <= 2e-6 * scale + 1e-14 and e_blk <= 2e-6 * scale + 1e-14, (e_pk, e_blk, scale)
                checked += 1
            lib.eip_set_three_center_variant(pc.la, pc.lb, ac.lc, 1, 0)
        assert checked >= len(classes)
    finally:
        apply_tuning()


def _item_max(plan, ref, pi, ai):
    """max |(ij|P)| of every (shell pair, block of 64 auxiliary shells) item of class (pi, ai), [n_sp, G]."""
    import torch
    pc, ac = plan.pairs[pi], plan.aux[ai]
    rows, cols = _class_index(plan, pi, ai)
    nij, nk = (2 * pc.la + 1) * (2 * pc.lb + 1), 2 * ac.lc + 1
    a = ref[torch.as_tensor(rows, device="cuda")].index_select(1, torch.as_tensor(cols, device="cuda")).abs()
    a = a.reshape(pc.n_sp, nij, ac.n_ks, nk).amax(dim=(1, 3))
    G = (ac.n_ks + 63) // 64
    a = torch.nn.functional.pad(a, (0, G * 64 - ac.n_ks))
    return a.reshape(pc.n_sp, G, 64).amax(dim=2)


@pytest.mark.parametrize("tier", ["tuned", "bf16"])
def test_checked_mixed_precision(tier):
    """Checked mixed precision: an item the float32 pass leaves unflagged has all its values below the
    threshold (their term scale is, and it bounds them), the float64 pass replaces exactly the flagged
    items, and the packed result stays within 1e-10 of float64 with a share of the items in float32 --
    with the tuned float32 variants and with the bf16 matrix-core ones on every class that has them."""
    import torch
    from electron_integral_playground_amd.engine import ThreeCenterPlan, apply_tuning
    mol = generators.make_molecule("mixed", n_atom=12, box=9.0)
    arrays = pair_utility.form_pair_arrays(mol, 1e-14)
    pairs = pair_utility.form_primitive_pair_data(mol, arrays)
    aux = pair_utility.form_auxiliary_pair_data(mol)
    plan = ThreeCenterPlan(mol, pairs, aux)
    lib = plan.lib
    try:
        if tier == "bf16":
            for pc in plan.pairs:
                for ac in plan.aux:
                    vs = [v for v in range(lib.eip_three_center_nvariants(pc.la, pc.lb, ac.lc))
                          if lib.eip_three_center_variant_mfma(pc.la, pc.lb, ac.lc, v) == 2]
                    if vs:
                        lib.eip_set_three_center_variant(pc.la, pc.lb, ac.lc, 1, vs[-1])
        ref = plan.run()
        for thr in (1e-6, 1e-4):
            plan.setup_mixed(arrays, None, threshold=thr, granularity="checked")
            got = torch.full_like(ref, float("nan"))
            plan.zero_unwritten(got)
            keys = [(pi, ai) for pi in range(len(plan.pairs)) for ai in range(len(plan.aux))]
            jobs = plan.job_table([plan.job_words(pi, ai, got.data_ptr(), plan.n_aux) for pi, ai in keys])
            n_items = n_f32 = 0
            for j, (pi, ai) in enumerate(keys):
                pc, ac = plan.pairs[pi], plan.aux[ai]
                items = plan.class_items(pi, ai)
                plan.launch_checked(pi, ai, items, jobs[j], layout=0)
                n = int(items.shape[0])
                n_items += n
                if lib.eip_three_center_f32_uses_bf16(pc.la, pc.lb, ac.lc) != 1:
                    continue
                stride = lib.eip_three_center_flag_stride(pc.la, pc.lb, ac.lc)
                flags = plan.flag_buffer(n * stride)[:n * stride].view(n, stride).ne(0).any(1)
                imax = _item_max(plan, ref, pi, ai)
                it = items.long()
                vmax = imax[it[:, 0], it[:, 1] // 64]
                bad = (~flags) & (vmax >= 1.01 * thr)
                assert int(bad.sum().item()) == 0, (pc.la, pc.lb, ac.lc, int(bad.sum().item()))
                n_f32 += int((~flags).sum().item())
            torch.cuda.synchronize()
            err = (got - ref).abs().max().item()
            print(f"case tier {tier} threshold {thr:g}: {n_f32}/{n_items} items kept in float32; "
                  f"max |mixed - f64| {err:.3e}", flush=True)
            assert err <= 1e-10
            any_bf16 = any(lib.eip_three_center_f32_uses_bf16(pc.la, pc.lb, ac.lc) == 1
                           for pc in plan.pairs for ac in plan.aux)
            if thr == 1e-4 and any_bf16:
                assert 0 < n_f32 < n_items
        # the plan's own checked paths: whole output and compact blocks
        got2 = plan.run_mixed()
        torch.cuda.synchronize()
        assert (got2 - ref).abs().max().item() <= 1e-10
        blocks = plan.row_blocks(4e9)
        A = torch.empty((max(plan.block_rows(*b) for b in blocks), mol.n_aux), dtype=torch.float64, device="cuda")
        B = torch.empty_like(A)
        for b in blocks:
            n = plan.block_rows(*b)
            plan.run_block(*b, A)
            plan.run_block(*b, B, mixed=True)
            torch.cuda.synchronize()
            assert (B[:n] - A[:n]).abs().max().item() <= 1e-10, b
    finally:
        apply_tuning()


@pytest.mark.parametrize("omega", [0.0, 0.3])
def test_envelope_scale_bounds_every_item(omega):
    """The a-priori term scale of every work item bounds the item's largest |(ij|P)| (it bounds the
    magnitude of every term summed into the outputs), on every class of a 12-atom molecule, and it is
    distance-aware: far from 1 for the item spread of a 9 A box."""
    import torch
    from electron_integral_playground_amd.engine import ThreeCenterPlan
    mol = generators.make_molecule("mixed", n_atom=12, box=9.0)
    arrays = pair_utility.form_pair_arrays(mol, 1e-14)
    pairs = pair_utility.form_primitive_pair_data(mol, arrays)
    aux = pair_utility.form_auxiliary_pair_data(mol)
    plan = ThreeCenterPlan(mol, pairs, aux, omega=omega)
    ref = plan.run()
    torch.cuda.synchronize()
    worst, spread = 0.0, []
    for pi, pc in enumerate(plan.pairs):
        for ai, ac in enumerate(plan.aux):
            print(f"case omega {omega} class ({pc.la}{pc.lb}|{ac.lc})", flush=True)
            items = plan.items[(pi, ai)]
            scale = plan.item_scales(pi, ai).double()
            imax = _item_max(plan, ref, pi, ai).reshape(-1)       # items are (sp, block) in sp-major order
            assert imax.shape == scale.shape
            # (the scale is float32: items below its range, |value| < 1e-30, are not compared)
            big = imax > 1e-30
            ratio = (imax[big] / scale[big]).max().item() if bool(big.any()) else 0.0
            worst = max(worst, ratio)
            assert ratio <= 1.0 + 1e-5, (pc.la, pc.lb, ac.lc, ratio)
            spread.append((scale.max() / scale.clamp(min=1e-30).min()).item())
    print(f"largest value / scale over all items {worst:.3f}; scale spread per class up to {max(spread):.3e}")
    assert max(spread) > 1e6


def test_envelope_split_within_bound():
    """The mixed split on the envelope estimate stays within 1e-10 of float64 (packed and compact-block
    paths) with a float32 share above the Schwarz-type split's, and threshold 0 reproduces float64 bit for
    bit."""
    import torch
    from electron_integral_playground_amd.engine import ThreeCenterPlan
    mol = generators.make_molecule("mixed", n_atom=12, box=9.0)
    arrays = pair_utility.form_pair_arrays(mol, 1e-14)
    pairs = pair_utility.form_primitive_pair_data(mol, arrays)
    aux = pair_utility.form_auxiliary_pair_data(mol)
    plan = ThreeCenterPlan(mol, pairs, aux)
    ref = plan.run()
    blocks = plan.row_blocks(4e9)
    A = torch.empty((max(plan.block_rows(*b) for b in blocks), mol.n_aux), dtype=torch.float64, device="cuda")
    B = torch.empty_like(A)
    for thr in (1e-6, 1e-5, 0.0):
        stats = plan.setup_mixed(arrays, None, threshold=thr, estimate="envelope")
        frac = stats["f32_triplets"] / (stats["f32_triplets"] + stats["f64_triplets"])
        got = plan.run_mixed()
        torch.cuda.synchronize()
        err = (got - ref).abs().max().item()
        print(f"case envelope threshold {thr:g}: float32 share {frac:.3f} of triplets; max |mixed - f64| {err:.3e}; "
              f"estimate {stats['estimate_seconds']:.3f} s", flush=True)
        if thr == 0.0:
            assert stats["f32_triplets"] == 0 and torch.equal(got, ref)
        else:
            assert err <= 1e-10 and frac > 0.05
        for b in blocks:
            n = plan.block_rows(*b)
            plan.run_block(*b, A)
            plan.run_block(*b, B, mixed=True)
            torch.cuda.synchronize()
            assert (B[:n] - A[:n]).abs().max().item() <= (1e-10 if thr else 0.0), (thr, b)
