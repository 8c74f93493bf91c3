"""Device-resident path (ThreeCenterPlan, packed output) against the drop-in ABI and the oracle."""

import numpy as np
import pytest

from tests.helpers import ATOL, RTOL, oracle_lib, random_molecule, run_three_center_oracle
from electron_integral_playground_amd import generators, pair_utility

pytestmark = pytest.mark.gpu


def _packed_from_dense(J):
    n_ao = J.shape[1]
    i, j = np.tril_indices(n_ao)
    return J[:, i, j].T


@pytest.mark.parametrize("workload", ["small", "random"])
def test_plan_matches_oracle(workload):
    import torch
    from electron_integral_playground_amd.engine import ThreeCenterPlan
    print(f"case workload={workload}", flush=True)
    if workload == "small":
        mol = generators.make_molecule("small")
    else:
        mol = random_molecule(n_atom=4, seed=21, orbital_l=(0, 1, 2, 3), aux_l=(0, 1, 2, 3, 4))
    pairs = pair_utility.form_pair_data(mol, 1e-14)
    aux = pair_utility.form_auxiliary_pair_data(mol)
    plan = ThreeCenterPlan(mol, pairs, aux)
    out = plan.run()
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    ref = _packed_from_dense(run_three_center_oracle(oracle_lib(), pairs, aux, mol))
    np.testing.assert_allclose(got, ref, rtol=RTOL, atol=ATOL)
    # a second run into the same buffer gives bit-identical results (no atomics, fixed order)
    out2 = plan.run(out=torch.full_like(out, float("nan")))
    torch.cuda.synchronize()
    assert torch.equal(out, out2)
    # launches spread over several streams (the default) equal the single-stream order bit for bit
    out3 = plan.run(out=torch.full_like(out, float("nan")), concurrent=False)
    torch.cuda.synchronize()
    assert torch.equal(out, out3)
    # the metric computed in the same call (on a side stream, or in order) equals two_center()
    ref2 = plan.two_center()
    for concurrent in (True, False):
        J = torch.full_like(ref2, float("nan"))
        out4 = plan.run(out=torch.full_like(out, float("nan")), concurrent=concurrent, metric_out=J)
        torch.cuda.synchronize()
        assert torch.equal(out, out4) and torch.equal(J, ref2), concurrent


def _q_aux_from_metric(plan, J2c):
    """sqrt(max diagonal (P|P)) per auxiliary shell, per class, in the plan's shell order."""
    diag = J2c.diagonal().cpu().numpy()
    out = []
    for ac in plan.aux:
        n = 2 * ac.lc + 1
        idx = ac.ks_start[:, None] + np.arange(n)[None, :]
        out.append(np.sqrt(np.abs(diag[idx]).max(axis=1)))
    return out


def test_two_center_device_matches_dropin():
    import torch
    from electron_integral_playground_amd.engine import ThreeCenterPlan
    from electron_integral_playground_amd.molecular_integral import two_center
    mol = random_molecule(n_atom=4, seed=22, orbital_l=(0, 1), aux_l=(0, 1, 2, 3, 4))
    pairs = pair_utility.form_pair_data(mol, 1e-14)
    aux = pair_utility.form_auxiliary_pair_data(mol)
    plan = ThreeCenterPlan(mol, pairs, aux)
    J = plan.two_center()
    torch.cuda.synchronize()
    np.testing.assert_allclose(J.cpu().numpy(), two_center(aux, mol), rtol=RTOL, atol=ATOL)


def test_mixed_precision_error_bound():
    """Float32 triplets (bound < 1e-6) keep the packed result within 1e-10 of the float64 path."""
    import torch
    from electron_integral_playground_amd.engine import ThreeCenterPlan
    mol = generators.make_molecule("mixed", n_atom=12, box=9.0)
    pl = pair_utility.form_primitive_pair_list(mol, 1e-14)
    pairs = pair_utility.form_primitive_pair_data(mol, pl)
    aux = pair_utility.form_auxiliary_pair_data(mol)
    plan = ThreeCenterPlan(mol, pairs, aux)
    ref = plan.run()
    stats = plan.setup_mixed(pl, _q_aux_from_metric(plan, plan.two_center()), threshold=1e-6, granularity="pair")
    print("mixed split", stats, flush=True)
    assert stats["f32_triplets"] > 0 and stats["f64_triplets"] > 0
    got = plan.run_mixed()
    torch.cuda.synchronize()
    err = (got - ref).abs().max().item()
    print("max abs error mixed vs float64:", err)
    assert err <= 1e-10


def test_batch_plan_matches_single_molecule_plans():
    """A batch of uneven molecules in one launch per class, replayed from a HIP graph, equals the
    per-molecule plans bit for bit."""
    import torch
    from electron_integral_playground_amd.engine import BatchPlan, ThreeCenterPlan
    mols = []
    for i, n in enumerate([4, 9, 6]):
        mol = generators.batched_molecule(i, n)
        mols.append((mol, pair_utility.form_pair_data(mol, 1e-14), pair_utility.form_auxiliary_pair_data(mol)))
    batch = BatchPlan(mols)
    out = batch.new_output()
    g = torch.cuda.CUDAGraph()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        batch.run(out)                        # warm-up outside capture
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    with torch.cuda.graph(g):
        batch.run(out)
    singles = [ThreeCenterPlan(mol, pd, ad) for mol, pd, ad in mols]
    ref3 = [p.run() for p in singles]
    ref2 = [p.two_center() for p in singles]
    torch.cuda.synchronize()
    for replay in range(3):
        out.fill_(float("nan"))
        g.replay()
        torch.cuda.synchronize()
        for m in range(len(mols)):
            assert torch.equal(batch.molecule_output(out, m), ref3[m]), (replay, m)
        # eager work between replays, as a caller would have
        torch.cuda.synchronize()
    out2 = batch.two_center(batch.new_two_center_output())
    torch.cuda.synchronize()
    for m in range(len(mols)):
        assert torch.equal(batch.molecule_two_center(out2, m), ref2[m]), m


def _block_plan(n_atom=10, box=8.0):
    from electron_integral_playground_amd.engine import ThreeCenterPlan
    mol = generators.make_molecule("mixed", n_atom=n_atom, box=box)
    arrays = pair_utility.form_pair_arrays(mol, 1e-14)
    pairs = pair_utility.form_primitive_pair_data(mol, arrays)
    aux = pair_utility.form_auxiliary_pair_data(mol)
    return mol, arrays, ThreeCenterPlan(mol, pairs, aux)


def test_row_blocks_match_packed_output():
    """Streaming by shell-pair blocks (compact rows, job-table launches) gives the packed path's values
    bit for bit (up to the i <-> j symmetry inside diagonal shell pairs), for blocks small enough that
    every class is split."""
    import torch
    mol, _, plan = _block_plan()
    ref = plan.run()
    blocks = plan.row_blocks(3 * 25 * mol.n_aux * 8)
    assert len(blocks) > len(plan.pairs)
    buf = torch.full((max(plan.block_rows(*b) for b in blocks), mol.n_aux), float("nan"), dtype=torch.float64,
                     device="cuda")
    for pi, sp0, sp1 in blocks:
        print(f"case block pi={pi} sp=[{sp0},{sp1})", flush=True)
        pc = plan.pairs[pi]
        plan.run_block(pi, sp0, sp1, buf)
        torch.cuda.synchronize()
        ni, nj = 2 * pc.la + 1, 2 * pc.lb + 1
        for sp in range(sp0, sp1):
            I, J = int(pc.sp_i[sp]), int(pc.sp_j[sp])
            for a in range(ni):
                for b in range(nj):
                    hi, lo = max(I + a, J + b), min(I + a, J + b)
                    row = (sp - sp0) * ni * nj + a * nj + b
                    want = ref[hi * (hi + 1) // 2 + lo]
                    if I == J and a < b:
                        # the packed path keeps the (b, a) element of a diagonal shell pair; (a, b) is
                        # the same integral through the transposed pair matrix entry
                        assert torch.allclose(buf[row], want, rtol=RTOL, atol=ATOL), (pi, sp, a, b)
                    else:
                        assert torch.equal(buf[row], want), (pi, sp, a, b)


def test_row_blocks_into_padded_rows():
    """A block buffer whose rows are longer than n_aux (a column slice of a wider tensor): run_block
    writes the same values at the tensor's row stride and leaves the padding alone; a buffer without
    contiguous columns is rejected."""
    import torch
    mol, _, plan = _block_plan()
    blocks = plan.row_blocks(3 * 25 * mol.n_aux * 8)
    rows = max(plan.block_rows(*b) for b in blocks)
    ref = torch.empty((rows, mol.n_aux), dtype=torch.float64, device="cuda")
    wide = torch.full((rows, mol.n_aux + 40), -7.0, dtype=torch.float64, device="cuda")
    padded = wide[:, :mol.n_aux]
    for b in blocks[::3]:
        print(f"case padded block {b}", flush=True)
        n = plan.block_rows(*b)
        plan.run_block(*b, ref)
        plan.run_block(*b, padded)
        torch.cuda.synchronize()
        assert torch.equal(padded[:n], ref[:n]), b
        assert bool((wide[:, mol.n_aux:] == -7.0).all())
    with pytest.raises(ValueError):
        plan.run_block(*blocks[0], wide.t())


def test_row_blocks_mixed_within_bound():
    """The mixed split (float64 prefix stored, float32 suffix added) per block stays within 1e-10 of
    float64, and a threshold of 0 (everything float64) reproduces float64 exactly."""
    import torch
    mol, arrays, plan = _block_plan()
    q = _q_aux_from_metric(plan, plan.two_center())
    blocks = plan.row_blocks(4e9)
    rows = max(plan.block_rows(*b) for b in blocks)
    A = torch.empty((rows, mol.n_aux), dtype=torch.float64, device="cuda")
    B = torch.empty_like(A)
    for thr, bound, gran, order in ((1e-6, 1e-10, "pair", "qp"), (1e-6, 1e-10, "item", "qp"),
                                    (1e-6, 1e-10, "item", "natural"), (1e-6, 1e-10, "pair", "natural"),
                                    (0.0, 0.0, "pair", "qp"), (0.0, 0.0, "item", "natural")):
        stats = plan.setup_mixed(arrays, q, threshold=thr, granularity=gran, aux_order=order)
        print(f"case threshold={thr} granularity={gran} aux_order={order} split={stats}", flush=True)
        if thr > 0 and gran == "pair":
            assert stats["f32_triplets"] > 0
        if thr == 0:
            assert stats["f32_triplets"] == 0
        for b in blocks:
            n = plan.block_rows(*b)
            plan.run_block(*b, A)
            plan.run_block(*b, B, mixed=True)
            torch.cuda.synchronize()
            err = (B[:n] - A[:n]).abs().max().item()
            assert err <= bound, (thr, b, err)


def test_metric_cholesky_on_device():
    """L L^T reproduces the float64 metric, and L^-1 (P|ij) whitens a block of the three-centre tensor."""
    import torch
    from electron_integral_playground_amd.engine import ThreeCenterPlan, metric_cholesky
    mol = generators.make_molecule("small")
    plan = ThreeCenterPlan(mol, pair_utility.form_pair_data(mol, 1e-14), pair_utility.form_auxiliary_pair_data(mol))
    J2c = plan.two_center()
    L = metric_cholesky(J2c)
    torch.cuda.synchronize()
    assert torch.allclose(L @ L.T, J2c, rtol=1e-12, atol=1e-10 * J2c.abs().max().item())
    J3c = plan.run()
    B = torch.linalg.solve_triangular(L, J3c.T.contiguous(), upper=False)
    torch.cuda.synchronize()
    # (ij|kl) ~ B^T B: the fitted Coulomb matrix is symmetric positive semi-definite
    V = B.T @ B
    assert torch.allclose(V, V.T, atol=1e-10)


@pytest.mark.parametrize("max_blocks", [0, 7, 256])
def test_zeroing_writes_exactly_the_screened_rows(max_blocks):
    """The zeroing kernels (full speed, and the paced background form on a capped grid) zero every
    screened-out packed row and touch nothing else."""
    import torch
    mol, _, plan = _block_plan()
    print(f"case max_blocks={max_blocks} zero rows={plan.zero_rows.numel()}", flush=True)
    assert plan.zero_rows.numel() > 0
    out = plan.new_output().fill_(1.0)
    plan.zero_unwritten(out, None, max_blocks)
    torch.cuda.synchronize()
    mask = torch.zeros(out.shape[0], dtype=torch.bool, device=out.device)
    mask[plan.zero_rows.long()] = True
    assert bool((out[mask] == 0).all().item())
    assert bool((out[~mask] == 1.0).all().item())



def test_matrix_core_variants_match_vector_kernels():
    """Every float64 matrix-core variant (v_mfma_f64_16x16x4f64 kernels over the dense pair matrix) gives
    the vector kernels' (ij|P) to 1e-12 + 1e-10 |x| on every class that has one, through the packed
    whole-class launch and the compact block launch, and the drop-in entry point matches the oracle."""
    import torch
    from electron_integral_playground_amd.engine import ThreeCenterPlan, apply_tuning
    from electron_integral_playground_amd.molecular_integral import three_center
    mol = random_molecule(n_atom=4, seed=23, orbital_l=(0, 1, 2, 3), aux_l=(0, 1, 2, 3, 4), nprim=(1, 4))
    pairs = pair_utility.form_pair_data(mol, 1e-14)
    aux = pair_utility.form_auxiliary_pair_data(mol)
    plan = ThreeCenterPlan(mol, pairs, aux)
    lib = plan.lib
    classes = [(pi, ai) for pi, pc in enumerate(plan.pairs) for ai, ac in enumerate(plan.aux)
               if any(lib.eip_three_center_variant_mfma(pc.la, pc.lb, ac.lc, v) == 1
                      for v in range(lib.eip_three_center_nvariants(pc.la, pc.lb, ac.lc)))]
    assert len(classes) >= 20
    try:
        ref = plan.new_output()
        for pi, ai in classes:
            pc, ac = plan.pairs[pi], plan.aux[ai]
            lib.eip_set_three_center_variant(pc.la, pc.lb, ac.lc, 0, 0)
            plan.launch_class(pi, ai, ref)
        torch.cuda.synchronize()
        checked = 0
        for pi, ai in classes:
            pc, ac = plan.pairs[pi], plan.aux[ai]
            nij = (2 * pc.la + 1) * (2 * pc.lb + 1)
            for v in range(lib.eip_three_center_nvariants(pc.la, pc.lb, ac.lc)):
                if lib.eip_three_center_variant_mfma(pc.la, pc.lb, ac.lc, v) != 1:
                    continue
                print(f"case class ({pc.la}{pc.lb}|{ac.lc}) matrix-core variant {v}", flush=True)
                assert lib.eip_set_three_center_variant(pc.la, pc.lb, ac.lc, 0, v) == 0
                assert lib.eip_set_three_center_variant(pc.la, pc.lb, ac.lc, 1, v) == 1   # float64 only
                assert lib.eip_three_center_uses_mfma(pc.la, pc.lb, ac.lc) == 1
                got = torch.full_like(ref, float("nan"))
                plan.launch_class(pi, ai, got)
                blk = torch.full((pc.n_sp * nij, mol.n_aux), float("nan"), dtype=torch.float64, device="cuda")
                plan.run_block(pi, 0, pc.n_sp, blk, aux_classes=[ai])
                torch.cuda.synchronize()
                cols = torch.as_tensor((ac.ks_start[:, None] + np.arange(2 * ac.lc + 1)[None, :]).ravel(), device="cuda")
                sp = np.arange(pc.n_sp)
                I = (pc.sp_i[sp][:, None, None] + np.arange(2 * pc.la + 1)[None, :, None]).astype(np.int64)
                J = (pc.sp_j[sp][:, None, None] + np.arange(2 * pc.lb + 1)[None, None, :]).astype(np.int64)
                hi, lo = np.maximum(I, J), np.minimum(I, J)
                packed = torch.as_tensor((hi * (hi + 1) // 2 + lo).ravel(), device="cuda")
                r = ref[packed].index_select(1, cols).cpu().numpy()
                np.testing.assert_allclose(got[packed].index_select(1, cols).cpu().numpy(), r, rtol=RTOL, atol=ATOL)
                np.testing.assert_allclose(blk.index_select(1, cols).cpu().numpy(), r, rtol=RTOL, atol=ATOL)
                checked += 1
            lib.eip_set_three_center_variant(pc.la, pc.lb, ac.lc, 0, 0)
        assert checked >= len(classes)
        # the drop-in ABI with every matrix-core class on its first matrix-core variant
        for pi, ai in classes:
            pc, ac = plan.pairs[pi], plan.aux[ai]
            v = next(v for v in range(lib.eip_three_center_nvariants(pc.la, pc.lb, ac.lc))
                     if lib.eip_three_center_variant_mfma(pc.la, pc.lb, ac.lc, v) == 1)
            lib.eip_set_three_center_variant(pc.la, pc.lb, ac.lc, 0, v)
        np.testing.assert_allclose(three_center(pairs, aux, mol), run_three_center_oracle(oracle_lib(), pairs, aux, mol),
                                   rtol=RTOL, atol=ATOL)
    finally:
        apply_tuning()
