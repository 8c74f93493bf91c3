"""Density-fitted Coulomb matrix consumed on the device (engine.CoulombFit).

J_ij = sum_P (ij|P) d_P with (P|Q) d = gamma, gamma_P = sum_ij (P|ij) D_ij, for a random symmetric
density D:
* on the 12-atom workload from the packed output, against the same contraction of the oracle's dense
  (ij|P) and (P|Q) in numpy (float64; the bar is relative to max |J|, since the metric solve amplifies the
  integrals' own rounding by the metric's condition number);
* the streamed form (two passes over compact blocks of shell pairs in one small reused buffer) equal to
  the packed form on a 30-atom molecule of the mixed-precision generators, plain and mixed precision.
"""

import numpy as np
import pytest

from tests.helpers import oracle_lib, run_three_center_oracle, run_two_center_oracle
from electron_integral_playground_amd import generators, pair_utility

pytestmark = pytest.mark.gpu


def _density(n, seed):
    rng = np.random.default_rng(seed)
    c = rng.standard_normal((n, max(1, n // 4)))
    return c @ c.T / n


def _plan(mol):
    from electron_integral_playground_amd.engine import ThreeCenterPlan
    pairs = pair_utility.form_pair_data(mol, 1e-14)
    aux = pair_utility.form_auxiliary_pair_data(mol)
    return ThreeCenterPlan(mol, pairs, aux), pairs, aux


def test_coulomb_fit_vs_oracle_small():
    import torch
    from electron_integral_playground_amd.engine import CoulombFit
    mol = generators.make_molecule("small")
    plan, pairs, aux = _plan(mol)
    D = _density(mol.n_ao, 5)
    Dt = torch.as_tensor(D, device="cuda")
    packed = plan.run()
    fit = CoulombFit(plan)
    lib = oracle_lib()
    T = run_three_center_oracle(lib, pairs, aux, mol)          # (n_aux, n_ao, n_ao)
    V = run_two_center_oracle(lib, aux, mol)
    # the two contractions, each against the oracle's integrals with the same operand
    g = fit.gamma_packed(packed, Dt).cpu().numpy()
    g_ref = np.einsum("pij,ij->p", T, D)
    g_scale = np.einsum("pij,ij->p", np.abs(T), np.abs(D))
    assert np.all(np.abs(g - g_ref) <= 1e-12 + 1e-12 * g_scale), float(np.max(np.abs(g - g_ref)))
    d_ref = np.linalg.solve(V, g_ref)
    Jd = fit.contract_packed(packed, torch.as_tensor(d_ref, device="cuda")).cpu().numpy()
    ref = np.einsum("pij,p->ij", T, d_ref)
    scale = np.einsum("pij,p->ij", np.abs(T), np.abs(d_ref))
    assert np.all(np.abs(Jd - ref) <= 1e-12 + 1e-12 * scale), float(np.max(np.abs(Jd - ref)))
    # the whole fit, solve included: the metric's condition number amplifies the integrals' rounding
    J = fit.from_packed(packed, Dt).cpu().numpy()
    cond = float(np.linalg.cond(V))
    err = float(np.max(np.abs(J - ref)))
    print(f"12-atom RI-J: max |J - oracle| = {err:.2e} (max |J| {np.max(np.abs(ref)):.3e}, cond(P|Q) {cond:.1e})",
          flush=True)
    assert err <= 1e-15 * cond * float(np.max(scale)) + 1e-12
    assert np.array_equal(J, J.T)


@pytest.mark.parametrize("mixed", [False, True])
def test_coulomb_fit_streamed_equals_packed(mixed):
    import torch
    from electron_integral_playground_amd.engine import CoulombFit
    mol = generators.make_molecule("mixed", n_atom=30, seed_offset=77, box=20.0 * (30 / 200) ** (1 / 3))
    plan, pairs, aux = _plan(mol)
    D = torch.as_tensor(_density(mol.n_ao, 6), device="cuda")
    fit = CoulombFit(plan)
    packed = plan.run()
    g_ref = fit.gamma_packed(packed, D)
    d = fit.solve(g_ref)
    J_ref = fit.contract_packed(packed, d)
    if mixed:
        stats = plan.setup_envelope(1e-6)
        assert stats["f32_triplets"] > 0
    block = 8 * plan.n_aux * 600                                # a few hundred rows per block
    n_blocks = len(plan.row_blocks(block))
    assert n_blocks > 10
    g = fit.gamma_streamed(D, block, mixed=mixed)
    J = fit.contract_streamed(d, block, mixed=mixed)
    eg = float((g - g_ref).abs().max() / g_ref.abs().max())
    eJ = float((J - J_ref).abs().max() / J_ref.abs().max())
    print(f"30-atom RI-J streamed ({n_blocks} blocks, mixed={mixed}) vs packed: gamma {eg:.2e}, J {eJ:.2e} (relative)",
          flush=True)
    bar = 1e-10 if mixed else 1e-13
    assert eg <= bar and eJ <= bar


@pytest.mark.parametrize("n_rows,n_cols,ld", [(100003, 1537, 1600), (64, 1024, 1024), (1, 7, 9), (0, 300, 300)])
def test_consumer_kernels_match_torch(n_rows, n_cols, ld):
    """eip_weighted_column_sums / eip_row_dots against float64 torch products, on ragged shapes (columns
    not a multiple of the tile, padded rows, a single row, no rows), both column-sum modes."""
    import ctypes
    import torch
    from electron_integral_playground_amd import library
    lib = library()
    g = torch.Generator(device="cuda").manual_seed(n_rows + n_cols)
    full = torch.randn((max(n_rows, 1), ld), dtype=torch.float64, device="cuda", generator=g)
    A = full[:n_rows, :n_cols]
    w = torch.randn(n_rows, dtype=torch.float64, device="cuda", generator=g)
    x = torch.randn(n_cols, dtype=torch.float64, device="cuda", generator=g)
    V = ctypes.c_void_p
    s = V(torch.cuda.current_stream().cuda_stream)
    need = lib.eip_weighted_column_sums_workspace(ctypes.c_longlong(n_rows), n_cols)
    work = torch.empty(max(need, 1), dtype=torch.float64, device="cuda")
    gamma = torch.full((n_cols,), 3.0, dtype=torch.float64, device="cuda")
    assert lib.eip_weighted_column_sums(V(A.data_ptr()), ctypes.c_longlong(n_rows), n_cols, ctypes.c_longlong(ld),
                                        V(w.data_ptr()), V(work.data_ptr()), ctypes.c_longlong(work.numel()), 0,
                                        V(gamma.data_ptr()), s) == 0
    ref = A.t() @ w if n_rows else torch.zeros(n_cols, dtype=torch.float64, device="cuda")
    scale = (A.abs().t() @ w.abs()) if n_rows else torch.zeros_like(ref)
    assert bool(((gamma - ref).abs() <= 1e-13 * scale + 1e-300).all())
    acc = gamma.clone()
    assert lib.eip_weighted_column_sums(V(A.data_ptr()), ctypes.c_longlong(n_rows), n_cols, ctypes.c_longlong(ld),
                                        V(w.data_ptr()), V(work.data_ptr()), ctypes.c_longlong(work.numel()), 1,
                                        V(acc.data_ptr()), s) == 0
    assert bool(((acc - 2 * ref).abs() <= 2e-13 * scale + 1e-300).all())
    # a workspace too small is rejected, not overrun
    if n_rows:
        assert lib.eip_weighted_column_sums(V(A.data_ptr()), ctypes.c_longlong(n_rows), n_cols, ctypes.c_longlong(ld),
                                            V(w.data_ptr()), V(work.data_ptr()), ctypes.c_longlong(need - 1), 0,
                                            V(gamma.data_ptr()), s) == 4
    out = torch.full((max(n_rows, 1),), 7.0, dtype=torch.float64, device="cuda")
    assert lib.eip_row_dots(V(A.data_ptr()), ctypes.c_longlong(n_rows), n_cols, ctypes.c_longlong(ld), V(x.data_ptr()),
                            V(out.data_ptr()), s) == 0
    if n_rows:
        refr = A @ x
        assert bool(((out[:n_rows] - refr).abs() <= 1e-13 * (A.abs() @ x.abs())).all())
    torch.cuda.synchronize()
