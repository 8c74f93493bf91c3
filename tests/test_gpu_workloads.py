"""The five benchmark workloads at their full sizes on the GPU, checked against the oracle.

Each workload runs through the same path its benchmark leg uses; the oracle (the reference's
per-primitive-triplet algorithm, reference three_center.cpp:291-317 for the three-centre loop) is
evaluated on the CPU for every orbital pair against a seeded sample of auxiliary shells, since the full
outputs (51 GB to 3.9 TB) exceed what a CPU check can produce:

1. small (12 atoms): S, V at the nuclei, (P|Q) and (ij|P) in full through the drop-in entry points;
2. plain (80 atoms, n_ao 1120, n_aux 10240): the device-resident packed (ij|P) of every pair against 16
   auxiliary shells per class, the full (P|Q), and S and V through the drop-in entry points;
3. batched (first wave of the 256-molecule workload in one HIP-graph-captured batch plan): the
   smallest molecule in full, the largest on 8 shells per class, every other molecule on one shell per
   class, and (P|Q) of the smallest and largest;
4. mixed (200 atoms, n_ao 6000, n_aux 40000): the float64 block stream and the mixed-precision stream on
   16 auxiliary shells (part of them in the float32 tier) against all 7.84 million pairs;
5. serving: 10-, 40- and 130-atom jobs served concurrently, each checked on sampled shells, and the
   sampled outputs equal when each job is served alone.

Tolerances: float64 paths |HIP - oracle| <= 1e-12 + 1e-10 |oracle|; the mixed-precision path max abs
error <= 1e-10 against the oracle (its stated bound).  The reference's Cartesian-then-spherical route
cancels heavily for a few shell pairs of very unbalanced exponents (e.g. 0.087 with 4790 at 5 Bohr in a
(df) pair: the float64 oracle and the compiled reference agree with each other and both differ from the
long-double oracle by 1e-10); an element outside its bar is therefore decided by the long-double oracle,
evaluated for that shell pair alone, within the same bar plus 10x the float64 oracle's own error there.
"""

import numpy as np
import pytest

from tests.helpers import (ATOL, RTOL, assert_parity, oracle_aux_shell_block, oracle_ld_lib, oracle_lib,
                           oracle_shell_pair_block, run_nuclear_oracle, run_overlap_oracle, run_three_center_oracle,
                           run_two_center_oracle, sample_aux_shells)
from electron_integral_playground_amd import generators, pair_utility

pytestmark = pytest.mark.gpu

MIXED_BOUND = 1e-10


def _say(*a):
    print(*a, flush=True)


def _free():
    import gc
    import torch
    gc.collect()
    torch.cuda.synchronize()
    torch.cuda.empty_cache()


def _packed_index(n_ao):
    return np.tril_indices(n_ao)


def _block_ao(pc, sp0, sp1):
    """AO indices (i, j) of the compact rows (sp - sp0) * nij + ij of shell pairs [sp0, sp1)."""
    ni, nj = 2 * pc.la + 1, 2 * pc.lb + 1
    sp = np.arange(sp0, sp1)
    I = pc.sp_i[sp].astype(np.int64)[:, None, None] + np.arange(ni)[None, :, None]
    J = pc.sp_j[sp].astype(np.int64)[:, None, None] + np.arange(nj)[None, None, :]
    shape = (len(sp), ni, nj)
    return np.broadcast_to(I, shape).ravel(), np.broadcast_to(J, shape).ravel()


def _compare(got, ref, what, bound_abs, bound_rel, arbitrate):
    """got, ref: (rows, nk) against the float64 oracle.  Rows with an element outside bound_abs +
    bound_rel |ref| are decided by ``arbitrate(rows) -> (long-double oracle, float64 oracle)`` values of
    those rows: they pass within bound_abs + bound_rel |ld| + 10 |oracle64 - ld|.  Returns (max error
    against the float64 oracle, rows decided by the long-double oracle)."""
    got, ref = np.asarray(got), np.asarray(ref)
    err = np.abs(got - ref)
    bad = np.unique(np.nonzero(err > bound_abs + bound_rel * np.abs(ref))[0])
    if len(bad):
        ld, o64 = arbitrate(bad)
        cond = np.abs(o64 - ld)
        e2 = np.abs(got[bad] - ld)
        ok = e2 <= bound_abs + bound_rel * np.abs(ld) + 10 * cond
        if not ok.all():
            k = np.unravel_index(int(np.argmax(np.where(ok, -1.0, e2))), e2.shape)
            raise AssertionError(f"{what}: row {int(bad[k[0]])} component {int(k[1])}: got {got[bad][k]!r}, "
                                 f"long-double oracle {ld[k]!r}, float64 oracle {o64[k]!r}")
    return (float(err.max()) if err.size else 0.0), len(bad)


def _arbiter(pd, order, sp_pp_off, pc, sp0, pk, lc, start):
    """arbitrate(rows) for compact rows of shell pairs from sp0 of pair class pc: the oracle of each
    involved shell pair alone, long double and float64."""
    nij, nj = (2 * pc.la + 1) * (2 * pc.lb + 1), 2 * pc.lb + 1

    def arbitrate(rows):
        ld = np.empty((len(rows), 2 * lc + 1))
        o64 = np.empty_like(ld)
        for sp in np.unique(sp0 + rows // nij):
            prims = order[sp_pp_off[sp]:sp_pp_off[sp + 1]]
            b_ld = oracle_shell_pair_block(oracle_ld_lib(), pd, prims, pc.la, pc.lb, pk, lc, int(start))
            b_64 = oracle_shell_pair_block(oracle_lib(), pd, prims, pc.la, pc.lb, pk, lc, int(start))
            for k in np.flatnonzero(sp0 + rows // nij == sp):
                mi, mj = divmod(int(rows[k]) % nij, nj)
                ld[k], o64[k] = b_ld[:, mi, mj], b_64[:, mi, mj]
        return ld, o64
    return arbitrate


def _q_aux(plan, J2c):
    diag = J2c.diagonal().cpu().numpy()
    out = []
    for ac in plan.aux:
        idx = ac.ks_start[:, None] + np.arange(2 * ac.lc + 1)[None, :]
        out.append(np.sqrt(np.abs(diag[idx]).max(axis=1)))
    return out


# ------------------------------------------------------------------------------------------------
# 1. small
# ------------------------------------------------------------------------------------------------

def test_small_workload_all_integrals_vs_oracle():
    from electron_integral_playground_amd.molecular_integral import (nuclear_attraction, overlap, three_center,
                                                                     two_center)
    mol = generators.make_molecule("small")
    _say(f"case small: n_ao={mol.n_ao} n_aux={mol.n_aux}")
    pairs = pair_utility.form_pair_data(mol, 1e-14)
    aux = pair_utility.form_auxiliary_pair_data(mol)
    lib = oracle_lib()
    assert_parity(overlap(pairs, mol), run_overlap_oracle(lib, pairs, mol), "S")
    charges = np.ascontiguousarray(mol.geometry)
    assert_parity(nuclear_attraction(pairs, mol, charges), run_nuclear_oracle(lib, pairs, mol, charges), "V")
    assert_parity(two_center(aux, mol), run_two_center_oracle(lib, aux, mol), "(P|Q)")
    assert_parity(three_center(pairs, aux, mol), run_three_center_oracle(lib, pairs, aux, mol), "(ij|P)")


# ------------------------------------------------------------------------------------------------
# 2. plain (80 atoms)
# ------------------------------------------------------------------------------------------------

@pytest.mark.timeout(900)
def test_plain_80_atoms_vs_oracle():
    import torch
    from electron_integral_playground_amd.engine import ThreeCenterPlan
    from electron_integral_playground_amd.molecular_integral import nuclear_attraction, overlap
    _free()
    mol = generators.make_molecule("plain")
    pairs = pair_utility.form_pair_data(mol, 1e-14)
    aux = pair_utility.form_auxiliary_pair_data(mol)
    n_pp = sum(len(p.P_p) for p in pairs.values())
    _say(f"case plain: n_ao={mol.n_ao} n_aux={mol.n_aux} primitive pairs={n_pp}")
    assert mol.n_ao == 1120 and mol.n_aux == 10240
    plan = ThreeCenterPlan(mol, pairs, aux)
    J2c = torch.zeros((mol.n_aux, mol.n_aux), dtype=torch.float64, device="cuda")
    out = plan.run(metric_out=J2c)
    torch.cuda.synchronize()
    lib = oracle_lib()
    assert_parity(J2c.cpu().numpy(), run_two_center_oracle(lib, aux, mol), "plain (P|Q)")
    del J2c
    sample = sample_aux_shells(aux, 16, seed=2)
    ti, tj = _packed_index(mol.n_ao)
    worst = 0.0
    for k_L, starts in sample.items():
        nk = 2 * k_L + 1
        cols = torch.as_tensor((starts[:, None] + np.arange(nk)[None, :]).ravel(), device="cuda")
        gpu = out.index_select(1, cols).cpu().numpy().reshape(-1, len(starts), nk)
        for s, start in enumerate(starts):
            ref = oracle_aux_shell_block(lib, pairs, aux[k_L], k_L, int(start), mol.n_ao)
            worst = max(worst, assert_parity(gpu[:, s, :], ref[:, ti, tj].T, f"plain (ij|P) l={k_L} shell {start}"))
        _say(f"  aux l={k_L}: {len(starts)} shells x {n_pp} pairs, worst |HIP - oracle| so far {worst:.3e}")
    del out, plan
    _free()
    # one-electron integrals of the same molecule through the drop-in entry points
    assert_parity(overlap(pairs, mol), run_overlap_oracle(lib, pairs, mol), "plain S")
    charges = np.ascontiguousarray(mol.geometry)
    assert_parity(nuclear_attraction(pairs, mol, charges), run_nuclear_oracle(lib, pairs, mol, charges), "plain V")


# ------------------------------------------------------------------------------------------------
# 3. batched (first wave)
# ------------------------------------------------------------------------------------------------

BATCH_BUDGET = 200e9


@pytest.mark.timeout(900)
def test_batched_first_wave_vs_oracle():
    import torch
    from electron_integral_playground_amd.engine import BatchPlan
    _free()
    sizes = generators.batched_sizes(n_mol=256)
    wave, acc = [], 0
    for i, n in enumerate(sizes):
        mol = generators.batched_molecule(i, int(n))
        b = 8 * (mol.n_ao * (mol.n_ao + 1) // 2 * mol.n_aux + mol.n_aux ** 2)
        if wave and acc + b > BATCH_BUDGET:
            break
        wave.append((mol, pair_utility.form_pair_data(mol, 1e-14), pair_utility.form_auxiliary_pair_data(mol)))
        acc += b
    atoms = [len(m.elements) for m, _, _ in wave]
    _say(f"case batched: first wave of {len(wave)} molecules ({min(atoms)}-{max(atoms)} atoms), {acc / 1e9:.1f} GB out")
    bp = BatchPlan(wave)
    out, out2 = bp.new_output(), bp.new_two_center_output()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        bp.run(out)                                   # warm-up outside capture
        bp.two_center(out2)
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        bp.run(out)
        bp.two_center(out2)
    out.fill_(float("nan"))
    out2.fill_(float("nan"))
    g.replay()
    torch.cuda.synchronize()
    lib = oracle_lib()
    small, large = int(np.argmin(atoms)), int(np.argmax(atoms))
    for m, (mol, pairs, aux) in enumerate(wave):
        got = bp.molecule_output(out, m)
        ti, tj = _packed_index(mol.n_ao)
        if m == small:
            ref = run_three_center_oracle(lib, pairs, aux, mol)
            assert_parity(got.cpu().numpy(), ref[:, ti, tj].T, f"batched molecule {m} (ij|P)")
        else:
            sample = sample_aux_shells(aux, 8 if m == large else 1, seed=300 + m)
            for k_L, starts in sample.items():
                nk = 2 * k_L + 1
                for start in starts:
                    cols = torch.arange(int(start), int(start) + nk, device="cuda")
                    g_cols = got.index_select(1, cols).cpu().numpy()
                    ref = oracle_aux_shell_block(lib, pairs, aux[k_L], k_L, int(start), mol.n_ao)
                    assert_parity(g_cols, ref[:, ti, tj].T, f"batched molecule {m} l={k_L} shell {start}")
        if m in (small, large):
            assert_parity(bp.molecule_two_center(out2, m).cpu().numpy(), run_two_center_oracle(lib, aux, mol),
                          f"batched molecule {m} (P|Q)")
    _say(f"  {len(wave)} molecules checked (molecule {small} in full, {large} on 8 shells per class)")
    del g, out, out2, bp
    _free()


# ------------------------------------------------------------------------------------------------
# 4. mixed precision (200 atoms)
# ------------------------------------------------------------------------------------------------

@pytest.mark.timeout(1200)
def test_mixed_200_atoms_float64_and_mixed_vs_oracle():
    import torch
    from electron_integral_playground_amd.engine import ThreeCenterPlan
    _free()
    mol = generators.make_molecule("mixed")
    arrays = pair_utility.form_pair_arrays(mol, 1e-14)
    pairs = pair_utility.form_primitive_pair_data(mol, arrays)
    aux = pair_utility.form_auxiliary_pair_data(mol)
    n_pp = sum(len(a) for a in arrays.values())
    _say(f"case mixed: n_ao={mol.n_ao} n_aux={mol.n_aux} primitive pairs={n_pp}")
    assert mol.n_ao == 6000 and mol.n_aux == 40000
    plan = ThreeCenterPlan(mol, pairs, aux)
    J2c = plan.two_center()
    q = _q_aux(plan, J2c)
    del J2c
    _free()
    stats = plan.setup_mixed(arrays, q, threshold=1e-6)
    _say(f"  split {stats}")
    assert stats["f32_triplets"] > 0 and stats["f64_triplets"] > 0
    # sample: per class, shells drawn at random plus shells of the 64-shell groups with the most
    # float32 work items (so the float32 tier is part of what is checked)
    rng = np.random.default_rng(4)
    sample = {}
    f32_cols = set()
    for ai, ac in enumerate(plan.aux):
        counts = np.zeros((ac.n_ks + 63) // 64, dtype=np.int64)
        for pi in range(len(plan.pairs)):
            it32 = plan.mixed[(pi, ai)]["it32"]
            if it32.shape[0]:
                counts += np.bincount(it32[:, 1].cpu().numpy() // 64, minlength=len(counts))
        n_rand = 2 if ac.lc > 0 else 3
        ks = list(rng.choice(ac.n_ks, size=n_rand, replace=False))
        top = int(np.argmax(counts))
        if counts[top] > 0:
            k = int(top * 64 + rng.integers(0, min(64, ac.n_ks - top * 64)))
            ks.append(k)
            f32_cols.add(int(ac.ks_start[k]))
        sample[ac.lc] = np.unique(ac.ks_start[np.array(ks)])
    assert f32_cols, "no float32 work in the sampled shells"
    col_list = np.concatenate([(st[:, None] + np.arange(2 * lc + 1)[None, :]).ravel() for lc, st in sample.items()])
    cols = torch.as_tensor(col_list, device="cuda")
    _say(f"  sampled shells {({lc: st.tolist() for lc, st in sample.items()})}, {len(col_list)} columns")
    blocks = plan.row_blocks(20e9)
    plan.prepare_blocks(blocks, mixed=True)
    rows = max(plan.block_rows(*b) for b in blocks)
    A = torch.empty((rows, mol.n_aux), dtype=torch.float64, device="cuda")
    B = torch.empty_like(A)
    got64, got32 = [], []
    for bi, b in enumerate(blocks):
        n = plan.block_rows(*b)
        plan.run_block(*b, A)
        plan.run_block(*b, B, mixed=True)
        got64.append(A[:n].index_select(1, cols))
        got32.append(B[:n].index_select(1, cols))
    torch.cuda.synchronize()
    _say(f"  streamed {len(blocks)} blocks (float64 and mixed, Schwarz-type split)")
    # the default split: the a-priori term-scale (envelope) estimate
    stats_e = plan.setup_mixed(arrays, q, threshold=1e-6, estimate="envelope")
    _say(f"  envelope split {stats_e}")
    assert stats_e["f32_triplets"] > 10 * stats["f32_triplets"]
    plan.prepare_blocks(blocks, mixed=True)
    gotE = []
    for bi, b in enumerate(blocks):
        n = plan.block_rows(*b)
        plan.run_block(*b, B, mixed=True)
        gotE.append(B[:n].index_select(1, cols))
    torch.cuda.synchronize()
    _say(f"  streamed {len(blocks)} blocks (mixed, envelope split)")
    del A, B
    _free()
    ao = [_block_ao(plan.pairs[pi], sp0, sp1) for pi, sp0, sp1 in blocks]
    lib = oracle_lib()
    c0 = 0
    worst64 = worst32 = worstE = 0.0
    sq32 = sqE = 0.0
    n_el = 0
    n_arb = 0
    for lc, starts in sample.items():
        nk = 2 * lc + 1
        ai = [a.lc for a in plan.aux].index(lc)
        for start in starts:
            ref = oracle_aux_shell_block(lib, pairs, aux[lc], lc, int(start), mol.n_ao)
            for bi, (I, J) in enumerate(ao):
                pi, sp0, _ = blocks[bi]
                pc = plan.pairs[pi]
                arb = _arbiter(pairs[(pc.la, pc.lb)], plan._pair_orders[pi], pc.sp_pp_off, pc, sp0, aux[lc], lc, start)
                r = ref[:, I, J].T
                g64 = got64[bi][:, c0:c0 + nk].cpu().numpy()
                g32 = got32[bi][:, c0:c0 + nk].cpu().numpy()
                e64, n64 = _compare(g64, r, f"mixed workload float64 l={lc} shell {start} block {bi}", ATOL, RTOL, arb)
                e32, n32 = _compare(g32, r, f"mixed workload mixed l={lc} shell {start} block {bi}", MIXED_BOUND, 0.0, arb)
                gE = gotE[bi][:, c0:c0 + nk].cpu().numpy()
                eE, nE = _compare(gE, r, f"mixed workload envelope split l={lc} shell {start} block {bi}", MIXED_BOUND,
                                  0.0, arb)
                worst64, worst32, worstE = max(worst64, e64), max(worst32, e32), max(worstE, eE)
                n_arb += n64 + n32 + nE
                d = np.abs(g32 - r)
                sq32 += float((d * d).sum())
                d = np.abs(gE - r)
                sqE += float((d * d).sum())
                n_el += d.size
            _say(f"  l={lc} shell {start} (aux class {ai}): worst vs float64 oracle: float64 path {worst64:.3e}, "
                 f"mixed {worst32:.3e}, envelope split {worstE:.3e}; rows decided by the long-double oracle so far "
                 f"{n_arb}")
            c0 += nk
    _say(f"  mixed vs oracle: Schwarz split max abs {worst32:.3e}, RMS {(sq32 / n_el) ** 0.5:.3e}; envelope split "
         f"max abs {worstE:.3e}, RMS {(sqE / n_el) ** 0.5:.3e}; over {n_el} sampled elements, {n_arb} rows decided "
         "by the long-double oracle")
    assert n_arb <= 1e-4 * n_el
    del got64, got32, gotE, plan
    _free()


# ------------------------------------------------------------------------------------------------
# 5. serving
# ------------------------------------------------------------------------------------------------

class _Sampler:
    """Server consumer: keeps the sampled auxiliary columns of every finished unit (on the device) with
    the AO indices of its rows, before the unit's memory is reused."""

    def __init__(self, cols_of_job):
        self.cols_of_job = cols_of_job
        self.got = {}

    def __call__(self, job, unit):
        import torch
        cols = torch.as_tensor(self.cols_of_job[job.job_id], device=unit.out.device)
        for (pi, sp0, sp1), r0 in zip(unit.blocks, unit.row_off):
            pc = job.plan.pairs[pi]
            n = (sp1 - sp0) * (2 * pc.la + 1) * (2 * pc.lb + 1)
            I, J = _block_ao(pc, sp0, sp1)
            self.got.setdefault(job.job_id, []).append(((pi, sp0, sp1), I, J,
                                                        unit.out[r0:r0 + n].index_select(1, cols).clone()))


@pytest.mark.timeout(1200)
def test_serving_mix_vs_oracle_and_alone():
    import torch
    from electron_integral_playground_amd.serving import IntegralServer
    _free()
    specs = [(0, 10, 61), (1, 40, 62), (2, 130, 63)]
    mols = {jid: generators.serving_molecule(n, seed) for jid, n, seed in specs}
    host = {jid: (pair_utility.form_pair_data(m, 1e-14), pair_utility.form_auxiliary_pair_data(m))
            for jid, m in mols.items()}
    samples = {jid: sample_aux_shells(host[jid][1], {0: 1, 1: 1, 2: 1, 3: 1, 4: 2}, seed=500 + jid) for jid in mols}
    cols = {jid: np.concatenate([(st[:, None] + np.arange(2 * lc + 1)[None, :]).ravel() for lc, st in s.items()])
            for jid, s in samples.items()}

    def serve(job_ids):
        sampler = _Sampler(cols)
        srv = IntegralServer(budget_bytes=200e9, mixed=True, keep_outputs=False, on_unit=sampler)
        first = True
        for jid in job_ids:
            srv.submit(jid, mols[jid], size_hint=len(mols[jid].elements))
            if first:
                srv.poll()                  # the first job may be on the GPU before the others arrive
                first = False
        srv.drain(timeout=600)
        srv.close()
        assert all(j.error is None for j in srv.completed), [j.error for j in srv.completed]
        return sampler.got, srv.n_steps

    mix, steps = serve([0, 1, 2])
    _say(f"case serving mix: jobs {[(j, n) for j, n, _ in specs]}, {steps} steps")
    lib = oracle_lib()
    for jid, mol in mols.items():
        pairs, aux = host[jid]
        c0 = 0
        worst = 0.0
        for lc, starts in samples[jid].items():
            nk = 2 * lc + 1
            for start in starts:
                ref = oracle_aux_shell_block(lib, pairs, aux[lc], lc, int(start), mol.n_ao)
                for _, I, J, g in mix[jid]:
                    worst = max(worst, float(np.abs(g[:, c0:c0 + nk].cpu().numpy() - ref[:, I, J].T).max()))
                c0 += nk
        _say(f"  job {jid} ({len(mol.elements)} atoms, {len(mix[jid])} blocks): max |served - oracle| {worst:.3e}")
        assert worst <= MIXED_BOUND, (jid, worst)
    # each job alone: the same sampled values bit for bit
    for jid, _, _ in specs:
        alone, _ = serve([jid])
        a = {blk: g for blk, _, _, g in alone[jid]}
        m = {blk: g for blk, _, _, g in mix[jid]}
        assert a.keys() == m.keys(), jid
        for blk in a:
            assert torch.equal(a[blk], m[blk]), (jid, blk)
        _say(f"  job {jid} alone equals mix ({len(a)} blocks)")
    _free()
