"""Continuous-batching server: results are the same alone and in a mix of concurrent jobs, and equal the
single-molecule engine's."""

import numpy as np
import pytest

from electron_integral_playground_amd import generators, pair_utility

pytestmark = pytest.mark.gpu


def _job(n_atom, seed):
    return generators.serving_molecule(n_atom, seed)


def _drain(server, jobs):
    for jid, mol in jobs:
        server.submit(jid, mol)
    server.drain(timeout=300)
    return server


@pytest.mark.parametrize("mixed,persistent", [(False, False), (True, False), (False, True), (True, True)])
def test_alone_equals_mix(mixed, persistent):
    import torch
    from electron_integral_playground_amd.serving import IntegralServer
    print(f"case mixed={mixed} persistent={persistent}", flush=True)
    a = _job(10, 31)
    alone = _drain(IntegralServer(budget_bytes=8e9, mixed=mixed, keep_outputs=True, persistent=persistent), [(0, a)])
    others = [(1, _job(12, 32)), (2, _job(10, 33)), (3, _job(11, 34))]
    mix = IntegralServer(budget_bytes=8e9, mixed=mixed, keep_outputs=True, persistent=persistent)
    mix.submit(1, others[0][1])
    mix.poll()                                   # job 1 may already be on the GPU when job 0 arrives
    mix.submit(0, a)
    for jid, mol in others[1:]:
        mix.submit(jid, mol)
    mix.drain(timeout=300)
    assert mix.n_steps >= 1 and all(j.error is None for j in mix.completed)
    got_alone, got_mix = alone.job_output(0), mix.job_output(0)
    assert got_alone.keys() == got_mix.keys()
    for pi in got_alone:
        assert torch.equal(got_alone[pi], got_mix[pi]), pi
    # the metric is computed per job in float64
    assert torch.equal(alone.jobs[0].J2c, mix.jobs[0].J2c)


@pytest.mark.parametrize("persistent", [False, True])
def test_server_matches_engine_blocks(persistent):
    """Plain float64 server output equals ThreeCenterPlan.run_block -- bit for bit with the class
    kernels, to a few roundings through the persistent queue kernels (the same class bodies compiled in
    another function); a budget small enough to split the job into shell-pair blocks gives the same
    values bit for bit."""
    import torch
    from electron_integral_playground_amd.engine import ThreeCenterPlan
    from electron_integral_playground_amd.serving import IntegralServer
    mol = _job(10, 41)
    srv = _drain(IntegralServer(budget_bytes=8e9, mixed=False, keep_outputs=True, persistent=persistent), [(0, mol)])
    small = _drain(IntegralServer(budget_bytes=8e9, mixed=False, keep_outputs=True, unit_cap_bytes=5e7,
                                  persistent=persistent), [(0, mol)])
    assert len(small.jobs[0].units) > len(srv.jobs[0].units)
    arrays = pair_utility.form_pair_arrays(mol, 1e-14)
    plan = ThreeCenterPlan(mol, pair_utility.form_primitive_pair_data(mol, arrays),
                           pair_utility.form_auxiliary_pair_data(mol))
    out_srv, out_small = srv.job_output(0), small.job_output(0)
    for pi, pc in enumerate(plan.pairs):
        ref = torch.empty((plan.block_rows(pi, 0, pc.n_sp), mol.n_aux), dtype=torch.float64, device="cuda")
        plan.run_block(pi, 0, pc.n_sp, ref)
        torch.cuda.synchronize()
        if persistent:
            assert bool(((out_srv[pi] - ref).abs() <= 1e-15 + 1e-13 * ref.abs()).all()), pi
        else:
            assert torch.equal(out_srv[pi], ref), pi
        assert torch.equal(out_small[pi], out_srv[pi]), pi


def test_consumed_outputs_match_kept_ones():
    """Without keep_outputs the server reduces every output -- the metric right after setup, each unit of
    (ij|P) when its step retires -- to (sum, 2-norm) on the device; these equal the same reductions of the
    outputs a keep_outputs server holds."""
    import torch
    from electron_integral_playground_amd.serving import IntegralServer
    mol = _job(10, 51)
    kept = _drain(IntegralServer(budget_bytes=8e9, mixed=True, keep_outputs=True), [(0, mol)])
    used = _drain(IntegralServer(budget_bytes=8e9, mixed=True, keep_outputs=False), [(0, mol)])
    J = kept.jobs[0].J2c
    assert used.jobs[0].J2c is None
    assert torch.equal(used.jobs[0].J2c_checksum, torch.stack([J.sum(), torch.linalg.vector_norm(J)]))
    assert len(kept.jobs[0].units) == len(used.jobs[0].units)
    for uk, uu in zip(kept.jobs[0].units, used.jobs[0].units):
        assert torch.equal(uk.checksum, uu.checksum)
