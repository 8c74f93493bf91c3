"""Persistent queue kernels (one kernel per register tier over a device work queue).

* A batch of molecules run through the tier kernels equals the per-class launches to the last bits (the
  tier kernels run the same class bodies, compiled in another function: a few roundings may differ),
  the smallest molecule equals the oracle in full, and the others on sampled auxiliary shells.
* The queue's control block is left ready by each launch: back-to-back launches and the replays of one
  captured graph give the same result bit for bit, also after a different batch's queue was uploaded into
  the same device buffers (the graph is replayed, not rebuilt).
* Every class of the specialised range (orbital l <= 3, auxiliary l <= 4, float64 and float32) runs
  through the queue on a small molecule and equals that class's own kernels (to rounding where the
  tuned kernel is a matrix-core variant; float32 to float32 rounding).
"""

import numpy as np
import pytest

from tests.helpers import (assert_parity, oracle_aux_shell_block, oracle_lib, random_molecule, run_three_center_oracle,
                           sample_aux_shells)
from electron_integral_playground_amd import generators, pair_utility

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _close(a, b):
    """Equal to a few roundings: |a - b| <= 1e-15 + 1e-13 |b| everywhere (NaN fails)."""
    import torch
    return bool(((a - b).abs() <= 1e-15 + 1e-13 * b.abs()).all())


def _batch(seed, n_mol, atoms):
    out = []
    for i in range(n_mol):
        mol = generators.batched_molecule(seed + i, int(atoms[i % len(atoms)]))
        out.append((mol, pair_utility.form_pair_data(mol, 1e-14), pair_utility.form_auxiliary_pair_data(mol)))
    return out


def test_queue_batch_equals_class_launches_and_oracle():
    import torch
    from electron_integral_playground_amd.engine import BatchPlan
    wave = _batch(500, 6, (4, 9, 14, 6, 11, 5))
    _say(f"case queue batch: {len(wave)} molecules, atoms {[len(m.elements) for m, _, _ in wave]}")
    bp = BatchPlan(wave)
    ref = bp.new_output()
    bp.run(ref)
    out = bp.new_output().fill_(float("nan"))
    q = bp.work_queue(out)
    q.upload()
    bp.run(out, queue=q)
    torch.cuda.synchronize()
    _say(f"max |queue - class launches| = {float((out - ref).abs().max()):.2e}")
    assert _close(out, ref)
    first = out.clone()
    # back to back: the control block was reset by the previous launch
    out.fill_(float("nan"))
    bp.run(out, queue=q)
    torch.cuda.synchronize()
    assert torch.equal(out, first)
    lib = oracle_lib()
    atoms = [len(m.elements) for m, _, _ in wave]
    small = int(np.argmin(atoms))
    ti, tj = np.tril_indices(wave[small][0].n_ao)
    mol, pairs, aux = wave[small]
    assert_parity(bp.molecule_output(out, small).cpu().numpy(), run_three_center_oracle(lib, pairs, aux, mol)[:, ti, tj].T,
                  "queue batch smallest molecule")
    for m, (mol, pairs, aux) in enumerate(wave):
        if m == small:
            continue
        ti, tj = np.tril_indices(mol.n_ao)
        got = bp.molecule_output(out, m)
        for k_L, starts in sample_aux_shells(aux, 1, seed=700 + m).items():
            for start in starts:
                cols = torch.arange(int(start), int(start) + 2 * k_L + 1, device="cuda")
                r = oracle_aux_shell_block(lib, pairs, aux[k_L], k_L, int(start), mol.n_ao)
                assert_parity(got.index_select(1, cols).cpu().numpy(), r[:, ti, tj].T, f"queue molecule {m} l={k_L}")


def test_queue_graph_replayed_over_new_work():
    import torch
    from electron_integral_playground_amd.engine import BatchPlan, WorkQueue
    plans = [BatchPlan(_batch(600, 4, (5, 8, 12, 7))), BatchPlan(_batch(650, 5, (9, 4, 6, 10, 13)))]
    outs = [bp.new_output() for bp in plans]
    queues = [bp.work_queue(o) for bp, o in zip(plans, outs)]
    refs = []
    for bp, o, q in zip(plans, outs, queues):
        q.upload()                            # eager queue runs: the reference of every replay
        bp.run(o, queue=q)
        torch.cuda.synchronize()
        refs.append(o.clone())
        r = bp.new_output()
        bp.run(r)
        assert _close(o, r)
        q.dev = None
    # device buffers large enough for both queues
    bufs = WorkQueue.buffers_for(queues)
    queues[0].upload(into=bufs)
    assert queues[0].dev is bufs
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        queues[0].run(s)                      # warm-up outside capture
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        queues[0].run()
    for rep in range(3):
        k = rep % 2
        queues[k].upload(into=bufs)
        assert queues[k].dev is bufs
        outs[k].fill_(float("nan"))
        plans[k].zero_unwritten(outs[k])      # rows of screened-out shell pairs (not in the queue)
        for pi, ai in queues[k].fallback:     # classes that keep their own kernels
            plans[k]._launch_class(pi, ai, outs[k], None)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(outs[k], refs[k]), f"replay {rep} (batch {k})"
        _say(f"replay {rep}: batch {k} equal")
        # eager work between replays
        plans[1 - k].run(outs[1 - k])
        torch.cuda.synchronize()


@pytest.mark.parametrize("use_f32", [False, True])
def test_queue_every_class_equals_class_kernels(use_f32):
    import torch
    from electron_integral_playground_amd.engine import ThreeCenterPlan, WorkQueue
    mol = random_molecule(n_atom=3, seed=31, orbital_l=(0, 1, 2, 3), aux_l=(0, 1, 2, 3, 4))
    plan = ThreeCenterPlan(mol, pair_utility.form_pair_data(mol, 1e-14), pair_utility.form_auxiliary_pair_data(mol))
    ref = plan.new_output()
    out = plan.new_output().fill_(float("nan"))
    q = WorkQueue(plan.lib, plan.device, use_f32=use_f32)
    entries = []
    for pi, pc in enumerate(plan.pairs):
        for ai, ac in enumerate(plan.aux):
            plan.launch_class(pi, ai, ref, use_f32=use_f32)
            it = plan.class_items(pi, ai)
            if it is None or it.shape[0] == 0:
                continue
            # the queue runs the vector bodies: the job reads the vector pair matrices
            dp, da = pc.dev, ac.dev
            entries.append([dp["off"].data_ptr(), dp["si"].data_ptr(), dp["sj"].data_ptr(), dp["P"].data_ptr(),
                            dp["E"].data_ptr(), dp["E32"].data_ptr(), 0, da["off"].data_ptr(), da["col"].data_ptr(),
                            da["C"].data_ptr(), da["c"].data_ptr(), ac.n_ks, out.data_ptr(), plan.n_aux, 0, 0])
    table = torch.tensor(np.array(entries, dtype=np.int64), device="cuda")
    k = 0
    own = []
    for pi, pc in enumerate(plan.pairs):
        for ai, ac in enumerate(plan.aux):
            it = plan.class_items(pi, ai)
            if it is None or it.shape[0] == 0:
                continue
            if q.supports(pc.la, pc.lb, ac.lc):
                q.add(pc.la, pc.lb, ac.lc, it, int(it.shape[0]), table.data_ptr() + 128 * k,
                      np.ones(int(it.shape[0])), layout=0)
            else:
                own.append((pi, ai))
            k += 1
    _say(f"{q.n_triplet_items} items of {k - len(own)} classes through the queue; "
         f"{len(own)} classes on their own kernels: {[(plan.pairs[p].la, plan.pairs[p].lb, plan.aux[a].lc) for p, a in own]}")
    assert k - len(own) >= 0.7 * k
    q.build().upload()
    plan.zero_unwritten(ref)
    plan.zero_unwritten(out)
    for pi, ai in own:
        plan.launch_class(pi, ai, out, use_f32=use_f32)
    q.run()
    torch.cuda.synchronize()
    if use_f32:
        err = float((out - ref).abs().max())
        _say(f"float32 tier through the queue: max |queue - class kernels| = {err:.2e}")
        assert err <= 1e-5 * max(1.0, float(ref.abs().max()))
    else:
        # classes whose tuned float64 variant runs on the matrix cores differ at rounding level (the queue
        # runs the vector bodies); every other class is bit for bit equal
        err = float((out - ref).abs().max())
        _say(f"float64 through the queue: max |queue - class kernels| = {err:.2e}")
        assert err <= 1e-13 * max(1.0, float(ref.abs().max()))


def test_run_persistent_shared_graph():
    """BatchPlan.run_persistent: one QueueGraph for several batches (captured again only while its buffers
    grow to the largest batch), with the zeroing and (P|Q) beside it, equal to the per-class path for
    every batch."""
    import torch
    from electron_integral_playground_amd.engine import BatchPlan, QueueGraph
    plans = [BatchPlan(_batch(800 + 10 * k, 3 + k, (6, 10, 4, 12))) for k in range(3)]
    graph = QueueGraph()
    s = torch.cuda.Stream()
    first_pass = None
    for rep in range(2):
        if rep == 1:
            first_pass = graph.captures
        for k, bp in enumerate(plans):
            ref, ref2 = bp.new_output(), bp.new_two_center_output()
            bp.run(ref)
            bp.two_center(ref2)
            out, out2 = bp.new_output().fill_(float("nan")), bp.new_two_center_output().fill_(float("nan"))
            q = bp.work_queue(out)
            torch.cuda.synchronize()
            s.wait_stream(torch.cuda.current_stream())
            bp.run_persistent(out, q, graph, out2c=out2, stream=s)
            torch.cuda.current_stream().wait_stream(s)
            torch.cuda.synchronize()
            assert _close(out, ref), f"batch {k} pass {rep}"
            assert torch.equal(out2, ref2), f"batch {k} pass {rep} (P|Q)"
    _say(f"3 batches x 2 passes through one graph, captured {first_pass}x in the first pass, "
         f"{graph.captures - first_pass}x in the second")
    # the buffers only grow: once every batch has been seen, every queue fits and the graph is replayed
    assert 1 <= first_pass <= 3 and graph.captures == first_pass
