"""Host logic of the persistent queue kernels (no GPU): every class of the specialised range (but the few
highest, whose bodies need more registers than the persistent loop leaves) has a queue plan whose
component chunks tile the auxiliary shell and whose tiers exist; slots carry their class's
kind and tier; the queue's entries (runs of consecutive items of one slot, cheap items grouped) cover
every (slot, item) exactly once, ordered by class (decreasing mean item cost) and longest-first within a
class, dealt in runs of QUEUE_RUN entries over the 8 XCD parts, with a control block (one cache line per
part counter) that describes the parts."""

import ctypes

import numpy as np

from electron_integral_playground_amd import library
from electron_integral_playground_amd.engine import QUEUE_PARTS, QUEUE_RUN, WorkQueue


def _plan(lib, la, lb, lc, f32):
    m0, m1, tier = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    out, c = [], 0
    while lib.eip_three_center_queue_chunk(la, lb, lc, f32, c, ctypes.byref(m0), ctypes.byref(m1),
                                           ctypes.byref(tier)) == 0:
        out.append((m0.value, m1.value, tier.value))
        c += 1
    return out


def test_queue_plans_tile_every_class():
    lib = library()
    missing = []
    for f32 in (0, 1):
        ntiers = lib.eip_three_center_queue_tiers(f32)
        assert ntiers >= 1
        for lb in range(4):
            for la in range(lb + 1):
                for lc in range(5):
                    plan = _plan(lib, la, lb, lc, f32)
                    if not plan:
                        # only classes whose bodies need more registers than the persistent loop leaves
                        assert la + lb >= 5 and lc >= 2, (la, lb, lc, f32)
                        missing.append((la, lb, lc, f32))
                        continue
                    assert plan[0][0] == 0 and plan[-1][1] == 2 * lc + 1
                    assert all(a[1] == b[0] for a, b in zip(plan, plan[1:]))
                    assert all(0 <= t < ntiers for _, _, t in plan)
    assert len(missing) <= 8, missing
    # every class of the batched workload (orbital l <= 2, auxiliary l <= 3) is in the queue kernels
    assert not [m for m in missing if m[1] <= 2 and m[2] <= 3]
    assert _plan(lib, 0, 4, 0, 0) == []          # outside the specialised range


def test_queue_order_parts_and_control_block():
    import torch
    from electron_integral_playground_amd.engine import QUEUE_ENTRY_COST, QUEUE_MAX_GROUP
    lib = library()
    q = WorkQueue(lib, device="cpu")
    rng = np.random.default_rng(3)
    dummy = torch.zeros(4, dtype=torch.int32)
    classes = [(0, 0, 0, 700), (1, 2, 3, 350), (2, 2, 2, 90), (0, 1, 4, 1)]
    costs = {}
    for la, lb, lc, n in classes:
        c = rng.integers(1, 50, n).astype(np.float64)
        costs[(la, lb, lc)] = c
        q.add(la, lb, lc, dummy, n, 0x1000, c)
    q.build()
    slot_class = [(la, lb, lc) for la, lb, lc, n in classes for _ in _plan(lib, la, lb, lc, 0)]
    assert len(q.host["slots"]) == len(q._slots) * lib.eip_three_center_queue_slot_bytes()
    words = lib.eip_three_center_queue_ctl_words()
    covered = {}
    for t in range(q.ntiers):
        ent = q.host["entries"][t]
        ctl = q.host["ctl"][t]
        assert ent.shape[1] == 4 and ctl.shape == (words,)
        # counters (one 128-byte line each) and the finish count are zero before a launch
        assert not ctl[0:256].any() and not ctl[288:].any()
        offs, cnts = ctl[256:264], ctl[264:272]
        assert cnts.sum() == len(ent)
        assert (offs == np.concatenate([[0], np.cumsum(cnts)[:-1]])).all()
        # undo the dealing: part p holds the runs p, p + 8, ... of QUEUE_RUN entries of the ordered list
        n = len(ent)
        if n == 0:                      # a tier none of these classes falls in
            continue
        part = (np.arange(n) // QUEUE_RUN) % QUEUE_PARTS
        ordered = np.empty_like(ent)
        ordered[np.argsort(part, kind="stable")] = ent
        # entries of one slot are contiguous; within it, runs by decreasing cost
        change = np.flatnonzero(np.diff(ordered[:, 0])) + 1
        runs = np.split(ordered, change)
        assert len({int(r[0, 0]) for r in runs}) == len(runs)
        for r in runs:
            c = costs[slot_class[int(r[0, 0])]]
            run_cost = [c[i:i + k].sum() for _, i, k, _ in r]
            assert all(a >= b for a, b in zip(run_cost, run_cost[1:]))
            assert (r[:, 2] >= 1).all() and (r[:, 2] <= QUEUE_MAX_GROUP).all()
            for _, i, k, _ in r:
                covered.setdefault((t, int(r[0, 0])), []).extend(range(i, i + k))
    # every item of every slot exactly once
    for s, cls in enumerate(slot_class):
        n = len(costs[cls])
        got = [v for (t, ss), items in covered.items() if ss == s for v in items]
        assert sorted(got) == list(range(n)), (s, cls)
    # cheap items are grouped: the (ss|s) slot's 700 items take far fewer entries
    s0 = slot_class.index((0, 0, 0))
    n0 = sum(len(r) for (t, ss), r in covered.items() if ss == s0)
    assert n0 == 700
