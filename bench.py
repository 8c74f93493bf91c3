"""Benchmark of the McMurchie-Davidson three-centre path on one MI355X; prints one JSON line.

value (headline): plain FP64 3c integrals/s of the 200-atom workload (20 A box, orbital 4s3p2d1f with
1-6 primitives per shell, auxiliary s-g with 8 even-tempered exponents per l; n_ao 6000, n_aux 40000,
7 840 675 primitive pairs) on its fixed auxiliary-block share: every primitive pair against the
auxiliary shells of the first 24 atoms (192 shells of each angular class l = 0..4, 4800 auxiliary
functions, 12 % of n_aux -- the share one of eight GPUs computes when the tensor is sharded by
auxiliary blocks).  One step writes the slice's compact (ij|P) -- 12.2 million significant-shell-pair
rows x 4800 columns, 469 GB -- through a 40 GB device buffer in blocks of shell pairs, inputs resident
in HBM before timing.  value = 4800 * n_ao (n_ao + 1) / 2 / (seconds per step).

Beside it, in the same line:
* mixed_precision: the same slice with the float32 tier (triplets whose bound is below 1e-6) --
  integrals/s, ratio to plain, max abs / RMS difference from plain over the whole slice, and max abs
  difference from the CPU reference on the sampled shells;
* roofline of the dominant kernel class (HIP events on the launch stream, algorithmic flops of the fixed
  per-triplet model, HBM traffic from the committed rocprofv3 counter summary);
* cpu_baseline: the reference's OpenMP C++ (oracle/_ref; the oracle's restatement where that build is
  absent) on a bounded sample of the same slice (whole auxiliary shells x all pairs), with its parity;
* plain80: the 80-atom workload's full step (n_ao 1120, n_aux 10240: every (ij|P) into the packed
  51.4 GB device output plus (P|Q)), with its own roofline;
* batched: 256 molecules (lognormal sizes) in waves of HIP-graph-captured batch plans;
* serving: a short Poisson stream through the continuous-batching server (p50/p99 latency);
  ``--serving --serving-jobs N`` runs a long stream alone, ``--mixed-full`` the full 3.9 TB comparison.
"""

import argparse
import csv
import ctypes
import json
import os
import pathlib
import re
import sys
import threading
import time

import numpy as np

# Hardware queues per process, read once when the HIP runtime starts: with HIP's default of 4, the
# serving loop's setup streams share in-order queues with the compute stream and a job's device setup
# waits behind a multi-second step of a large job (measured: small jobs 14-20 s in setup with 4 queues,
# none with 16).  Raised here, before anything initialises the runtime.
if int(os.environ.get("GPU_MAX_HW_QUEUES", "4")) < 16:
    os.environ["GPU_MAX_HW_QUEUES"] = "16"

ROOT = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))
import __graft_entry__  # noqa: E402

PEAK_FP64 = 78.6e12          # MI355X FP64 vector peak, FLOP/s
PEAK_HBM = 8.0e12            # bytes/s
SLICE_ATOMS = 24             # auxiliary shells of the first 24 atoms: 192 per angular class
BLOCK_BYTES = 40e9           # device buffer of one streamed block
CPU_SECONDS = 25.0           # bound of the CPU-baseline sample
PMC_SUMMARY = ROOT / "profiles" / "j3c_slice200_plain_pmc.md"
PMC_SUMMARY_PLAIN80 = ROOT / "profiles" / "j3c_plain80_pmc.md"
METRIC = "3c integrals/s, mixed vs plain FP64 at <=1e-10 err; p50/p99 job latency"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def measured_traffic(summary, cls, column=3):
    """HBM bytes of class (la, lb, lc) per step from a committed rocprofv3 FETCH_SIZE / WRITE_SIZE
    summary written by --pmc-summary (rows ``| j3c_<la><lb><lc> | fetch KiB | write KiB | calibrated
    bytes | ...``); None if absent."""
    if not summary.exists():
        return None
    key = "j3c_%d%d%d" % cls
    for line in summary.read_text().splitlines():
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) >= 4 and cells[0] == key:
            try:
                return float(cells[column])
            except ValueError:
                return None
    return None


def q_aux_from_metric(plan, J2c):
    """Q_P = sqrt(max_m (P_m|P_m)) per auxiliary shell, per class, in the plan's shell order."""
    diag = J2c.diagonal().cpu().numpy()
    out = []
    for ac in plan.aux:
        idx = ac.ks_start[:, None] + np.arange(2 * ac.lc + 1)[None, :]
        out.append(np.sqrt(np.abs(diag[idx]).max(axis=1)))
    return out


def slice_workload(n_slice_atoms=SLICE_ATOMS):
    """The 200-atom workload with its auxiliary basis cut to the shells of the first atoms (their
    functions are the first columns, since auxiliary offsets follow the atom order)."""
    import copy
    from electron_integral_playground_amd import generators, pair_utility
    t0 = time.perf_counter()
    mol = generators.make_molecule("mixed")
    arrays = pair_utility.form_pair_arrays(mol, 1e-14)
    pairs = pair_utility.form_primitive_pair_data(mol, arrays)
    aux_full = pair_utility.form_auxiliary_pair_data(mol)
    t_pairs = time.perf_counter() - t0
    aux = {}
    n_aux = 0
    for lc, pk in aux_full.items():
        sel = pk.i_atom < n_slice_atoms
        sub = copy.copy(pk)
        sub.A_a = np.ascontiguousarray(pk.A_a[sel])
        sub.coefficient = np.ascontiguousarray(pk.coefficient[sel])
        sub.i_ao_start = np.ascontiguousarray(pk.i_ao_start[sel])
        sub.i_atom = np.ascontiguousarray(pk.i_atom[sel])
        aux[lc] = sub
        n_aux = max(n_aux, int(sub.i_ao_start.max()) + 2 * lc + 1)
    mol_s = copy.copy(mol)
    mol_s.n_aux = n_aux
    return mol, mol_s, arrays, pairs, aux, t_pairs


def class_times(plan, blocks, buf, mixed, stream):
    """Per (la, lb, lc): summed HIP-event time (s) of its launches over one step, every launch on the
    launch stream (one class at a time, so durations do not overlap)."""
    import torch
    ev = []
    for b in blocks:
        for ai, ac in enumerate(plan.aux):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            plan.run_block(*b, buf, mixed=mixed, aux_classes=[ai])
            e1.record(stream)
            pc = plan.pairs[b[0]]
            ev.append(((pc.la, pc.lb, ac.lc), e0, e1))
    torch.cuda.synchronize()
    out = {}
    for key, e0, e1 in ev:
        out[key] = out.get(key, 0.0) + e0.elapsed_time(e1) * 1e-3
    return out


def cpu_baseline(mol, pairs, aux, gpu_cols, blocks_ao, max_seconds=CPU_SECONDS):
    """The reference C++ (or the oracle's restatement if the reference build is absent) on whole
    auxiliary shells of the slice against every pair, one shell per angular class in a fixed order until
    max_seconds have passed.  Returns the baseline record and the parity of the GPU columns
    (gpu_cols[name][lc][start] -> list of per-block (rows, nk) arrays) against it."""
    from tests.helpers import REF_SO, ORACLE_SO, dp, ip
    kind, path, fn = ("reference", REF_SO, "three_center") if REF_SO.exists() else ("port", ORACLE_SO, "oracle_three_center")
    cores = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
    lib = ctypes.CDLL(str(path))
    n_ao = mol.n_ao
    n_packed = n_ao * (n_ao + 1) // 2
    total_t, total_int, shells = 0.0, 0, []
    err = {name: 0.0 for name in gpu_cols}
    for lc, start in sample_shells(aux):
        pk = aux[lc]
        sel = pk.i_ao_start == start
        A = np.ascontiguousarray(pk.A_a[sel])
        c = np.ascontiguousarray(pk.coefficient[sel])
        st = np.zeros(int(sel.sum()), dtype=np.int32)
        nk = 2 * lc + 1
        J = np.zeros((nk, n_ao, n_ao))
        t0 = time.perf_counter()
        for (i_L, j_L), pd in pairs.items():
            rc = getattr(lib, fn)(i_L, j_L, lc, dp(pd.P_p), dp(pd.A_a), dp(pd.B_b), dp(pd.coefficient),
                                  ip(pd.i_ao_start), ip(pd.j_ao_start), pd.P_p.shape[0], dp(A), dp(c), ip(st),
                                  A.shape[0], dp(J), n_ao, nk, ctypes.c_bool(True), ctypes.c_double(0.0))
            assert rc == 0
        total_t += time.perf_counter() - t0
        total_int += nk * n_packed
        shells.append((int(lc), int(start)))
        for name, cols in gpu_cols.items():
            for bi, (I, Jx) in enumerate(blocks_ao):
                err[name] = max(err[name], float(np.abs(cols[(lc, start)][bi] - J[:, I, Jx].T).max()))
        if total_t > max_seconds:
            break
    n_pp = sum(len(p.P_p) for p in pairs.values())
    rec = {"value": total_int / total_t, "unit": "3c integrals/s", "cores": cores, "kind": kind,
           "sample": f"{len(shells)} whole auxiliary shells (l, first function) {shells} of the slice x all {n_pp} "
                     "primitive pairs, dense-equivalent integrals", "seconds": round(total_t, 3)}
    return rec, err


def sample_shells(aux):
    """(l, first function) of the CPU-baseline sample: the first shell of atom 1 for l = 4, 2, 0, 3, 1."""
    out = []
    for lc in (4, 2, 0, 3, 1):
        if lc in aux:
            pk = aux[lc]
            starts = np.unique(pk.i_ao_start[pk.i_atom == 1])
            if len(starts):
                out.append((lc, int(starts[0])))
    return out


COPY_PROBE_BYTES = 256 << 20    # each of the copy kernel's source and destination
COPY_PROBE_RUNS = 8
# three-centre kernels of a plain FP64 pass: vector (j3c_) or matrix-core (j3cm_) chunk kernels, direct or
# job-table forms; the float32 kernels carry an _f32 suffix and do not match
J3C_F64_KERNEL = re.compile(r"^j3cm?_(\d)(\d)(\d)_\d+_\d+(_jobs)?$")
# float32-tier kernels of a mixed pass: vector (_f32 suffix) or bf16 matrix-core (j3cb_)
J3C_F32_KERNEL = re.compile(r"^(?:j3c_(\d)(\d)(\d)_\d+_\d+_f32|j3cb_(\d)(\d)(\d)_\d+_\d+)(_jobs)?$")


def copy_probe():
    """COPY_PROBE_RUNS launches of the library's plain copy kernel over COPY_PROBE_BYTES: the known-size
    dispatches whose FETCH_SIZE / WRITE_SIZE readings calibrate the counter summaries."""
    import torch
    from electron_integral_playground_amd import library
    from electron_integral_playground_amd.engine import _stream_handle
    lib = library()
    n = COPY_PROBE_BYTES // 8
    a = torch.arange(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(COPY_PROBE_RUNS):
        rc = lib.eip_copy_probe(ctypes.c_void_p(a.data_ptr()), ctypes.c_void_p(b.data_ptr()),
                                ctypes.c_longlong(n), _stream_handle(None))
        assert rc == 0, rc
    torch.cuda.synchronize()
    assert torch.equal(a, b)
    del a, b


def profile_plain_leg(args):
    """The headline's plain FP64 steps alone (W warmup + K timed), then one per-class HIP-event pass on
    the launch stream: W + K + 1 plain passes over every class, the command the kernel-trace and counter
    summaries are made from.  Returns per-class event times, model flops and algorithmic bytes."""
    import torch
    from electron_integral_playground_amd.engine import ThreeCenterPlan, three_center_flops_per_triplet
    mol, mol_s, arrays, pairs, aux, _ = slice_workload(args.slice_atoms)
    plan = ThreeCenterPlan(mol_s, pairs, aux)
    blocks = plan.row_blocks(BLOCK_BYTES)
    mixed = bool(args.profile_mixed)
    if mixed:
        stats = plan.setup_mixed(arrays, None, threshold=args.threshold, estimate="envelope")
        log(f"profile-mixed: split {stats}")
    plan.prepare_blocks(blocks, mixed=mixed)
    rows = max(plan.block_rows(*b) for b in blocks)
    A = torch.empty((rows, mol_s.n_aux), dtype=torch.float64, device="cuda")
    stream = torch.cuda.current_stream()
    log(f"profile-plain: {len(blocks)} blocks per step, {max(1, args.warmup)} warmup + {args.steps} timed steps + "
        "1 event pass")
    for _ in range(max(1, args.warmup)):
        for b in blocks:
            plan.run_block(*b, A, mixed=mixed)
    torch.cuda.synchronize()
    log("profile-plain: warmup done")
    t0 = time.perf_counter()
    for _ in range(args.steps):
        for b in blocks:
            plan.run_block(*b, A, mixed=mixed)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    log(f"profile-plain: {dt * 1e3:.1f} ms/step")
    times = class_times(plan, blocks, A, mixed, stream)
    trip = plan.primitive_triplets()
    out_bytes, in_bytes = {}, {}
    for pi, sp0, sp1 in blocks:
        pc = plan.pairs[pi]
        for ai, ac in enumerate(plan.aux):
            k = (pc.la, pc.lb, ac.lc)
            out_bytes[k] = out_bytes.get(k, 0) + plan.block_rows(pi, sp0, sp1) * ac.n_ks * (2 * ac.lc + 1) * 8
            E = plan.kernel_E(pi, ai)
            in_bytes[k] = in_bytes.get(k, 0) + E.numel() * E.element_size() * (sp1 - sp0) // max(1, pc.n_sp)
    classes = {"%d%d%d" % k: {"event_s_per_pass": t, "model_flops": trip[k] * three_center_flops_per_triplet(*k),
                              "store_bytes": out_bytes[k], "pair_matrix_bytes": in_bytes[k]}
               for k, t in times.items()}
    del A, plan
    torch.cuda.empty_cache()
    return {"profile_plain": {"passes": max(1, args.warmup) + args.steps + 1, "ms_per_step": dt * 1e3,
                              "blocks_per_step": len(blocks), "mode": "mixed" if mixed else "plain",
                              "threshold": args.threshold if mixed else None, "classes": classes}}


def _csv_rows(path):
    with open(path, newline="") as f:
        return list(csv.DictReader(f))


def _class_of(kernel, tier="f64"):
    if tier == "f64":
        m = J3C_F64_KERNEL.match(kernel.strip())
        return None if m is None else "%s%s%s" % m.groups()[:3]
    m = J3C_F32_KERNEL.match(kernel.strip())
    if m is None:
        return None
    g = m.groups()
    return "%s%s%s" % (g[:3] if g[0] is not None else g[3:6])


def trace_summary(stats_csv, run_json):
    """Markdown table: per class, dispatches, total and per-pass kernel time from a kernel_stats.csv of a
    --profile-plain run, beside that run's own HIP-event time per pass."""
    run = json.loads(pathlib.Path(run_json).read_text())["profile_plain"]
    passes = run["passes"]
    agg = {}
    for r in _csv_rows(stats_csv):
        name = r.get("Name") or r.get("Kernel_Name") or ""
        for tier in ("f64", "f32"):
            cls = _class_of(name, tier)
            if cls is None:
                continue
            a = agg.setdefault(cls, {"f64": [0, 0.0], "f32": [0, 0.0]})[tier]
            a[0] += int(r["Calls"])
            a[1] += float(r["TotalDurationNs"]) * 1e-9
    if run.get("mode") == "mixed":
        lines = [f"mixed passes per class: {passes} (threshold {run['threshold']:g}); bench ms/step "
                 f"{run['ms_per_step']:.1f}", "",
                 "| class | float64 dispatches | float64 s per pass | float32 dispatches | float32 s per pass | "
                 "s per pass (trace) | s per pass (HIP events) | trace / events |",
                 "|---|---|---|---|---|---|---|---|"]
        for cls, t in sorted(agg.items(), key=lambda kv: -(kv[1]["f64"][1] + kv[1]["f32"][1])):
            c = run["classes"].get(cls)
            ev = c["event_s_per_pass"] if c else float("nan")
            per = (t["f64"][1] + t["f32"][1]) / passes
            lines.append(f"| j3c_{cls} | {t['f64'][0]} | {t['f64'][1] / passes:.5f} | {t['f32'][0]} | "
                         f"{t['f32'][1] / passes:.5f} | {per:.5f} | {ev:.5f} | {per / ev:.3f} |")
        return "\n".join(lines)
    lines = [f"plain passes per class: {passes}; bench ms/step {run['ms_per_step']:.1f}", "",
             "| class | dispatches | kernel s (trace) | s per pass (trace) | s per pass (HIP events) | "
             "trace / events | model TFLOP/s (trace) | frac of 78.6 |",
             "|---|---|---|---|---|---|---|---|"]
    for cls, t in sorted(agg.items(), key=lambda kv: -kv[1]["f64"][1]):
        calls, tot = t["f64"]
        if not calls:
            continue
        c = run["classes"].get(cls)
        ev = c["event_s_per_pass"] if c else float("nan")
        per = tot / passes
        tf = c["model_flops"] / per / 1e12 if c else float("nan")
        lines.append(f"| j3c_{cls} | {calls} | {tot:.4f} | {per:.5f} | {ev:.5f} | {per / ev:.3f} | {tf:.2f} | "
                     f"{tf / 78.6:.3f} |")
    return "\n".join(lines)


SQ_PASS_A = ("SQ_WAVE_CYCLES", "SQ_BUSY_CYCLES", "SQ_WAIT_ANY", "SQ_WAIT_INST_ANY", "SQ_ACTIVE_INST_VALU",
             "SQ_INSTS_VALU", "SQ_INSTS_SMEM", "SQ_INSTS_LDS")
SQ_PASS_B = ("SQ_LDS_BANK_CONFLICT", "SQ_WAIT_INST_LDS", "SQ_ACTIVE_INST_LDS", "SQ_INSTS_VALU_FMA_F64",
             "SQ_INSTS_VALU_MUL_F64", "SQ_INSTS_VALU_ADD_F64", "SQ_ACTIVE_INST_SCA", "SQ_INSTS_SALU")


def sq_summary(csv_a, csv_b, top=14):
    """Markdown table: per float64 class (its chunk kernels' dispatches summed), from two rocprofv3 --pmc
    passes of the same run (counters SQ_PASS_A, then SQ_PASS_B): the fraction of wave cycles spent waiting
    on s_waitcnt (SQ_WAIT_ANY), stalled on dependencies (SQ_WAIT_INST_ANY) and issuing VALU
    (SQ_ACTIVE_INST_VALU), the VALU and LDS instructions per wave cycle, and the share of VALU
    instructions that are float64 FMAs.  Classes ordered by wave cycles, the top ``top`` shown."""
    tot = {}
    for path in (csv_a, csv_b):
        for r in _csv_rows(path):
            cls = _class_of(r.get("Kernel_Name") or "")
            if cls is None:
                continue
            d = tot.setdefault(cls, {})
            d[r["Counter_Name"]] = d.get(r["Counter_Name"], 0.0) + float(r["Counter_Value"])
            if r["Counter_Name"] == "SQ_WAVE_CYCLES":
                d["dispatches"] = d.get("dispatches", 0) + 1
    lines = ["| class | dispatches | wave cycles | wait (s_waitcnt) | dependency stall | VALU active | "
             "VALU instr / wave cycle | FP64 FMA share of VALU | LDS instr / VALU instr | LDS bank conflict / LDS active |",
             "|---|---|---|---|---|---|---|---|---|---|"]
    for cls, d in sorted(tot.items(), key=lambda kv: -kv[1].get("SQ_WAVE_CYCLES", 0))[:top]:
        wc = d.get("SQ_WAVE_CYCLES", 0) or 1
        valu = d.get("SQ_INSTS_VALU", 0) or 1
        lds_act = d.get("SQ_ACTIVE_INST_LDS", 0) or 1
        lines.append(f"| j3c_{cls} | {d.get('dispatches', 0)} | {wc:.3e} | {d.get('SQ_WAIT_ANY', 0) / wc:.2f} | "
                     f"{d.get('SQ_WAIT_INST_ANY', 0) / wc:.2f} | {d.get('SQ_ACTIVE_INST_VALU', 0) / wc:.2f} | "
                     f"{d.get('SQ_INSTS_VALU', 0) / wc:.3f} | {d.get('SQ_INSTS_VALU_FMA_F64', 0) / valu:.2f} | "
                     f"{d.get('SQ_INSTS_LDS', 0) / valu:.3f} | {d.get('SQ_LDS_BANK_CONFLICT', 0) / lds_act:.3f} |")
    return "\n".join(lines)


def pmc_summary(fetch_csv, write_csv, run_json):
    """Markdown table: per class, FETCH_SIZE and WRITE_SIZE per pass (KiB, summed over the class's
    dispatches), each scaled by the factor the copy probe's dispatches give (known bytes / reading), beside
    the algorithmic bytes; rows ``| j3c_<cls> | ... | calibrated bytes | ...``."""
    run = json.loads(pathlib.Path(run_json).read_text())["profile_plain"]
    passes = run["passes"]

    def read(path, counter):
        per_cls, probe = {}, []
        for r in _csv_rows(path):
            if r.get("Counter_Name", counter) != counter:
                continue
            name = r.get("Kernel_Name") or r.get("Name") or ""
            v = float(r["Counter_Value"])
            if "copy_probe_kernel" in name:
                probe.append(v)
            cls = _class_of(name)
            if cls is not None:
                per_cls[cls] = per_cls.get(cls, 0.0) + v
        return per_cls, probe

    F, pf = read(fetch_csv, "FETCH_SIZE")
    W, pw = read(write_csv, "WRITE_SIZE")
    known_kib = COPY_PROBE_BYTES / 1024
    fF = known_kib / (sum(pf) / len(pf))
    fW = known_kib / (sum(pw) / len(pw))
    lines = [f"calibration: copy_probe_kernel ({len(pf)} + {len(pw)} dispatches, {COPY_PROBE_BYTES >> 20} MiB read "
             f"+ {COPY_PROBE_BYTES >> 20} MiB written each): FETCH_SIZE {sum(pf) / len(pf):.1f} KiB -> factor "
             f"{fF:.4f}; WRITE_SIZE {sum(pw) / len(pw):.1f} KiB -> factor {fW:.4f}",
             f"plain passes per class: {passes}; values are per pass (one step's launches of the class)", "",
             "| class | FETCH_SIZE KiB/pass | WRITE_SIZE KiB/pass | HBM bytes/pass (calibrated) | algorithmic bytes/pass "
             "(stores + pair matrix once) | traffic / algorithmic |",
             "|---|---|---|---|---|---|"]
    rows = []
    for cls in set(F) | set(W):
        fk, wk = F.get(cls, 0.0) / passes, W.get(cls, 0.0) / passes
        hbm = (fk * fF + wk * fW) * 1024
        c = run["classes"].get(cls)
        alg = c["store_bytes"] + c["pair_matrix_bytes"] if c else float("nan")
        rows.append((hbm, f"| j3c_{cls} | {fk:.1f} | {wk:.1f} | {hbm:.4e} | {alg:.4e} | {hbm / alg:.3f} |"))
    lines += [r for _, r in sorted(rows, reverse=True)]
    return "\n".join(lines)


def headline_leg(args):
    """Plain and mixed-precision steps over the 200-atom workload's auxiliary-block share."""
    import torch
    from electron_integral_playground_amd.engine import ThreeCenterPlan, three_center_flops_per_triplet
    mol, mol_s, arrays, pairs, aux, t_pairs = slice_workload(args.slice_atoms)
    plan = ThreeCenterPlan(mol_s, pairs, aux)
    if args.tune:
        from electron_integral_playground_amd.engine import autotune
        if not args.tune_f32_only:
            autotune(plan, save=args.tune, log=log, block_bytes=4e9)
        autotune(plan, save=args.tune, log=log, block_bytes=4e9, use_f32=True)
    J2c = plan.two_center()
    q = q_aux_from_metric(plan, J2c)
    del J2c
    stats = plan.setup_mixed(arrays, q, threshold=args.threshold, granularity=args.mixed_granularity,
                             estimate=args.mixed_estimate if args.mixed_granularity == "item" else "schwarz")
    blocks = plan.row_blocks(BLOCK_BYTES)
    plan.prepare_blocks(blocks, mixed=True)
    rows = max(plan.block_rows(*b) for b in blocks)
    A = torch.empty((rows, mol_s.n_aux), dtype=torch.float64, device="cuda")
    A2 = torch.empty_like(A) if not args.single_stream else None
    stream = torch.cuda.current_stream()
    lane_sets = [[torch.cuda.Stream() for _ in range(6)] for _ in range(2)]
    n_pp = sum(len(a) for a in arrays.values())
    log(f"headline: n_ao={mol.n_ao} n_aux(slice)={mol_s.n_aux} of {mol.n_aux} primitive pairs={n_pp} "
        f"blocks={len(blocks)} host pair list {t_pairs:.1f}s; split {stats}")

    def step(mixed):
        """One pass over every block.  Blocks alternate between two 40 GB buffers; each block's launches
        (auxiliary classes, and float64/float32 tiers) are spread over its buffer's 6 streams, and a block
        waits only for the previous block of the same buffer (two blocks back), so one block's tail
        overlaps the next block.  --single-stream: one buffer, every launch in order on one stream."""
        if args.single_stream:
            for b in blocks:
                plan.run_block(*b, A, mixed=mixed)
            return
        last = [[], []]
        for s_ in lane_sets[0] + lane_sets[1]:
            s_.wait_stream(stream)
        for k, b in enumerate(blocks):
            lanes = lane_sets[k % 2]
            for s_ in lanes:
                for e in last[k % 2]:
                    s_.wait_event(e)
            plan.run_block(*b, (A, A2)[k % 2], mixed=mixed, lanes=lanes)
            last[k % 2] = [s_.record_event() for s_ in lanes]
        for s_ in lane_sets[0] + lane_sets[1]:
            stream.wait_stream(s_)

    for _ in range(max(1, args.warmup)):
        step(False)
        step(True)
    torch.cuda.synchronize()
    log("headline warmup done")
    if args.tune_step:
        # variants chosen by the step time itself: float64 on the plain step, float32 on the mixed step,
        # the classes in order of their share of the step (per-class event pass), the top ones only
        from electron_integral_playground_amd.engine import autotune_in_step

        def synced(mixed):
            def run():
                step(mixed)
                torch.cuda.synchronize()
            return run
        for mixed in (False, True):
            t = class_times(plan, blocks, A, mixed, stream)
            top = [k for k, v in sorted(t.items(), key=lambda kv: -kv[1]) if v >= 0.01 * sum(t.values())]
            log(f"step-tune {'float32' if mixed else 'float64'}: {len(top)} classes")
            autotune_in_step(plan.lib, top, synced(mixed), use_f32=mixed, save=args.tune_step, log=log)
    # step boundaries as events on the launch stream (every lane joins it at the end of a step), so the
    # per-step durations come without synchronising between steps
    marks = {False: [], True: []}
    t0 = time.perf_counter()
    for mixed in (False, True):
        if mixed:
            torch.cuda.synchronize()
            dt = (time.perf_counter() - t0) / args.steps
            t0 = time.perf_counter()
        marks[mixed].append(stream.record_event(torch.cuda.Event(enable_timing=True)))
        for _ in range(args.steps):
            step(mixed)
            marks[mixed].append(stream.record_event(torch.cuda.Event(enable_timing=True)))
    torch.cuda.synchronize()
    dt_mixed = (time.perf_counter() - t0) / args.steps
    each = {m: [a.elapsed_time(b) for a, b in zip(ev[:-1], ev[1:])] for m, ev in marks.items()}
    log("headline steps (ms): plain " + " ".join(f"{x:.0f}" for x in each[False]) + "; mixed "
        + " ".join(f"{x:.0f}" for x in each[True]))
    n_int = mol_s.n_aux * mol.n_ao * (mol.n_ao + 1) // 2
    log(f"headline step plain {dt * 1e3:.1f} ms -> {n_int / dt:.4e} integrals/s; mixed {dt_mixed * 1e3:.1f} ms "
        f"-> {n_int / dt_mixed:.4e} ({dt / dt_mixed:.3f}x)")

    # per-class kernel time (one class at a time on the launch stream) and model flops
    trip = plan.primitive_triplets()
    times = class_times(plan, blocks, A, False, stream)
    times32 = class_times(plan, blocks, A, True, stream)
    per_class = sorted(((k, t, trip[k] * three_center_flops_per_triplet(*k)) for k, t in times.items()),
                       key=lambda r: -r[1])
    for k, t, fl in per_class[:10]:
        log(f"  class {k}: {t * 1e3:.2f} ms/step  {fl / t / 1e12:.2f} TFLOP/s (model)  mixed {times32[k] * 1e3:.2f} ms")
    key, t, fl = per_class[0]
    roofline = {"bound": "fp64-valu", "kernel": "j3c_%d%d%d (%d component-chunk kernels, all blocks of one step)"
                % (key + (plan.lib.eip_three_center_nchunks(*key),)),
                "achieved": fl / t / 1e12, "peak": PEAK_FP64 / 1e12, "unit": "TFLOP/s", "frac": fl / t / PEAK_FP64,
                "traffic": measured_traffic(PMC_SUMMARY, key),
                "traffic_unit": "HBM bytes per step for this class (rocprofv3 FETCH_SIZE x 1.9998 + WRITE_SIZE x 1.0000, "
                                "factors from the library's plain copy kernel)",
                "traffic_algorithmic": measured_traffic(PMC_SUMMARY, key, column=4),
                "algorithmic_flops_per_step": fl, "class_seconds_per_step": t}

    # error pass over the whole slice (mixed vs plain on the device), keeping the CPU sample's columns
    B = torch.empty_like(A)
    sample = sample_shells(aux)
    cols = np.concatenate([np.arange(s, s + 2 * lc + 1) for lc, s in sample])
    cols_t = torch.as_tensor(cols, device="cuda")
    max_abs = torch.zeros((), dtype=torch.float64, device="cuda")
    sum_sq = torch.zeros((), dtype=torch.float64, device="cuda")
    max_ref = torch.zeros((), dtype=torch.float64, device="cuda")
    got = {"plain": [], "mixed": []}
    n_el = 0
    if plan.checked:
        plan.checked_tally = {}
    for b in blocks:
        n = plan.block_rows(*b)
        plan.run_block(*b, A)
        plan.run_block(*b, B, mixed=True)
        for r in range(0, n, 8192):
            r1 = min(n, r + 8192)
            d = B[r:r1] - A[r:r1]
            max_abs = torch.maximum(max_abs, d.abs().max())
            sum_sq += (d * d).sum()
            max_ref = torch.maximum(max_ref, A[r:r1].abs().max())
        got["plain"].append(A[:n].index_select(1, cols_t).cpu().numpy())
        got["mixed"].append(B[:n].index_select(1, cols_t).cpu().numpy())
        n_el += n * mol_s.n_aux
    torch.cuda.synchronize()
    if plan.checked:
        tally = {k: (int(a.item()), int(b.item())) for k, (a, b) in plan.checked_tally.items()}
        plan.checked_tally = None
        stats["f64_triplets"] = sum(b for _, b in tally.values())
        stats["f32_triplets"] = sum(a - b for a, b in tally.values())
        log("checked split: float64 share per class " +
            ", ".join(f"{''.join(map(str, k))} {b / max(a, 1):.3f}" for k, (a, b) in sorted(tally.items())))
    del A, A2, B
    torch.cuda.empty_cache()
    gpu_cols = {name: {} for name in got}
    c0 = 0
    for lc, s in sample:
        nk = 2 * lc + 1
        for name in got:
            gpu_cols[name][(lc, s)] = [g[:, c0:c0 + nk] for g in got[name]]
        c0 += nk
    blocks_ao = []
    for pi, sp0, sp1 in blocks:
        pc = plan.pairs[pi]
        ni, nj = 2 * pc.la + 1, 2 * pc.lb + 1
        sp = np.arange(sp0, sp1)
        I = np.broadcast_to(pc.sp_i[sp].astype(np.int64)[:, None, None] + np.arange(ni)[None, :, None], (len(sp), ni, nj))
        J = np.broadcast_to(pc.sp_j[sp].astype(np.int64)[:, None, None] + np.arange(nj)[None, None, :], (len(sp), ni, nj))
        blocks_ao.append((I.ravel(), J.ravel()))
    mixed = {"value": n_int / dt_mixed, "unit": "3c integrals/s", "ms_per_step": dt_mixed * 1e3,
             "mixed_over_plain": dt / dt_mixed, "threshold": args.threshold, "granularity": args.mixed_granularity,
             "estimate": stats.get("estimate", "schwarz"), "estimate_seconds": stats.get("estimate_seconds"),
             "f32_triplet_fraction": stats["f32_triplets"] / (stats["f32_triplets"] + stats["f64_triplets"]),
             "max_abs_err_vs_plain_fp64": max_abs.item(), "rms_err_vs_plain_fp64": (sum_sq.item() / n_el) ** 0.5,
             "max_abs_value": max_ref.item(), "compared_elements": n_el,
             "class_seconds_per_step": {"%d%d%d" % k: v for k, v in times32.items()},
             "ms_each_step": [round(x, 2) for x in each[True]]}
    cpu = None
    if not args.no_cpu:
        cpu, err = cpu_baseline(mol_s, pairs, aux, gpu_cols, blocks_ao)
        mixed["max_abs_err_vs_cpu_reference_sample"] = err["mixed"]
        log(f"cpu baseline {cpu}; sample max |GPU - CPU| plain {err['plain']:.3e} mixed {err['mixed']:.3e}")
        cpu_err = err["plain"]
    else:
        cpu_err = None
    res = {"value": n_int / dt, "ms_per_step": dt * 1e3, "n_ao": mol.n_ao, "n_aux_slice": mol_s.n_aux,
           "n_aux_full": mol.n_aux, "primitive_pairs": n_pp, "blocks_per_step": len(blocks),
           "roofline": roofline, "cpu_baseline": cpu, "mixed": mixed,
           "max_abs_err_vs_cpu_reference_sample": cpu_err,
           "class_seconds_per_step": {"%d%d%d" % k: v for k, v, _ in per_class},
           "ms_each_step": [round(x, 2) for x in each[False]]}
    del plan
    torch.cuda.empty_cache()
    return res


def dropin_leg(repeats=3):
    """The reference-style drop-in loop -- molecular_integral.three_center / two_center over every class
    through the C ABI, host arrays in and the dense reference layout out -- on the 12-atom workload,
    timed end to end (best of ``repeats`` after one warm-up) beside the compiled reference (oracle/_ref,
    OpenMP) called the same way, and their largest difference."""
    from electron_integral_playground_amd import generators, pair_utility
    from electron_integral_playground_amd.molecular_integral import three_center, two_center
    from tests.helpers import REF_SO, ORACLE_SO, run_three_center_oracle, run_two_center_oracle
    mol = generators.make_molecule("small")
    pairs = pair_utility.form_pair_data(mol, 1e-14)
    aux = pair_utility.form_auxiliary_pair_data(mol)
    kind, path = ("reference", REF_SO) if REF_SO.exists() else ("port", ORACLE_SO)
    lib = ctypes.CDLL(str(path))
    f3 = "three_center" if kind == "reference" else "oracle_three_center"
    f2 = "two_center" if kind == "reference" else "oracle_two_center"

    def best(fn):
        fn()
        ts = []
        for _ in range(repeats):
            t0 = time.perf_counter()
            out = fn()
            ts.append(time.perf_counter() - t0)
        return min(ts), out

    t3, J3 = best(lambda: three_center(pairs, aux, mol))
    r3, R3 = best(lambda: run_three_center_oracle(lib, pairs, aux, mol, fn=f3))
    t2, J2 = best(lambda: two_center(aux, mol))
    r2, R2 = best(lambda: run_two_center_oracle(lib, aux, mol, fn=f2))
    n3 = mol.n_aux * mol.n_ao * (mol.n_ao + 1) // 2
    res = {"workload": "12 atoms (small), n_ao %d, n_aux %d; dense (n_aux, n_ao, n_ao) host output" % (mol.n_ao, mol.n_aux),
           "three_center_s": t3, "three_center_reference_s": r3, "speedup_three_center": r3 / t3,
           "three_center_integrals_per_s": n3 / t3, "two_center_s": t2, "two_center_reference_s": r2,
           "max_abs_diff_three_center": float(np.abs(J3 - R3).max()), "max_abs_diff_two_center": float(np.abs(J2 - R2).max()),
           "reference_kind": kind, "reference_threads": int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)}
    log(f"drop-in loop: {res}")
    return res


def plain80_leg(args, omega=0.0):
    """The 80-atom workload: every (ij|P) into the packed device output plus (P|Q), per step (range-
    separated erf(omega r)/r operator when omega > 0)."""
    import torch
    from electron_integral_playground_amd import generators, pair_utility
    from electron_integral_playground_amd.engine import ThreeCenterPlan, three_center_flops_per_triplet
    mol = generators.make_molecule("plain")
    pairs = pair_utility.form_pair_data(mol, 1e-14)
    aux = pair_utility.form_auxiliary_pair_data(mol)
    plan = ThreeCenterPlan(mol, pairs, aux, omega=omega)
    out = plan.new_output()
    J2c = torch.zeros((mol.n_aux, mol.n_aux), dtype=torch.float64, device="cuda")
    stream = torch.cuda.current_stream()
    for _ in range(max(1, args.warmup)):
        plan.run(out, concurrent=not args.single_stream, metric_out=J2c)
    torch.cuda.synchronize()
    ev = []
    for pi, pc in enumerate(plan.pairs):
        for ai, ac in enumerate(plan.aux):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            plan.launch_class(pi, ai, out)
            e1.record(stream)
            ev.append(((pc.la, pc.lb, ac.lc), pc.n_pp * ac.n_prim, e0, e1))
    torch.cuda.synchronize()
    per_class = sorted(((k, e0.elapsed_time(e1) * 1e-3, n * three_center_flops_per_triplet(*k)) for k, n, e0, e1 in ev),
                       key=lambda r: -r[1])
    t0 = time.perf_counter()
    for _ in range(args.steps):
        plan.run(out, concurrent=not args.single_stream, metric_out=J2c)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    n_int = plan.n_integrals()
    if args.profile_plain:
        classes = {}
        for pi, pc in enumerate(plan.pairs):
            ni, nj = 2 * pc.la + 1, 2 * pc.lb + 1
            diag = int((np.asarray(pc.sp_i) == np.asarray(pc.sp_j)).sum())
            rows = (pc.n_sp - diag) * ni * nj + diag * ni * (ni + 1) // 2
            for ai, ac in enumerate(plan.aux):
                k = (pc.la, pc.lb, ac.lc)
                E = plan.kernel_E(pi, ai)
                classes["%d%d%d" % k] = {"store_bytes": rows * ac.n_ks * (2 * ac.lc + 1) * 8,
                                         "pair_matrix_bytes": E.numel() * E.element_size()}
        for k, t, fl in per_class:
            classes["%d%d%d" % k].update(event_s_per_pass=t, model_flops=fl)
        del out, plan, J2c
        torch.cuda.empty_cache()
        return {"profile_plain": {"passes": max(1, args.warmup) + 1 + args.steps, "ms_per_step": dt * 1e3,
                                  "classes": classes}}
    key, t, fl = per_class[0]
    log(f"plain80 step {dt * 1e3:.2f} ms -> {n_int / dt:.4e} integrals/s; dominant class {key} {t * 1e3:.3f} ms "
        f"{fl / t / 1e12:.2f} TFLOP/s (model)")
    coulomb = None
    if omega == 0.0:
        coulomb = coulomb_fit_timing(plan, out, J2c, args.steps)
    res = {"workload": "80 atoms, 14 A box, 3s2p1d (1-6 prim), aux s-f 8 exp/l ratio 2.2" +
                       (f"; erf(omega r)/r, omega {omega}" if omega else ""),
           "n_ao": mol.n_ao, "n_aux": mol.n_aux, "value": n_int / dt, "unit": "3c integrals/s",
           "ms_per_step": dt * 1e3,
           "output": "packed (n_ao(n_ao+1)/2, n_aux) float64 (ij|P) + dense (n_aux, n_aux) (P|Q) on the device, every step",
           "roofline": {"bound": "fp64-valu", "kernel": "j3c_%d%d%d" % key, "achieved": fl / t / 1e12,
                        "peak": PEAK_FP64 / 1e12, "unit": "TFLOP/s", "frac": fl / t / PEAK_FP64,
                        "traffic": measured_traffic(PMC_SUMMARY_PLAIN80, key)},
           "top_classes_ms": {"%d%d%d" % k: t * 1e3 for k, t, _ in per_class[:8]},
           "coulomb_fit": coulomb}
    del out, plan, J2c
    torch.cuda.empty_cache()
    return res


def coulomb_fit_timing(plan, packed, J2c, steps):
    """The density-fitted Coulomb matrix consumed from the packed (ij|P) on the device (engine.CoulombFit):
    gamma = (P|ij) D, (P|Q) d = gamma through the float64 Cholesky factor, J = (ij|P) d, for a random
    symmetric density; the factorisation (Cholesky and L^-1) is timed once, the contraction per step,
    beside the same fit through two triangular vector solves per step."""
    import torch
    from electron_integral_playground_amd.engine import CoulombFit
    rng = np.random.default_rng(9)
    c = rng.standard_normal((plan.n_ao, plan.n_ao // 4))
    D = torch.as_tensor(c @ c.T / plan.n_ao, device="cuda")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    fit = CoulombFit(plan, J2c=J2c)
    torch.cuda.synchronize()
    t_chol = time.perf_counter() - t0

    def timed(f):
        J = f.from_packed(packed, D)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            J = f.from_packed(packed, D)
        torch.cuda.synchronize()
        return J, (time.perf_counter() - t0) / steps

    J, t_fit = timed(fit)
    J2, t_trsv = timed(CoulombFit(plan, L=fit.L, inverse=False))
    diff = float((J - J2).abs().max() / J2.abs().max())
    res = {"what": "J = (ij|P) (P|Q)^-1 (P|kl) D_kl from the packed output on the device (two streaming "
                   "passes over it, eip_weighted_column_sums and eip_row_dots, and a Cholesky solve)",
           "factorisation_s": t_chol, "ms_per_fit": t_fit * 1e3,
           "ms_per_fit_triangular_solves": t_trsv * 1e3, "rel_diff_inverse_vs_triangular_solves": diff,
           "packed_GB_read_per_fit": 2 * packed.numel() * 8 / 1e9,
           "GB_per_s": 2 * packed.numel() * 8 / t_fit / 1e9, "J_symmetric": bool(torch.equal(J, J.T))}
    log(f"coulomb fit: {res}")
    return res


def mixed_leg(threshold=1e-6, block_bytes=40e9, n_atom=None, tune=None, aux_order="natural"):
    """Mixed-precision workload (200 atoms, 20 A box, 4s3p2d1f, auxiliary s-g, n_ao 6000, n_aux 40000):
    the (ij|P) tensor (3.9 TB over the significant shell pairs) is streamed through two device buffers
    in blocks of shell pairs; each block is computed by the plain float64 kernels and by the mixed
    split, and the two are compared on the device (max abs, RMS).  Times are HIP events on the launch
    stream around each block's launches."""
    import torch
    from electron_integral_playground_amd import generators, pair_utility
    from electron_integral_playground_amd.engine import ThreeCenterPlan
    t0 = time.perf_counter()
    mol = generators.make_molecule("mixed", n_atom=n_atom)
    arrays = pair_utility.form_pair_arrays(mol, 1e-14)
    pairs = pair_utility.form_primitive_pair_data(mol, arrays)
    aux = pair_utility.form_auxiliary_pair_data(mol)
    t_pairs = time.perf_counter() - t0
    plan = ThreeCenterPlan(mol, pairs, aux)
    stream = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    J2c = torch.empty((mol.n_aux, mol.n_aux), dtype=torch.float64, device="cuda")
    plan.two_center(out=J2c.zero_())                           # warm-up
    J2c.zero_()
    e0.record(stream)
    plan.two_center(out=J2c)
    e1.record(stream)
    torch.cuda.synchronize()
    t_metric = e0.elapsed_time(e1) * 1e-3
    q_aux = q_aux_from_metric(plan, J2c)
    # the downstream factorisation of the metric, on the device in float64
    from electron_integral_playground_amd.engine import metric_cholesky
    chol = {}
    try:
        e0.record(stream)
        L = metric_cholesky(J2c)
        e1.record(stream)
        torch.cuda.synchronize()
        chol = {"metric_cholesky_s": e0.elapsed_time(e1) * 1e-3,
                "metric_cholesky_residual": float(((L @ L.T) - J2c).abs().max().item() / J2c.abs().max().item())}
        del L
    except ValueError as exc:
        chol = {"metric_cholesky": str(exc)}
    del J2c
    torch.cuda.empty_cache()
    if tune:
        from electron_integral_playground_amd.engine import autotune
        autotune(plan, save=tune, log=log, block_bytes=4e9)
        autotune(plan, save=tune, log=log, block_bytes=4e9, use_f32=True)
    stats = plan.setup_mixed(arrays, q_aux, threshold=threshold, aux_order=aux_order)
    t_setup = time.perf_counter() - t0
    log(f"mixed workload: n_ao={mol.n_ao} n_aux={mol.n_aux} primitive pairs={sum(len(p) for p in arrays.values())}"
        f" host pair list {t_pairs:.1f}s, setup {t_setup:.1f}s; split {stats}")
    blocks = plan.row_blocks(block_bytes)
    plan.prepare_blocks(blocks, mixed=True)
    max_rows = max(plan.block_rows(*b) for b in blocks)
    A = torch.empty((max_rows, mol.n_aux), dtype=torch.float64, device="cuda")
    B = torch.empty_like(A)
    # warm-up: one block of every class in both modes
    seen = set()
    for b in blocks:
        if b[0] not in seen:
            seen.add(b[0])
            plan.run_block(*b, A)
            plan.run_block(*b, B, mixed=True)
    torch.cuda.synchronize()
    t_plain = t_mixed = 0.0
    max_abs = torch.zeros((), dtype=torch.float64, device="cuda")
    sum_sq = torch.zeros((), dtype=torch.float64, device="cuda")
    max_ref = torch.zeros((), dtype=torch.float64, device="cuda")
    n_el = 0
    ev = []
    for bi, b in enumerate(blocks):
        rows = plan.block_rows(*b)
        a0, a1, a2 = (torch.cuda.Event(enable_timing=True) for _ in range(3))
        a0.record(stream)
        plan.run_block(*b, A)
        a1.record(stream)
        plan.run_block(*b, B, mixed=True)
        a2.record(stream)
        ev.append((a0, a1, a2))
        for r in range(0, rows, 4096):
            r1 = min(rows, r + 4096)
            d = B[r:r1] - A[r:r1]
            max_abs = torch.maximum(max_abs, d.abs().max())
            sum_sq += (d * d).sum()
            max_ref = torch.maximum(max_ref, A[r:r1].abs().max())
        n_el += rows * mol.n_aux
        if bi % 10 == 0:
            log(f"  block {bi}/{len(blocks)}")
    torch.cuda.synchronize()
    for a0, a1, a2 in ev:
        t_plain += a0.elapsed_time(a1) * 1e-3
        t_mixed += a1.elapsed_time(a2) * 1e-3
    n_int = mol.n_aux * mol.n_ao * (mol.n_ao + 1) // 2
    res = {"workload": f"mixed: {len(mol.elements)} atoms, 4s3p2d1f (1-6 prim), aux s-g 8 exp/l",
           "n_ao": mol.n_ao, "n_aux": mol.n_aux,
           "primitive_pairs": int(sum(len(p) for p in arrays.values())),
           "threshold": threshold, "aux_order": aux_order, "f32_triplet_fraction": stats["f32_triplets"] / (stats["f32_triplets"] + stats["f64_triplets"]),
           "plain_fp64_integrals_per_s": n_int / t_plain, "mixed_integrals_per_s": n_int / t_mixed,
           "mixed_over_plain": t_plain / t_mixed, "plain_s": t_plain, "mixed_s": t_mixed,
           "max_abs_err_vs_fp64": max_abs.item(), "rms_err_vs_fp64": (sum_sq.item() / n_el) ** 0.5,
           "max_abs_value": max_ref.item(), "metric_fp64_s": t_metric, **chol,
           "output": f"{n_el * 8 / 1e12:.2f} TB of significant-shell-pair rows, streamed in {len(blocks)} blocks"}
    log(f"mixed leg: {res}")
    return res


def serving_leg(n_jobs, rate, budget=200e9, mixed=True, part=0, latency_file=None, persistent=False):
    """Poisson stream of integral jobs (atom counts lognormal, median 40, clipped 10-250; 4s3p2d1f
    orbital and s-g auxiliary generators) through the continuous-batching server (each step one work
    queue per precision through the persistent tier kernels, or -- persistent=False -- one launch per
    class and precision)."""
    import torch
    from electron_integral_playground_amd import generators
    from electron_integral_playground_amd.serving import IntegralServer, run_poisson
    # part k of a longer stream: its own arrival/size seed and job seeds
    jobs = generators.serving_jobs(n_jobs=n_jobs, rate=rate, seed=generators.SEED_BASE + 5 + 1000 * part)
    jobs = [(t, n, seed + 100000 * part) for t, n, seed in jobs]
    # warm-up: one small job through a throw-away server (kernel loading, allocator)
    warm = IntegralServer(budget_bytes=budget, mixed=mixed, persistent=persistent)
    warm.submit(0, generators.serving_molecule(10, 1))
    warm.drain(timeout=300)
    warm.close()
    del warm
    torch.cuda.empty_cache()
    server = IntegralServer(budget_bytes=budget, mixed=mixed, persistent=persistent)
    span = run_poisson(server, jobs, generators.serving_molecule, log=log)
    server.close()
    done = [j for j in server.completed if j.error is None]
    lat = np.array([j.latency for j in done])
    n_int = sum(j.n_aux * j.n_ao * (j.n_ao + 1) // 2 for j in done)
    ph = [j.phases() for j in done if j.phases() is not None]
    phases = {k: {"p50": float(np.percentile([x[k] for x in ph], 50)),
                  "p90": float(np.percentile([x[k] for x in ph], 90)),
                  "p99": float(np.percentile([x[k] for x in ph], 99))} for k in (ph[0] if ph else {})}
    res = {"jobs": n_jobs, "completed": len(done), "errors": len(server.completed) - len(done),
           "latency_phases_s": phases,
           "rate_jobs_per_s": rate, "budget_GB": budget / 1e9, "mixed_precision": mixed,
           "step_kernels": "persistent tier kernels (one work queue per precision)" if persistent
                           else "one launch per class and precision",
           "p50_latency_s": float(np.percentile(lat, 50)), "p99_latency_s": float(np.percentile(lat, 99)),
           "mean_latency_s": float(lat.mean()), "max_latency_s": float(lat.max()),
           "span_s": span, "jobs_per_s": len(done) / span, "integrals_per_s": n_int / span,
           "steps": server.n_steps, "launches": server.n_launches,
           "median_atoms": float(np.median([n for _, n, _ in jobs])),
           "atoms_range": [int(min(n for _, n, _ in jobs)), int(max(n for _, n, _ in jobs))]}
    if latency_file:
        pathlib.Path(latency_file).write_text(json.dumps({"part": part, "latencies_s": lat.tolist(),
                                                          "job_ids": [j.job_id for j in done],
                                                          "atoms": [jobs[j.job_id][1] for j in done],
                                                          "phases_s": [j.phases() for j in done],
                                                          "integrals": n_int, "span_s": span}))
    log(f"serving: {res}")
    return res


BATCH_BUILDERS = 4      # host threads building wave plans concurrently


def tune_queue_leg(path, budget=200e9, n_mol=256):
    """Autotune the persistent queue (grid occupancy cap x entry grouping) on the first wave of the
    batched workload and write the table to ``path`` (engine.autotune_queue)."""
    import torch
    from electron_integral_playground_amd import generators, pair_utility
    from electron_integral_playground_amd.engine import BatchPlan, autotune_queue
    sizes = generators.batched_sizes(n_mol=n_mol)
    wave, acc = [], 0
    for i, n in enumerate(sizes):
        mol = generators.batched_molecule(i, int(n))
        b = 8 * (mol.n_ao * (mol.n_ao + 1) // 2 * mol.n_aux + mol.n_aux ** 2)
        if wave and acc + b > budget:
            break
        wave.append((mol, pair_utility.form_pair_data(mol, 1e-14), pair_utility.form_auxiliary_pair_data(mol)))
        acc += b
    bp = BatchPlan(wave)
    out = bp.new_output()
    times = autotune_queue(bp, out, save=path, log=log)
    return {"molecules": len(wave), "ms": {f"waves {w}, entry cost {c:.0e}": t for (w, c), t in times.items()}}



def batched_leg(budget=200e9, n_mol=256, repeats=2):
    """Batched workload: 256 molecules, atom counts lognormal (median 20, sigma 0.5, clipped 4-120), plain
    generators (orbital l <= 2, auxiliary l <= 3), box scaled with the atom count.  Molecules are grouped
    in arrival order into waves whose outputs fit the device budget, all written into one shared output
    buffer.  Host work is pipelined: 8 threads form the pair lists, BATCH_BUILDERS threads build the wave
    plans and work queues ahead while the device runs the current wave.  Each wave's (ij|P) runs in the
    persistent tier kernels over its work queue, from ONE captured HIP graph replayed for every wave (the
    queue is uploaded into the graph's buffers); the zeroing of screened rows and every molecule's (P|Q)
    run beside it on side streams.  Every wave then runs once more as one launch per class (the previous
    form) for comparison (not counted in the wall time); the two outputs' checksums are compared (relative
    to the sum of |values|; the GPU tests compare them element by element).  Device time: HIP events around each wave's work, fastest
    of ``repeats``; outputs reduced to a checksum on the device."""
    import concurrent.futures as cf
    import torch
    from electron_integral_playground_amd import generators, pair_utility
    from electron_integral_playground_amd.engine import BatchPlan, QueueGraph
    t0 = time.perf_counter()
    sizes = generators.batched_sizes(n_mol=n_mol)
    mols = [generators.batched_molecule(i, int(sizes[i])) for i in range(n_mol)]
    out_bytes = [8 * (m.n_ao * (m.n_ao + 1) // 2 * m.n_aux + m.n_aux ** 2) for m in mols]
    waves, cur, acc = [], [], 0
    for i, b in enumerate(out_bytes):
        if cur and acc + b > budget:
            waves.append(cur)
            cur, acc = [], 0
        cur.append(i)
        acc += b
    waves.append(cur)
    n3 = [m.n_ao * (m.n_ao + 1) // 2 * m.n_aux for m in mols]
    n2 = [m.n_aux ** 2 for m in mols]
    out = torch.empty(max(sum(n3[i] for i in w) for w in waves), dtype=torch.float64, device="cuda")
    out2 = torch.empty(max(sum(n2[i] for i in w) for w in waves), dtype=torch.float64, device="cuda")
    prep_pool = cf.ThreadPoolExecutor(8)
    pair_futs = [prep_pool.submit(lambda m: (pair_utility.form_pair_data(m, 1e-14),
                                             pair_utility.form_auxiliary_pair_data(m)), mol) for mol in mols]
    builder = cf.ThreadPoolExecutor(BATCH_BUILDERS)
    tls = threading.local()
    host_wait = {"pairs": 0.0, "plan": 0.0}

    def build(w):
        t = time.perf_counter()
        items = [(mols[i],) + pair_futs[i].result() for i in w]
        host_wait["pairs"] += time.perf_counter() - t
        t = time.perf_counter()
        if not hasattr(tls, "stream"):
            tls.stream = torch.cuda.Stream()
        with torch.cuda.stream(tls.stream):
            bp = BatchPlan(items)
            q = bp.work_queue(out)
            tls.stream.synchronize()
        host_wait["plan"] += time.perf_counter() - t
        return bp, q

    stream = torch.cuda.Stream()
    graph = QueueGraph()
    t_dev = t_cls = 0.0
    checksum = checksum_cls = 0.0
    rel_diff = 0.0
    t_wait = t_cmp = 0.0
    # plans of up to BATCH_BUILDERS waves ahead are built concurrently (their inputs are small beside the
    # shared output buffer)
    futs = [builder.submit(build, w) for w in waves[:BATCH_BUILDERS]]
    for k, w in enumerate(waves):
        tw = time.perf_counter()
        bp, q = futs[k].result()
        t_wait += time.perf_counter() - tw
        if k + BATCH_BUILDERS < len(waves):
            futs.append(builder.submit(build, waves[k + BATCH_BUILDERS]))
        stream.wait_stream(torch.cuda.current_stream())
        best = None
        for _ in range(repeats):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            bp.run_persistent(out, q, graph, out2c=out2, stream=stream)
            e1.record(stream)
            e1.synchronize()
            best = e0.elapsed_time(e1) if best is None else min(best, e0.elapsed_time(e1))
        t_dev += best * 1e-3
        with torch.cuda.stream(stream):
            n_out = bp.n_out
            wave_sum = float(out[:n_out].sum().item())
            wave_abs = float(torch.linalg.vector_norm(out[:n_out], ord=1).item())
            checksum += wave_sum + float(out2[:int(bp.out2c_base[-1])].sum().item())
        # the same wave as one launch per class (comparison; its time is not part of the wall time)
        tc = time.perf_counter()
        best = None
        for _ in range(repeats):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            with torch.cuda.stream(stream):
                bp.run(out, stream=stream)
                bp.two_center(out2, stream=stream)
            e1.record(stream)
            e1.synchronize()
            best = e0.elapsed_time(e1) if best is None else min(best, e0.elapsed_time(e1))
        t_cls += best * 1e-3
        with torch.cuda.stream(stream):
            cls_sum = float(out[:n_out].sum().item())
            checksum_cls += cls_sum + float(out2[:int(bp.out2c_base[-1])].sum().item())
            rel_diff = max(rel_diff, abs(cls_sum - wave_sum) / max(wave_abs, 1e-300))
        torch.cuda.synchronize()
        t_cmp += time.perf_counter() - tc
        del bp, q
    builder.shutdown()
    prep_pool.shutdown()
    n_int = sum(n3)
    res = {"workload": f"batched: {n_mol} molecules, atoms lognormal(median 20, sigma 0.5) in [4, 120], "
                       "3s2p1d / aux s-f", "molecules": n_mol, "atoms_min_median_max": [int(sizes.min()),
                       float(np.median(sizes)), int(sizes.max())],
           "waves": len(waves), "three_center_integrals": n_int,
           "path": "persistent tier kernels over one work queue per wave, one HIP graph captured "
                   f"{graph.captures}x and replayed for all {len(waves)} waves",
           "integrals_per_s_device": n_int / t_dev, "device_s": t_dev,
           "class_launches_device_s": t_cls, "class_launches_integrals_per_s_device": n_int / t_cls,
           "checksum_diff_vs_class_launches_rel": rel_diff,
           "wall_s_incl_setup": time.perf_counter() - t0 - t_cmp,
           "integrals_per_s_wall": n_int / (time.perf_counter() - t0 - t_cmp),
           "host_wait_for_next_wave_s": t_wait, "host_pair_lists_wait_s": host_wait["pairs"],
           "host_plan_build_s": host_wait["plan"], "checksum": checksum, "checksum_class_launches": checksum_cls}
    log(f"batched leg: {res}")
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--threshold", type=float, default=1e-6, help="mixed split threshold: on the values (checked) or on the bound (item, pair)")
    ap.add_argument("--slice-atoms", type=int, default=SLICE_ATOMS,
                    help="auxiliary shells of this many atoms form the headline slice")
    ap.add_argument("--headline-only", action="store_true", help="only the headline (profiling runs)")
    ap.add_argument("--no-plain80", action="store_true", help="skip the 80-atom block")
    ap.add_argument("--plain80-only", action="store_true")
    ap.add_argument("--mixed-full", action="store_true",
                    help="run only the full 200-atom comparison (all 3.9 TB streamed by both paths)")
    ap.add_argument("--mixed-aux-order", choices=("qp", "natural"), default="natural")
    ap.add_argument("--mixed-granularity", choices=("checked", "item", "pair"), default="item",
                    help="headline mixed split: whole items ('item', default), the float64 prefix / float32 "
                         "suffix of an item's primitive pairs ('pair', Schwarz estimate only), or 'checked' "
                         "(bf16 matrix-core float32 pass first, float64 for the items it flags)")
    ap.add_argument("--mixed-estimate", choices=("envelope", "schwarz"), default="envelope",
                    help="triplet estimate of the split: the a-priori term scale with distance decay "
                         "('envelope', default) or the reference's Schwarz-type bound ub_pp Q_P ('schwarz')")
    ap.add_argument("--tune", metavar="JSON",
                    help="autotune the chunk-plan variants (vector and matrix-core) on the headline slice first "
                         "(or, with --mixed-full, on the full workload) and save the table to JSON")
    ap.add_argument("--tune-f32-only", action="store_true", help="with --tune: the float32 tier's variants only")
    ap.add_argument("--tune-step", metavar="JSON",
                    help="before the timed headline steps, choose the variants of the classes above 1 %% of the "
                         "step by the step time itself (engine.autotune_in_step) and write them to JSON")
    ap.add_argument("--tuning", metavar="JSON", help="apply this variant table on top of the shipped one")
    ap.add_argument("--serving", action="store_true", help="run only the serving workload")
    ap.add_argument("--batched-only", action="store_true", help="run only the batched workload")
    ap.add_argument("--tune-queue", metavar="JSON",
                    help="autotune the persistent queue on the first batched wave, write the table to JSON, exit")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched workload")
    ap.add_argument("--serving-jobs", type=int, default=2000)
    ap.add_argument("--no-serving", action="store_true", help="skip the short serving run of the default line")
    ap.add_argument("--line-serving-jobs", type=int, default=100,
                    help="jobs of the short serving run included in the default line")
    ap.add_argument("--serving-rate", type=float, default=1.8, help="Poisson arrival rate, jobs/s")
    ap.add_argument("--serving-plain", action="store_true", help="serve in plain float64 instead of mixed")
    ap.add_argument("--serving-persistent", action="store_true",
                    help="serving steps as one work queue per precision through the persistent tier kernels "
                         "instead of one launch per class and precision")
    ap.add_argument("--serving-part", type=int, default=0, help="part index of a stream run in several calls")
    ap.add_argument("--serving-latencies", metavar="JSON", help="write every job's latency to this file")
    ap.add_argument("--single-stream", action="store_true",
                    help="every launch of a step on one stream (80-atom block and headline steps; profiling runs: "
                         "kernel durations do not overlap)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="first run the library's plain copy kernel %d times over %d MiB (calibrates rocprofv3 "
                         "FETCH_SIZE/WRITE_SIZE)" % (COPY_PROBE_RUNS, COPY_PROBE_BYTES >> 20))
    ap.add_argument("--profile-plain", action="store_true",
                    help="headline (or, with --plain80-only, the 80-atom block): only the plain FP64 passes and "
                         "one per-class event pass, printing the per-class figures the --trace-summary and "
                         "--pmc-summary tables use (the command profiled for them)")
    ap.add_argument("--profile-mixed", action="store_true",
                    help="with --profile-plain: the mixed-precision step (envelope split at --threshold) instead")
    ap.add_argument("--pmc-summary", nargs=3, metavar=("FETCH_CSV", "WRITE_CSV", "RUN_JSON"),
                    help="write the per-class HBM traffic table from two rocprofv3 --pmc counter_collection.csv "
                         "files of a --profile-plain run whose printed line is in RUN_JSON, then exit")
    ap.add_argument("--sq-summary", nargs=2, metavar=("PASS_A_CSV", "PASS_B_CSV"),
                    help="write the per-class SQ counter table (stalls, VALU use, FP64 FMA share) from the two "
                         "rocprofv3 --pmc counter_collection.csv files of a --profile-plain run, then exit")
    ap.add_argument("--trace-summary", nargs=2, metavar=("KERNEL_STATS_CSV", "RUN_JSON"),
                    help="write the per-class kernel-time table from a rocprofv3 --kernel-trace --stats "
                         "kernel_stats.csv of a --profile-plain run, then exit")
    args = ap.parse_args()
    if args.pmc_summary:
        print(pmc_summary(*args.pmc_summary))
        return
    if args.sq_summary:
        print(sq_summary(*args.sq_summary))
        return
    if args.trace_summary:
        print(trace_summary(*args.trace_summary))
        return
    if args.gpus != 1:
        print("bench.py runs on one GPU only: the integral path here is a single-GPU build "
              "(multi-GPU sharding by auxiliary blocks is a design note, not implemented).", file=sys.stderr)
        sys.exit(2)

    import torch
    __graft_entry__.load_package()
    from electron_integral_playground_amd.engine import apply_tuning

    torch.cuda.set_device(0)
    variants = apply_tuning().get("variants", {})
    if args.tuning:
        variants.update(apply_tuning(args.tuning).get("variants", {}))
    if args.copy_probe:
        copy_probe()
    if args.serving:
        print(json.dumps({"serving": serving_leg(args.serving_jobs, args.serving_rate, mixed=not args.serving_plain,
                                                 part=args.serving_part, latency_file=args.serving_latencies,
                                                 persistent=args.serving_persistent)}),
              flush=True)
        return
    if args.batched_only:
        print(json.dumps({"batched": batched_leg()}), flush=True)
        return
    if args.tune_queue:
        print(json.dumps({"tune_queue": tune_queue_leg(args.tune_queue)}), flush=True)
        return
    if args.profile_plain:
        print(json.dumps(plain80_leg(args) if args.plain80_only else profile_plain_leg(args)), flush=True)
        return
    if args.mixed_full:
        print(json.dumps({"mixed_full": mixed_leg(tune=args.tune, aux_order=args.mixed_aux_order)}), flush=True)
        return
    if args.plain80_only:
        print(json.dumps({"plain80": plain80_leg(args)}), flush=True)
        return
    # The drop-in loop first, in the state a caller of the C ABI finds the process in: timed after the
    # headline (its 2 x 40 GB buffers in torch's cache, the CPU baseline's 16-thread OpenMP pool) the same
    # calls measured 11 ms instead of 4.3-5.2 ms.
    dropin = None if args.headline_only else dropin_leg()
    head = headline_leg(args)
    plain80 = batched = serving = omega80 = None
    if not args.headline_only:
        if not args.no_plain80:
            plain80 = plain80_leg(args)
            omega80 = plain80_leg(argparse.Namespace(**{**vars(args), "steps": 2}), omega=0.2)
        if not args.no_batched:
            batched = batched_leg()
        if not args.no_serving:
            torch.cuda.empty_cache()
            serving = serving_leg(args.line_serving_jobs, args.serving_rate)
    line = {"metric": METRIC, "value": head["value"], "unit": "3c integrals/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": head["ms_per_step"], "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded generators)",
            "config": {"workload": "mixed-precision workload (200 atoms, 20 A box, 4s3p2d1f 1-6 prim, aux s-g 8 exp/l "
                                   "ratio 2.2; n_ao 6000, n_aux 40000), auxiliary-block share of the first "
                                   f"{args.slice_atoms} atoms (n_aux {head['n_aux_slice']}) x all "
                                   f"{head['primitive_pairs']} primitive pairs; plain FP64",
                       "n_ao": head["n_ao"], "n_aux_slice": head["n_aux_slice"], "n_aux_full": head["n_aux_full"],
                       "primitive_pairs": head["primitive_pairs"],
                       "output": "compact (ij|P) rows of significant shell pairs x slice columns, float64, streamed "
                                 f"through a {BLOCK_BYTES / 1e9:.0f} GB device buffer in {head['blocks_per_step']} "
                                 "blocks per step; inputs resident before timing"},
            "roofline": head["roofline"], "cpu_baseline": head["cpu_baseline"],
            "mixed_precision": head["mixed"],
            "max_abs_err_vs_cpu_reference_sample": head["max_abs_err_vs_cpu_reference_sample"],
            "class_seconds_per_step": head["class_seconds_per_step"], "ms_each_step": head["ms_each_step"],
            "plain80": plain80, "plain80_omega": omega80, "dropin_loop": dropin, "batched": batched, "serving": serving,
            "three_center_variants": variants}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
