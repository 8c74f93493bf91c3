"""Benchmark of the McMurchie-Davidson three-centre path on one MI355X; prints one JSON line.

value: the plain FP64 workload -- every (ij|P) of 80 atoms in a 14 A box, orbital basis 3s2p1d with 1-6
primitives per shell, auxiliary s-f with 8 even-tempered exponents per l (n_ao = 1120, n_aux = 10240),
written into the device-resident packed output J3c_packed[i >= j, P] (51.4 GB) with inputs resident in
HBM before timing.  value = n_aux * n_ao (n_ao + 1) / 2 / (seconds per step).

Beside it, in the same line:
* roofline of the dominant kernel class (HIP-event time on the launch stream, algorithmic flops from
  the fixed per-triplet model, HBM traffic from the committed rocprofv3 counter summary);
* cpu_baseline: the reference's OpenMP C++ (oracle/_ref) on a bounded sample, with parity of the sample;
* mixed_precision: the 200-atom workload (n_ao 6000, n_aux 40000, 3.9 TB of output streamed in blocks),
  plain FP64 against the FP32/FP64 split, max abs and RMS error on the device, metric Cholesky;
* batched: 256 molecules (lognormal sizes) in waves of HIP-graph-captured batch plans;
* serving: a short Poisson stream through the continuous-batching server (p50/p99 latency);
  `--serving --serving-jobs N` runs the long stream alone.
"""

import argparse
import ctypes
import json
import os
import pathlib
import sys
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

PEAK_FP64 = 78.6e12          # MI355X FP64 vector (and matrix) peak, FLOP/s
PEAK_HBM = 8.0e12            # bytes/s
SAMPLE_SHELLS = 96           # auxiliary shells per angular class in the CPU-baseline sample
PMC_SUMMARY = ROOT / "profiles" / "j3c_plain_pmc.md"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(mol, pairs, aux, gpu_packed, max_seconds=20.0):
    """Time the reference C++ (or the oracle's restatement if the reference build is absent) on the
    first auxiliary shells of every class, all orbital pairs; return throughput and parity."""
    from tests.helpers import REF_SO, ORACLE_SO, dp, ip
    kind, path, fn = ("reference", REF_SO, "three_center") if REF_SO.exists() else ("port", ORACLE_SO, "oracle_three_center")
    cores = os.cpu_count() or 1
    os.environ.setdefault("OMP_NUM_THREADS", str(cores))
    lib = ctypes.CDLL(str(path))
    n_ao = mol.n_ao
    # sample: aux shells whose functions start in [0, n_blk) of each class (rebased block output)
    n_shell_per_class = SAMPLE_SHELLS
    total_t, total_int = 0.0, 0
    max_err = 0.0
    for k_L, pk in aux.items():
        starts = np.unique(pk.i_ao_start)[:n_shell_per_class]
        sel = np.isin(pk.i_ao_start, starts)
        A = np.ascontiguousarray(pk.A_a[sel]); c = np.ascontiguousarray(pk.coefficient[sel])
        base = {int(s): k * (2 * k_L + 1) for k, s in enumerate(starts)}
        st = np.array([base[int(s)] for s in pk.i_ao_start[sel]], dtype=np.int32)
        nblk = len(starts) * (2 * k_L + 1)
        J = np.zeros((nblk, n_ao, n_ao))
        t0 = time.perf_counter()
        for (i_L, j_L), pd in pairs.items():
            rc = getattr(lib, fn)(i_L, j_L, k_L, dp(pd.P_p), dp(pd.A_a), dp(pd.B_b), dp(pd.coefficient),
                                  ip(pd.i_ao_start), ip(pd.j_ao_start), pd.P_p.shape[0], dp(A), dp(c), ip(st),
                                  A.shape[0], dp(J), n_ao, nblk, ctypes.c_bool(True), ctypes.c_double(0.0))
            assert rc == 0
        total_t += time.perf_counter() - t0
        total_int += nblk * n_ao * (n_ao + 1) // 2
        # parity of the sample against the GPU packed output
        tri_i, tri_j = np.tril_indices(n_ao)
        for k, s in enumerate(starts if gpu_packed is not None else []):
            for m in range(2 * k_L + 1):
                gpu_col = gpu_packed[:, int(s) + m]
                ref_col = J[k * (2 * k_L + 1) + m][tri_i, tri_j]
                max_err = max(max_err, float(np.max(np.abs(gpu_col - ref_col))))
        if total_t > max_seconds:
            break
    sample = f"{n_shell_per_class} auxiliary shells per angular class x all {sum(len(p.P_p) for p in pairs.values())} primitive pairs"
    return {"value": total_int / total_t, "unit": "3c integrals/s", "cores": int(os.environ["OMP_NUM_THREADS"]),
            "kind": kind, "sample": sample, "seconds": round(total_t, 3)}, max_err


def measured_traffic(la, lb, lc):
    """HBM bytes of one pass of class (la lb | lc) (all its component kernels), from the committed
    rocprofv3 WRITE_SIZE / FETCH_SIZE summary (calibrated against a device copy); None if absent."""
    if not PMC_SUMMARY.exists():
        return None
    total, found = 0.0, False
    for line in PMC_SUMMARY.read_text().splitlines():
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        # rows of the traffic table: | kernel | WRITE_SIZE KiB | FETCH_SIZE KiB | calibrated bytes (x.xxxxe+yy) |
        if (len(cells) == 4 and cells[0].startswith(f"j3c_{la}{lb}{lc}_") and not cells[0].endswith("_f32")
                and "e+" in cells[3]):
            try:
                total += float(cells[3])
                found = True
            except ValueError:
                pass
    return total if found else None


def q_aux_from_metric(plan, J2c):
    """Q_P = sqrt(max_m (P_m|P_m)) per auxiliary shell, per class, in the plan's shell order."""
    diag = J2c.diagonal().cpu().numpy()
    out = []
    for ac in plan.aux:
        idx = ac.ks_start[:, None] + np.arange(2 * ac.lc + 1)[None, :]
        out.append(np.sqrt(np.abs(diag[idx]).max(axis=1)))
    return out


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


def serving_leg(n_jobs, rate, budget=200e9, mixed=True, part=0, latency_file=None):
    """Poisson stream of integral jobs (atom counts lognormal, median 40, clipped 10-250; 4s3p2d1f
    orbital and s-g auxiliary generators) through the continuous-batching server."""
    import torch
    from electron_integral_playground_amd import generators
    from electron_integral_playground_amd.serving import IntegralServer, run_poisson
    # part k of a longer stream: its own arrival/size seed and job seeds
    jobs = generators.serving_jobs(n_jobs=n_jobs, rate=rate, seed=generators.SEED_BASE + 5 + 1000 * part)
    jobs = [(t, n, seed + 100000 * part) for t, n, seed in jobs]
    # warm-up: one small job through a throw-away server (kernel loading, allocator)
    warm = IntegralServer(budget_bytes=budget, mixed=mixed)
    warm.submit(0, generators.serving_molecule(10, 1))
    warm.drain(timeout=300)
    warm.close()
    del warm
    torch.cuda.empty_cache()
    server = IntegralServer(budget_bytes=budget, mixed=mixed)
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


def batched_leg(budget=200e9, n_mol=256):
    """Batched workload: 256 molecules, atom counts lognormal (median 20, sigma 0.5, clipped 4-120), plain
    generators (orbital l <= 2, auxiliary l <= 3), box scaled with the atom count.  Molecules are grouped
    in arrival order into waves whose outputs fit the device budget; every wave is one BatchPlan (one
    launch per class and component chunk for all its molecules, plus their (P|Q)), captured once in a
    HIP graph and replayed; each wave's outputs are reduced to a checksum on the device, then freed."""
    import concurrent.futures as cf
    import torch
    from electron_integral_playground_amd import generators, pair_utility
    from electron_integral_playground_amd.engine import BatchPlan
    t0 = time.perf_counter()
    sizes = generators.batched_sizes(n_mol=n_mol)

    def prep(i):
        mol = generators.batched_molecule(i, int(sizes[i]))
        return mol, pair_utility.form_pair_data(mol, 1e-14), pair_utility.form_auxiliary_pair_data(mol)

    with cf.ThreadPoolExecutor(8) as ex:
        mols = list(ex.map(prep, range(n_mol)))
    t_host = time.perf_counter() - t0
    out_bytes = [8 * (m.n_ao * (m.n_ao + 1) // 2 * m.n_aux + m.n_aux ** 2) for m, _, _ in mols]
    waves, cur, acc = [], [], 0
    for i, b in enumerate(out_bytes):
        if cur and acc + b > budget:
            waves.append(cur)
            cur, acc = [], 0
        cur.append(i)
        acc += b
    waves.append(cur)
    stream = torch.cuda.Stream()
    t_dev = 0.0
    checksum = 0.0
    for w in waves:
        bp = BatchPlan([mols[i] for i in w])
        out = bp.new_output()
        out2 = bp.new_two_center_output()
        stream.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(stream):
            bp.run(out)                                # warm-up outside capture
            bp.two_center(out2)
        torch.cuda.current_stream().wait_stream(stream)
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            bp.run(out)
            bp.two_center(out2)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        g.replay()
        e1.record()
        torch.cuda.synchronize()
        t_dev += e0.elapsed_time(e1) * 1e-3
        checksum += float(out.sum().item()) + float(out2.sum().item())
        del g, out, out2, bp
        torch.cuda.empty_cache()
    n_int = sum(m.n_aux * m.n_ao * (m.n_ao + 1) // 2 for m, _, _ in mols)
    res = {"workload": f"batched: {n_mol} molecules, atoms lognormal(median 20, sigma 0.5) in [4, 120], "
                       "3s2p1d / aux s-f", "molecules": n_mol, "atoms_min_median_max": [int(sizes.min()),
                       float(np.median(sizes)), int(sizes.max())],
           "waves": len(waves), "three_center_integrals": n_int,
           "integrals_per_s_device": n_int / t_dev, "device_s": t_dev, "wall_s_incl_setup": time.perf_counter() - t0,
           "host_pair_lists_s": t_host, "checksum": checksum}
    log(f"batched leg: {res}")
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--tune", metavar="JSON", help="autotune chunk-plan variants first and save them to JSON")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision workload")
    ap.add_argument("--mixed-only", action="store_true", help="run only the mixed-precision workload")
    ap.add_argument("--mixed-atoms", type=int, default=None, help="atoms of the mixed workload (default 200)")
    ap.add_argument("--mixed-aux-order", choices=("qp", "natural"), default="natural",
                    help="auxiliary-shell grouping of the mixed split (see ThreeCenterPlan.setup_mixed)")
    ap.add_argument("--serving", action="store_true", help="run only the serving workload")
    ap.add_argument("--batched-only", action="store_true", help="run only the batched workload")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched workload")
    ap.add_argument("--serving-jobs", type=int, default=2000)
    ap.add_argument("--no-serving", action="store_true", help="skip the short serving run of the default line")
    ap.add_argument("--line-serving-jobs", type=int, default=100,
                    help="jobs of the short serving run included in the default line")
    ap.add_argument("--serving-rate", type=float, default=1.8, help="Poisson arrival rate, jobs/s")
    ap.add_argument("--serving-plain", action="store_true", help="serve in plain float64 instead of mixed")
    ap.add_argument("--serving-part", type=int, default=0, help="part index of a stream run in several calls")
    ap.add_argument("--serving-latencies", metavar="JSON", help="write every job's latency to this file")
    ap.add_argument("--single-stream", action="store_true",
                    help="time the steps with every launch on one stream (profiling runs: kernel durations "
                         "then do not overlap)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="also run one 4 GiB device copy (calibrates rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: the integral path here is a single-GPU build "
              "(multi-GPU sharding by auxiliary blocks is a design note, not implemented).", file=sys.stderr)
        sys.exit(2)

    import torch
    __graft_entry__.load_package()
    from electron_integral_playground_amd import generators, pair_utility
    from electron_integral_playground_amd.engine import (N_SIDE_STREAMS, ThreeCenterPlan, apply_tuning, autotune,
                                                          three_center_flops_per_triplet)

    torch.cuda.set_device(0)
    if args.serving:
        print(json.dumps({"serving": serving_leg(args.serving_jobs, args.serving_rate, mixed=not args.serving_plain,
                                                 part=args.serving_part, latency_file=args.serving_latencies)}),
              flush=True)
        return
    if args.batched_only:
        print(json.dumps({"batched": batched_leg()}), flush=True)
        return
    if args.mixed_only:
        print(json.dumps({"mixed_precision": mixed_leg(n_atom=args.mixed_atoms, tune=args.tune,
                                                       aux_order=args.mixed_aux_order)}), flush=True)
        return
    t0 = time.perf_counter()
    mol = generators.make_molecule("plain")
    pairs = pair_utility.form_pair_data(mol, 1e-14)
    aux = pair_utility.form_auxiliary_pair_data(mol)
    t_host = time.perf_counter() - t0
    log(f"workload plain: n_ao={mol.n_ao} n_aux={mol.n_aux} pairs={sum(len(p.P_p) for p in pairs.values())} "
        f"host pair list {t_host:.1f}s")
    plan = ThreeCenterPlan(mol, pairs, aux)
    out = plan.new_output()
    J2c = torch.zeros((mol.n_aux, mol.n_aux), dtype=torch.float64, device="cuda")   # (P|Q), part of each step
    stream = torch.cuda.current_stream()
    if args.copy_probe:
        a = torch.empty(1 << 29, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        b.copy_(a)                                   # 4 GiB read + 4 GiB written
        torch.cuda.synchronize()
        del a, b
    if args.tune:
        variants = autotune(plan, save=args.tune, log=log)
    else:
        variants = apply_tuning().get("variants", {})

    for _ in range(args.warmup):
        plan.run(out, concurrent=not args.single_stream, metric_out=J2c)
    torch.cuda.synchronize()
    log("warmup done")

    # per-class kernel times with HIP events on the launch stream (one extra pass)
    ev = []
    for pi, pc in enumerate(plan.pairs):
        for ai, ac in enumerate(plan.aux):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            plan.launch_class(pi, ai, out)
            e1.record(stream)
            ev.append(((pc.la, pc.lb, ac.lc), pc.n_pp * ac.n_prim, e0, e1))
    torch.cuda.synchronize()
    per_class = []
    for key, ntrip, e0, e1 in ev:
        ms = e0.elapsed_time(e1)
        fl = ntrip * three_center_flops_per_triplet(*key)
        per_class.append((key, ms, fl))
    per_class.sort(key=lambda r: -r[1])
    for key, ms, fl in per_class[:8]:
        log(f"  class {key}: {ms:.3f} ms  {fl / ms / 1e9:.2f} TFLOP/s (model)")

    # the same step with every launch on one stream, for comparison (not the reported value)
    dt_serial = None
    if not args.single_stream:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(max(1, args.steps // 2)):
            plan.run(out, concurrent=False, metric_out=J2c)
        torch.cuda.synchronize()
        dt_serial = (time.perf_counter() - t0) / max(1, args.steps // 2)
        log(f"single-stream step {dt_serial * 1e3:.2f} ms")

    # timed region
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        plan.run(out, concurrent=not args.single_stream, metric_out=J2c)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    n_int = plan.n_integrals()
    value = n_int / dt
    log(f"step {dt * 1e3:.2f} ms -> {value:.4e} integrals/s; model flops/step {plan.model_flops():.3e}")

    key, ms, fl = per_class[0]
    nchunks = plan.lib.eip_three_center_nchunks(*key)
    roofline = {"bound": "mfma", "pipe": "fp64-valu", "kernel": "j3c_%d%d%d (all %d component chunks)" % (key + (nchunks,)),
                "achieved": fl / (ms * 1e-3) / 1e12, "peak": PEAK_FP64 / 1e12, "unit": "TFLOP/s",
                "frac": fl / (ms * 1e-3) / PEAK_FP64, "traffic": measured_traffic(*key),
                "traffic_unit": "bytes per class pass (rocprofv3 FETCH_SIZE + WRITE_SIZE, calibrated)"}

    cpu, err = (None, None)
    if not args.no_cpu:
        gpu_packed = out.cpu().numpy() if out.numel() * 8 < 8e9 else None
        if gpu_packed is None:
            # copy only the columns the sample needs
            cols = sorted({int(s) + m for k_L, pk in aux.items() for s in np.unique(pk.i_ao_start)[:SAMPLE_SHELLS]
                           for m in range(2 * k_L + 1)})
            sub = out[:, cols].cpu().numpy()
            gpu_packed = np.zeros((out.shape[0], max(cols) + 1))
            gpu_packed[:, cols] = sub
        cpu, err = cpu_baseline(mol, pairs, aux, gpu_packed)
        log(f"cpu baseline {cpu}; sample max |GPU - reference| = {err:.3e}")

    mixed = batched = None
    del out, plan, J2c
    torch.cuda.empty_cache()
    if not args.no_mixed:
        mixed = mixed_leg(n_atom=args.mixed_atoms, aux_order=args.mixed_aux_order)
    if not args.no_batched:
        batched = batched_leg()
    serving = None
    if not args.no_serving:
        torch.cuda.empty_cache()
        serving = serving_leg(args.line_serving_jobs, args.serving_rate)
    line = {"metric": "3c integrals/s, mixed vs plain FP64 at <=1e-10 err; p50/p99 job latency", "value": value,
            "unit": "3c integrals/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": dt * 1e3, "single_stream_ms_per_step": dt_serial * 1e3 if dt_serial else None,
            "streams": 1 if args.single_stream else N_SIDE_STREAMS + 2, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f64", "data": "synthetic (seeded generators, plain workload)",
            "config": {"workload": "plain: 80 atoms, 14 A box, 3s2p1d (1-6 prim), aux s-f 8 exp/l ratio 2.2",
                       "n_ao": mol.n_ao, "n_aux": mol.n_aux,
                       "primitive_pairs": int(sum(len(p.P_p) for p in pairs.values())),
                       "output": "packed (n_ao(n_ao+1)/2, n_aux) float64 (ij|P) plus dense (n_aux, n_aux) float64 "
                                 "(P|Q), on device, every step"},
            "roofline": roofline, "cpu_baseline": cpu,
            "three_center_variants": variants,
            "mixed_precision": mixed,
            "batched": batched,
            "serving": serving,
            "max_abs_err_vs_reference_sample": err}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
