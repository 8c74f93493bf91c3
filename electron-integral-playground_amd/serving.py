"""Continuous-batching integral server: jobs (geometry + basis) arrive over time; their pair tables,
metric and three-centre outputs stay resident in HBM; admission is by the projected footprint of each
job's (ij|P) tensor against a device budget; all admitted work of one step runs as ONE launch per
angular class (work items of every job in the step, each item pointing at its job through a job table),
and a new step is formed while the previous one is still on the GPU, so arrivals join as memory and
slots free up.

Unit of work: a whole job when its compact (ij|P) tensor (significant shell pairs x n_aux, float64)
fits the per-unit cap, otherwise blocks of shell pairs of one class (a job larger than device memory is
streamed block by block; every block is consumed -- reduced to a checksum -- and freed when it is done).
A job is complete when all its units are; latency = completion (observed on the host) - arrival.

Determinism: every output element is written by exactly one lane of one float64 launch (and, in the
mixed mode, added to by one lane of one float32 launch) with a fixed summation order, so a job's result
does not depend on which other jobs share its launches.
"""

import concurrent.futures as cf
import os
import threading
import time
import warnings
from collections import deque
from dataclasses import dataclass, field
from typing import Dict, List, Optional

import numpy as np

from . import pair_utility
from .engine import ThreeCenterPlan, WorkQueue

UNIT_CAP_FRACTION = 0.25      # a unit's output may use at most this share of the budget
INPUT_FRACTION = 0.25         # admitted jobs' resident inputs (pair tables, metric) use at most this share
BYPASS_SECONDS = 5.0          # a blocked head unit may be overtaken by smaller units for this long
SETUP_CONCURRENCY = 3         # jobs in device setup at once (their transient allocations are not gated)
LARGE_JOB_ATOMS = 80          # jobs at least this large (by the submitter's size hint) prepare in their own
LARGE_PREP_WORKERS = 4        # pool, so a burst of long preparations never holds the small jobs
# Device memory that PyTorch keeps cached (freed, not returned) is unavailable to the HIP runtime's own
# allocations -- e.g. the scratch memory a queue needs when it first runs a kernel with a private
# segment.  Finished jobs' inputs are returned to the device once the cache holds more than
# CACHE_SLACK_BYTES, at a moment when no step is in flight: returning memory (hipFree) waits for the
# whole device, so a release while a step of a large job runs holds the scheduler, and every thread
# that allocates, until that step ends.  Past CACHE_SLACK_MAX_BYTES it is returned at once.
CACHE_SLACK_BYTES = 8e9
CACHE_SLACK_MAX_BYTES = 48e9
# The server uses a compute stream and one setup stream per setup thread.  The HIP runtime maps streams
# onto GPU_MAX_HW_QUEUES in-order hardware queues (default 4, fixed when the runtime starts); a setup
# stream that shares the compute stream's queue waits behind whole steps (measured 14-20 s for small
# jobs next to a 250-atom one).  Processes that serve should start with at least this many.
MIN_HW_QUEUES = 8


def _q_aux(plan, J2c):
    diag = J2c.diagonal().cpu().numpy()
    out = []
    for ac in plan.aux:
        idx = ac.ks_start[:, None] + np.arange(2 * ac.lc + 1)[None, :]
        out.append(np.sqrt(np.abs(diag[idx]).max(axis=1)))
    return out


class InputGate:
    """Admission of jobs' resident inputs (pair tables, metric) within a byte limit.

    A job reaches the gate when its host preparation is done and passes as soon as its inputs fit (or
    nothing is resident) -- unless a job that holds an earlier ticket is already waiting at the gate for
    room: waiting jobs pass in ticket order, so a large job is never starved by smaller ones behind it.
    Jobs still being prepared hold no place in line, so a slow preparation never delays the jobs behind
    it.  Every job past the gate runs to completion and releases its inputs, so waiting always ends."""

    def __init__(self, limit, on_pass=None):
        self.limit = float(limit)
        self.used = 0.0
        self._cv = threading.Condition()
        self._tickets = 0
        self._waiting = set()         # tickets of jobs waiting at the gate
        self._on_pass = on_pass       # called with (ticket, tickets still waiting) under the lock

    def ticket(self):
        with self._cv:
            t = self._tickets
            self._tickets += 1
            return t

    def enter(self, ticket, nbytes, failed=False):
        """Block until no earlier ticket waits at the gate and nbytes fit (a failed job reserves nothing
        and does not wait)."""
        if failed:
            return
        with self._cv:
            self._waiting.add(ticket)
            self._cv.wait_for(lambda: min(self._waiting) == ticket and (
                self.used == 0 or self.used + nbytes <= self.limit))
            self._waiting.discard(ticket)
            self.used += nbytes
            if self._on_pass is not None:
                self._on_pass(ticket, frozenset(self._waiting))
            self._cv.notify_all()

    def adjust(self, delta):
        """Change the reservation (estimate -> actual, or release when a job finishes)."""
        with self._cv:
            self.used += delta
            self._cv.notify_all()


class OutputArena:
    """One device allocation for all unit outputs, carved first-fit (no allocator fragmentation between
    jobs of different sizes).  Offsets and sizes are in float64 elements."""

    def __init__(self, torch, n_elems, device):
        self.buf = torch.empty(int(n_elems), dtype=torch.float64, device=device)
        self.free = [(0, int(n_elems))]            # sorted, non-adjacent (offset, size)

    def alloc(self, n):
        for k, (off, size) in enumerate(self.free):
            if size >= n:
                if size == n:
                    self.free.pop(k)
                else:
                    self.free[k] = (off + n, size - n)
                return off
        return None

    def release(self, off, n):
        self.free.append((off, n))
        self.free.sort()
        merged = []
        for o, sz in self.free:
            if merged and merged[-1][0] + merged[-1][1] == o:
                merged[-1] = (merged[-1][0], merged[-1][1] + sz)
            else:
                merged.append((o, sz))
        self.free = merged


@dataclass
class Unit:
    job: "Job"
    blocks: List[tuple]           # (pi, sp0, sp1)
    rows: int                     # compact output rows
    bytes: int
    row_off: List[int]            # first output row of each block
    out: object = None            # device tensor (rows, n_aux) while in flight / kept (a view of the arena)
    arena_off: int = -1
    checksum: object = None
    blocked_since: Optional[float] = None


@dataclass
class Job:
    job_id: int
    arrival: float
    molecule: object
    n_ao: int = 0
    n_aux: int = 0
    plan: Optional[ThreeCenterPlan] = None
    J2c: object = None                # (P|Q) on the device (keep_outputs), else reduced to J2c_checksum
    J2c_checksum: object = None
    units: List[Unit] = field(default_factory=list)
    units_done: int = 0
    completion: Optional[float] = None
    input_bytes: int = 0
    error: Optional[str] = None
    ticket: int = 0
    # host clock: preparation started (a pool thread took the job) and done, past the input gate,
    # a setup slot taken, device setup done, first unit launched
    t_prep_start: Optional[float] = None
    t_prepared: Optional[float] = None
    t_gate: Optional[float] = None
    t_setup_start: Optional[float] = None
    t_setup: Optional[float] = None
    t_first_launch: Optional[float] = None

    @property
    def latency(self):
        return None if self.completion is None else self.completion - self.arrival

    PHASES = ("prep_wait", "prepare", "gate", "setup_wait", "setup", "queue", "run")

    def phases(self):
        """Seconds spent waiting for a preparation thread, in host preparation, waiting at the input
        gate, waiting for a setup slot, in device setup, waiting for a step, and from the first launch
        to completion (the seven add up to the latency)."""
        t = [self.arrival, self.t_prep_start, self.t_prepared, self.t_gate, self.t_setup_start, self.t_setup,
             self.t_first_launch, self.completion]
        if any(x is None for x in t):
            return None
        return {k: t[i + 1] - t[i] for i, k in enumerate(self.PHASES)}


class IntegralServer:
    """Single-GPU continuous-batching server for (ij|P) (mixed precision or plain float64) plus the
    float64 metric (P|Q) of every job."""

    def __init__(self, budget_bytes=200e9, mixed=True, threshold=1e-6, keep_outputs=False, prep_workers=12,
                 max_steps_in_flight=2, unit_cap_bytes=None, granularity="item", device="cuda", log=None,
                 on_unit=None, estimate="envelope", persistent=False):
        """``on_unit(job, unit)``, if given, consumes every finished unit: it is called when the unit's
        step has retired, with the compute stream current and ``unit.out`` still valid (rows of
        ``unit.blocks`` from ``unit.row_off``), before the unit's output memory is reused.
        ``estimate``: the mixed split's triplet estimate -- "envelope" (a-priori term scale of every work
        item, distance-aware; whole items) or "schwarz" (the reference's bound ub_pp Q_P, item or pair
        granularity).  ``persistent``: each step runs the classes the persistent tier kernels hold as one
        work queue per precision (one launch per register tier), the others with their own kernels; off
        by default, since on the 2000-job stream one launch per class and precision measured 5-10 %
        lower latency at every percentile (the queue kernels' bodies are compiled without loop-invariant
        code motion)."""
        import torch
        self.torch = torch
        self.persistent = persistent
        self.budget = float(budget_bytes)
        self.unit_cap = min(UNIT_CAP_FRACTION * self.budget, unit_cap_bytes or float("inf"))
        self.mixed = mixed
        self.threshold = threshold
        self.granularity = granularity
        if estimate not in ("envelope", "schwarz"):
            raise ValueError("estimate is 'envelope' or 'schwarz'")
        if estimate == "envelope" and granularity != "item":
            raise ValueError("the envelope estimate splits whole items (granularity 'item')")
        self.estimate = estimate
        self.keep_outputs = keep_outputs
        self.max_steps_in_flight = max_steps_in_flight
        self.device = torch.device(device)
        self.compute_stream = torch.cuda.Stream(device=self.device)
        self._tls = threading.local()                # one setup stream per preparation thread
        self.gate = InputGate(INPUT_FRACTION * self.budget)
        self._setup_slots = threading.Semaphore(SETUP_CONCURRENCY)
        self.pool = cf.ThreadPoolExecutor(prep_workers)
        self.large_pool = cf.ThreadPoolExecutor(LARGE_PREP_WORKERS)
        self.log = log
        self.on_unit = on_unit
        if int(os.environ.get("GPU_MAX_HW_QUEUES", "4")) < MIN_HW_QUEUES:
            warnings.warn(f"GPU_MAX_HW_QUEUES below {MIN_HW_QUEUES}: device setups may wait behind whole "
                          "steps; start the serving process with a larger value", RuntimeWarning)
        self.prepared = deque()       # host-prepared jobs awaiting device setup and admission
        self.pending_units = deque()  # units of admitted jobs not yet launched
        self.in_flight = deque()      # (event, [units])
        self._release_pending = False  # finished jobs freed inputs; the cache may need returning
        self.outputs_used = 0.0       # bytes of unit outputs in flight (or kept)
        # unit outputs live in one arena: the budget minus the share reserved for resident inputs
        self.arena = OutputArena(torch, (1.0 - INPUT_FRACTION) * self.budget // 8, self.device)
        self.jobs: Dict[int, Job] = {}
        self.completed: List[Job] = []
        self.n_launches = 0
        self.n_steps = 0

    # -- host preparation (thread pool) --------------------------------------------------------
    def submit(self, job_id, molecule, arrival=None, size_hint=None):
        """Accept a job (a Molecule, or a callable returning one); its geometry, basis and pair tables
        are built on a host thread.  ``size_hint`` (atoms) routes jobs of LARGE_JOB_ATOMS or more to the
        large-job preparation pool."""
        job = Job(job_id, time.perf_counter() if arrival is None else arrival, molecule)
        self.jobs[job_id] = job
        job.ticket = self.gate.ticket()
        pool = self.large_pool if size_hint is not None and size_hint >= LARGE_JOB_ATOMS else self.pool
        fut = pool.submit(self._prepare, job)
        self.prepared.append((job, fut))
        return job

    def _prepare(self, job):
        """Host thread: geometry/basis, pair tables, then -- once the job's resident inputs pass the input
        gate -- device setup on this thread's own stream.  Errors are re-raised through the future."""
        host, error, est = None, None, 0
        job.t_prep_start = time.perf_counter()
        try:
            if callable(job.molecule):
                job.molecule = job.molecule()
            mol = job.molecule
            arrays = pair_utility.form_pair_arrays(mol, pair_utility.REFERENCE_SCREENING_THRESHOLD)
            host = (arrays, pair_utility.form_primitive_pair_data(mol, arrays), pair_utility.form_auxiliary_pair_data(mol))
            est = self.estimate_input_bytes(mol, host)
        except Exception as exc:                          # still takes its turn at the gate
            error = exc
        job.t_prepared = time.perf_counter()
        self.gate.enter(job.ticket, est, failed=error is not None)
        job.t_gate = time.perf_counter()
        if error is not None:
            raise error
        try:
            with self._setup_slots:
                job.t_setup_start = time.perf_counter()
                self._setup_device(job, host)
        finally:
            self.gate.adjust((job.input_bytes if job.plan is not None else 0) - est)
        job.t_setup = time.perf_counter()

    # -- device setup and admission -----------------------------------------------------------------
    def _setup_device(self, job, host):
        torch = self.torch
        arrays, pairs, aux = host
        mol = job.molecule
        job.n_ao, job.n_aux = mol.n_ao, mol.n_aux
        stream = getattr(self._tls, "stream", None)
        if stream is None:
            stream = self._tls.stream = torch.cuda.Stream(device=self.device)
        with torch.cuda.stream(stream):
            plan = ThreeCenterPlan(mol, pairs, aux, device=self.device)
            job.J2c = plan.two_center()
            q = _q_aux(plan, job.J2c)                     # synchronises this thread's stream only
            if not self.keep_outputs:
                # the metric is an output too: consumed (checksum) now instead of staying resident --
                # at n_aux = 40 000 it is 12.8 GB of the input budget
                job.J2c_checksum = torch.stack([job.J2c.sum(), torch.linalg.vector_norm(job.J2c)])
                job.J2c = None
            if self.mixed:
                plan.setup_mixed(arrays, q, threshold=self.threshold, granularity=self.granularity,
                                 estimate=self.estimate)
            stream.synchronize()
        job.plan = plan
        job.input_bytes = plan.device_bytes() + (job.J2c.numel() * 8 if job.J2c is not None else 0)
        # units: the whole job, or blocks of shell pairs
        total_rows = sum(plan.block_rows(pi, 0, pc.n_sp) for pi, pc in enumerate(plan.pairs))
        if total_rows * mol.n_aux * 8 <= self.unit_cap:
            blocks = [(pi, 0, pc.n_sp) for pi, pc in enumerate(plan.pairs) if pc.n_sp > 0]
            groups = [blocks]
        else:
            groups = [[b] for b in plan.row_blocks(self.unit_cap)]
        for g in groups:
            offs, r = [], 0
            for b in g:
                offs.append(r)
                r += plan.block_rows(*b)
            job.units.append(Unit(job, g, r, r * mol.n_aux * 8, offs))

    @staticmethod
    def estimate_input_bytes(molecule, host):
        """Resident inputs of a job before it is set up: pair matrices (float64 + float32 copies, at most
        2143 entries per primitive pair for (ff)), primitive data, and the float64 metric."""
        _, pairs, aux = host
        n_pp = sum(len(pd.P_p) for pd in pairs.values())
        return n_pp * (120 + 12 * 400) + molecule.n_aux ** 2 * 8

    def _admit(self):
        """Every job whose preparation (host and device) has finished joins the pending units, whatever
        the state of jobs submitted before it (their inputs may be waiting for room that only finishing
        jobs release)."""
        waiting = deque()
        while self.prepared:
            job, fut = self.prepared.popleft()
            if not fut.done():
                waiting.append((job, fut))
                continue
            exc = fut.exception()
            if exc is not None:                           # a job that cannot run is reported, not fatal
                job.error = repr(exc)
                job.completion = time.perf_counter()
                self.completed.append(job)
                continue
            self.pending_units.extend(job.units)
        self.prepared = waiting

    # -- steps ----------------------------------------------------------------------------------------
    def _form_step(self):
        """Pending units whose outputs the arena can place, taken round-robin across jobs (each job's
        units in order): a job with many units (one larger than the per-unit cap) shares every step
        with the jobs behind it instead of holding them until all its units have run.  A unit is at most
        UNIT_CAP_FRACTION of the budget and the arena is (1 - INPUT_FRACTION) of it, so a blocked unit
        fits once the steps in flight have retired; after BYPASS_SECONDS blocked, no later unit is taken
        before it."""
        now = time.perf_counter()
        queues, index = [], {}
        for u in self.pending_units:
            k = id(u.job)
            if k not in index:
                index[k] = len(queues)
                queues.append(deque())
            queues[index[k]].append(u)
        units, taken, closed, stop = [], set(), set(), False
        progress = True
        while progress and not stop:
            progress = False
            for qi, q in enumerate(queues):
                if not q or qi in closed:
                    continue
                u = q[0]
                off = self.arena.alloc(u.rows * u.job.n_aux)
                if off is None:
                    if u.blocked_since is None:
                        u.blocked_since = now
                    if now - u.blocked_since > BYPASS_SECONDS:
                        stop = True              # the blocked unit has waited long enough: no bypass
                        break
                    closed.add(qi)               # this job's later units wait behind its head
                    continue
                q.popleft()
                u.arena_off = off
                u.out = self.arena.buf[off:off + u.rows * u.job.n_aux].view(u.rows, u.job.n_aux)
                units.append(u)
                taken.add(id(u))
                progress = True
        self.pending_units = deque(u for u in self.pending_units if id(u) not in taken)
        return units

    def _launch_step(self, units):
        torch = self.torch
        step_bytes = 0
        with torch.cuda.stream(self.compute_stream):
            for u in units:
                step_bytes += u.bytes
            self.outputs_used += step_bytes
            # per angular class: items of every unit, one job-table entry per (unit, pair class, aux class)
            work = {}
            entries, entries_vec = [], []
            for u in units:
                plan = u.job.plan
                for (pi, sp0, sp1), r0 in zip(u.blocks, u.row_off):
                    pc = plan.pairs[pi]
                    nij = (2 * pc.la + 1) * (2 * pc.lb + 1)
                    base = u.out.data_ptr() + 8 * (r0 - sp0 * nij) * u.job.n_aux
                    for ai, ac in enumerate(plan.aux):
                        perm = plan.mixed[(pi, ai)]["perm"] if self.mixed else None
                        entries.append(plan.job_words(pi, ai, base, u.job.n_aux, perm))
                        entries_vec.append(plan.job_words(pi, ai, base, u.job.n_aux, perm, vector=True))
                        j = len(entries) - 1
                        key = (pc.la, pc.lb, ac.lc)
                        w = work.setdefault(key, {"f64": [], "f32": [], "pp64": [], "pp32": [], "j64": [], "j32": [],
                                                  "ref": (plan, pi, ai), "acc32": 0})
                        if self.mixed:
                            it, pp, it32, pp32, w["acc32"] = plan.mixed_items(pi, ai, sp0, sp1)
                        else:
                            G = (ac.n_ks + 63) // 64
                            it, pp = plan.items[(pi, ai)][sp0 * G:sp1 * G], None
                            it32 = pp32 = None
                        if it.shape[0]:
                            w["f64"].append(it); w["pp64"].append(pp); w["j64"].append((j, it.shape[0]))
                        if it32 is not None and it32.shape[0]:
                            w["f32"].append(it32); w["pp32"].append(pp32); w["j32"].append((j, it32.shape[0]))
            keep = []

            def upload(a):
                """Host array -> device without synchronising: pinned staging, asynchronous copy on the
                compute stream; the staging buffer is kept until the step retires."""
                h = torch.from_numpy(np.ascontiguousarray(a)).pin_memory()
                keep.append(h)
                return h.to(self.device, non_blocking=True)

            jobs = upload(np.array(entries, dtype=np.int64).reshape(-1, 16))
            keep.append(jobs)
            queues = {}
            if self.persistent:
                # every class the persistent tier kernels hold goes into one queue per precision: the
                # step is then a handful of launches whatever the number of jobs and classes
                jobs_vec = upload(np.array(entries_vec, dtype=np.int64).reshape(-1, 16))
                keep.append(jobs_vec)
                queues = {prec: WorkQueue(units[0].job.plan.lib, self.device, use_f32=(prec == "f32"))
                          for prec in ("f64", "f32")}
            for key, w in work.items():
                plan, pi, ai = w["ref"]
                for prec in ("f64", "f32"):
                    if not w[prec]:
                        continue
                    items = torch.cat(w[prec])
                    pps = w["pp" + prec[1:]]
                    pp = torch.cat(pps) if pps and pps[0] is not None else None
                    jid = upload(np.repeat(np.array([j for j, _ in w["j" + prec[1:]]], dtype=np.int32),
                                           [n for _, n in w["j" + prec[1:]]]))
                    keep += [items, jid] + ([pp] if pp is not None else [])
                    acc = w["acc32"] if prec == "f32" else 0
                    q = queues.get(prec)
                    if q is not None and acc == 0 and q.supports(*key):
                        q.add(*key, items, int(items.shape[0]), jobs_vec.data_ptr(), plan.mean_pairs_per_shell_pair(pi),
                              pp=pp, job=jid, layout=1, accumulate=0, omega=plan.omega)
                        continue
                    plan.launch_work(pi, ai, items, jobs, job=jid, pp=pp, layout=1, accumulate=acc,
                                     use_f32=(prec == "f32"), stream=self.compute_stream)
                    self.n_launches += 1
            for q in queues.values():
                if q.n_triplet_items:
                    q.build_device(self.compute_stream)
                    q.run(self.compute_stream)
                    keep.append(q)
                    self.n_launches += q.ntiers
            ev = torch.cuda.Event()
            ev.record(self.compute_stream)
        self.in_flight.append((ev, units, keep))
        self.n_steps += 1
        now = time.perf_counter()
        for u in units:
            if u.job.t_first_launch is None:
                u.job.t_first_launch = now

    def _retire(self):
        """Consume the units of finished steps (checksum on the device) and free their memory."""
        torch = self.torch
        released = False
        while self.in_flight and self.in_flight[0][0].query():
            now = time.perf_counter()
            _, units, _ = self.in_flight.popleft()
            with torch.cuda.stream(self.compute_stream):
                for u in units:
                    # consumed on the device: sum and 2-norm (reductions, no full-size temporaries)
                    u.checksum = torch.stack([u.out.sum(), torch.linalg.vector_norm(u.out)])
                    if self.on_unit is not None:
                        self.on_unit(u.job, u)
                    if not self.keep_outputs:
                        u.out = None
                        # stream-ordered reuse: later launches on this stream run after the checksum
                        self.arena.release(u.arena_off, u.rows * u.job.n_aux)
                        self.outputs_used -= u.bytes
                    job = u.job
                    job.units_done += 1
                    if job.units_done == len(job.units):
                        job.completion = now
                        self.completed.append(job)
                        if not self.keep_outputs:
                            job.plan = None
                            job.J2c = None
                            self.gate.adjust(-job.input_bytes)
                            released = True
        self._release_pending |= released
        if self._release_pending:
            slack = torch.cuda.memory_reserved(self.device) - torch.cuda.memory_allocated(self.device)
            if slack <= CACHE_SLACK_BYTES:
                self._release_pending = False
            elif not self.in_flight or slack > CACHE_SLACK_MAX_BYTES:
                torch.cuda.empty_cache()
                self._release_pending = False

    def poll(self):
        """One scheduling iteration: retire finished steps, admit prepared jobs, launch a step."""
        self._retire()
        self._admit()
        if len(self.in_flight) < self.max_steps_in_flight and self.pending_units:
            units = self._form_step()
            if units:
                self._launch_step(units)
                return True
        return False

    def drain(self, timeout=None):
        t0 = time.perf_counter()
        while self.prepared or self.pending_units or self.in_flight:
            if not self.poll():
                time.sleep(2e-4)
            if timeout is not None and time.perf_counter() - t0 > timeout:
                raise TimeoutError("server did not drain")
        self.torch.cuda.synchronize()

    def close(self):
        self.pool.shutdown(wait=True)
        self.large_pool.shutdown(wait=True)

    # -- results ------------------------------------------------------------------------------------
    def job_output(self, job_id):
        """(ij|P) of a job kept on the device (keep_outputs=True, whole-job units): the compact tensor
        rows (pair class pi, shell pair sp, ij) -> {pi: tensor (n_sp * nij, n_aux)}."""
        job = self.jobs[job_id]
        out = {}
        for u in job.units:
            for (pi, sp0, sp1), r0 in zip(u.blocks, u.row_off):
                pc = job.plan.pairs[pi]
                nij = (2 * pc.la + 1) * (2 * pc.lb + 1)
                out.setdefault(pi, []).append(u.out[r0:r0 + (sp1 - sp0) * nij])
        return {pi: self.torch.cat(v) for pi, v in out.items()}


def run_poisson(server, jobs, make_molecule, log=None):
    """Feed (arrival offset s, n_atom, seed) jobs to the server at their arrival times (wall clock);
    returns per-job latencies and the span."""
    t0 = time.perf_counter()
    i = logged = 0
    n = len(jobs)
    while i < n or server.prepared or server.pending_units or server.in_flight:
        now = time.perf_counter() - t0
        while i < n and jobs[i][0] <= now:
            t_arr, n_atom, seed = jobs[i]
            server.submit(i, (lambda n=n_atom, sd=seed: make_molecule(n, sd)), arrival=t0 + t_arr, size_hint=n_atom)
            i += 1
        if not server.poll():
            time.sleep(2e-4)
        if log and len(server.completed) >= logged + 50:
            logged = len(server.completed)
            log(f"  {logged}/{n} jobs done at {time.perf_counter() - t0:.1f}s, steps in flight "
                f"{len(server.in_flight)}, pending units {len(server.pending_units)}, inputs "
                f"{server.gate.used / 1e9:.1f} GB, outputs {server.outputs_used / 1e9:.1f} GB, device memory "
                f"reserved {server.torch.cuda.memory_reserved(server.device) / 1e9:.1f} GB")
    server.torch.cuda.synchronize()
    return time.perf_counter() - t0
