"""Device-resident three-centre / two-centre engine (the fast path behind the drop-in ABI).

A ``ThreeCenterPlan`` uploads one molecule's primitive pairs and auxiliary primitives to HBM once,
grouped by contracted shell pair and auxiliary shell, builds every pair matrix on the GPU and keeps it
resident; ``run`` then writes all (ij|P) into a device tensor in the packed layout

    J3c_packed[tri(i, j), P],   tri(i, j) = i (i + 1) / 2 + j  for i >= j,

i.e. rows are the i >= j AO pairs and the auxiliary index is contiguous (a column-major
(n_aux, n_ao (n_ao + 1) / 2) matrix, the form a whitening or density contraction consumes).  Element
(P, i, j) of the reference's dense (n_aux, n_ao, n_ao) array is J3c_packed[tri(max, min), P].

PyTorch only provides device memory and the stream; every integral is computed by the HIP library
through its C entry points (eip_pair_matrix, eip_three_center_items, eip_two_center_device).
"""

import ctypes
from dataclasses import dataclass, field
from typing import Dict, List, Tuple

import time as _time

import numpy as np

from . import library

_c_int = ctypes.c_int
_c_ll = ctypes.c_longlong
_c_double = ctypes.c_double
_vp = ctypes.c_void_p


def _ptr(t):
    return _vp(t.data_ptr()) if t is not None else _vp(0)


def _stream_handle(stream):
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return _vp(s.cuda_stream)


_SIDE_STREAMS = {}
# Streams the independent launches of one run are spread over (besides the caller's stream).  Class
# kernels write disjoint output regions, so they may run concurrently; the tail of one launch is then
# filled by the next class instead of idling.  All streams have the default priority.
N_SIDE_STREAMS = 5
# Workgroups of the paced (background) zeroing kernel beside the class kernels; 0 runs the full-speed
# kernel (one workgroup per span).  Plain workload, one box: no zeroing 14.6 ms/step, full-speed 19.4,
# paced 128 or 256 workgroups 18.9, 64 / 512 / 1024 workgroups 19.6-20.7.
ZERO_BLOCKS_BESIDE_COMPUTE = 256


class _fork:
    """Fork ``n`` side streams (default priority, cached per device) from ``stream`` on entry and join
    them back on exit: work queued on a side stream inside the block runs concurrently with the work on
    ``stream``, and everything queued on ``stream`` after the block waits for all of it.  Fork and join
    are event waits, so the pattern is legal inside HIP graph capture.  Yields [] when n == 0."""

    def __init__(self, stream, n):
        self.stream, self.n = stream, n

    def __enter__(self):
        if self.n <= 0:
            self.sides = []
            return []
        import torch
        main = self.stream if self.stream is not None else torch.cuda.current_stream()
        dev = main.device.index if main.device.index is not None else torch.cuda.current_device()
        pool = _SIDE_STREAMS.setdefault(dev, [])
        while len(pool) < self.n:
            pool.append(torch.cuda.Stream(device=dev))
        self.main, self.sides = main, pool[:self.n]
        for s in self.sides:
            s.wait_stream(main)
        return self.sides

    def __exit__(self, *exc):
        for s in self.sides:
            self.main.wait_stream(s)
        return False


def _assign_lpt(costs, n):
    """Longest-processing-time assignment of independent launches to n streams: index lists per stream."""
    load = [0.0] * n
    out = [[] for _ in range(n)]
    for i in sorted(range(len(costs)), key=lambda i: -costs[i]):
        k = min(range(n), key=load.__getitem__)
        out[k].append(i)
        load[k] += costs[i]
    return out


def xcd_logical(n, n_xcd=8):
    """Item index that block b of an n-block launch reads (the kernels' eip::xcd_remap): XCD x = b % 8
    owns the contiguous logical range [x per + min(x, rem), ...)."""
    b = np.arange(n, dtype=np.int64)
    per, rem = divmod(n, n_xcd)
    x = b % n_xcd
    return x * per + np.minimum(x, rem) + b // n_xcd


def longest_first(costs):
    """Order in which to store work items so that the launch dispatches them longest-first: block b
    (dispatched b-th, on XCD b % 8) processes the b-th most expensive item.  Long items then start first
    and short ones fill the tail, instead of one long item finishing alone at the end."""
    costs = np.asarray(costs)
    order = np.argsort(-costs, kind="stable")
    stored = np.empty(len(costs), dtype=np.int64)
    stored[xcd_logical(len(costs))] = order
    return stored


def longest_first_t(costs):
    """longest_first on a torch tensor, on its device (the same permutation; int64 result)."""
    import torch
    n = costs.shape[0]
    order = torch.argsort(-costs.to(torch.int64), stable=True)
    b = torch.arange(n, dtype=torch.int64, device=costs.device)
    per, rem = divmod(n, 8)
    x = b % 8
    stored = torch.empty(n, dtype=torch.int64, device=costs.device)
    stored[x * per + torch.clamp(x, max=rem) + b // 8] = order
    return stored


def three_center_flops_per_triplet(la, lb, lc):
    """Fixed algorithmic flop model of one primitive triplet (FMA = 2):
    Boys 20(L+1)+20, R recursion 2 C(L+4, 4), pair E 9(la+1)(lb+1)(la+lb+2)/2, auxiliary contraction
    sum_k 2 N_herm(la+lb) prod(floor(k_d/2)+1), pair contraction sum_ij 2 ncart(lc) prod(i_d+j_d+1)."""
    from math import comb
    L = la + lb + lc

    def carts(l):
        return [(x, y, l - x - y) for x in range(l, -1, -1) for y in range(l - x, -1, -1)]

    nherm = comb(la + lb + 3, 3)
    f = 20 * (L + 1) + 20 + 2 * comb(L + 4, 4) + 9 * (la + 1) * (lb + 1) * (la + lb + 2) / 2
    f += sum(2 * nherm * (k[0] // 2 + 1) * (k[1] // 2 + 1) * (k[2] // 2 + 1) for k in carts(lc))
    nc = (lc + 1) * (lc + 2) // 2
    f += sum(2 * nc * (i[0] + j[0] + 1) * (i[1] + j[1] + 1) * (i[2] + j[2] + 1) for i in carts(la) for j in carts(lb))
    return f


@dataclass
class PairClass:
    la: int
    lb: int
    n_pp: int
    n_sp: int
    sp_i: np.ndarray
    sp_j: np.ndarray
    sp_pp_off: np.ndarray
    dev: Dict[str, object] = field(default_factory=dict)


@dataclass
class AuxClass:
    lc: int
    n_ks: int
    n_prim: int
    ks_start: np.ndarray
    ks_prim_off: np.ndarray
    dev: Dict[str, object] = field(default_factory=dict)


def _dense_pair_matrix(lib, la, lb, E, device):
    """The dense [n_pp][n_t][16 * blocks] float64 pair matrix from the packed one (eip_pair_matrix_dense),
    on the current stream, which is then synchronised (launches may go to other streams)."""
    import torch
    w = lib.eip_pair_matrix_dense_width(la, lb)
    if w < 0:
        raise ValueError(f"pair class ({la},{lb}) has no matrix-core kernels")
    D = torch.empty((E.shape[0], w), dtype=torch.float64, device=device)
    rc = lib.eip_pair_matrix_dense(la, lb, _ptr(E), int(E.shape[0]), _ptr(D), _stream_handle(None))
    if rc != 0:
        raise RuntimeError(f"eip_pair_matrix_dense({la},{lb}) failed with {rc}")
    torch.cuda.current_stream(device).synchronize()
    return D


def aux_transform_cmax(lc):
    """max_m sum_s |c_m(s)| of the spherical auxiliary transform of angular momentum lc (coefficient
    sqrt(r_m) q_s of output m on Hermite index s; Hobson's theorem: |s| = lc)."""
    from .codegen.tables import aux_transform
    return max(sum(abs(float(q)) for q, _ in terms.values()) * float(r) ** 0.5
               for r, terms in aux_transform(lc, True))


def _bf16_pair_matrix(lib, la, lb, E, device):
    """The split bfloat16 pair matrix of the float32-tier matrix-core kernels from the packed float64 one
    (eip_pair_matrix_bf16), on the current stream, which is then synchronised."""
    import torch
    w = lib.eip_pair_matrix_bf16_width(la, lb)
    if w < 0:
        raise ValueError(f"pair class ({la},{lb}) has no bf16 matrix-core kernels")
    D = torch.empty((E.shape[0], w), dtype=torch.bfloat16, device=device)
    rc = lib.eip_pair_matrix_bf16(la, lb, _ptr(E), int(E.shape[0]), _ptr(D), _stream_handle(None))
    if rc != 0:
        raise RuntimeError(f"eip_pair_matrix_bf16({la},{lb}) failed with {rc}")
    torch.cuda.current_stream(device).synchronize()
    return D


# Primitive pairs with p = a + b below this exponent are "diffuse": the mixed-precision split always
# evaluates them in float64.
DIFFUSE_PAIR_EXPONENT = 0.2


def zero_spans(rows, ld, max_len=16384):
    """Sorted packed rows -> (offset, count) spans of doubles covering them, consecutive rows merged and
    spans cut at max_len (one block each)."""
    rows = np.asarray(rows, dtype=np.int64)
    if rows.size == 0:
        return np.zeros((0, 2), dtype=np.int64)
    brk = np.flatnonzero(np.diff(rows) != 1) + 1
    starts = rows[np.concatenate([[0], brk])]
    lens = np.diff(np.concatenate([[0], brk, [rows.size]]))
    off, cnt = starts * ld, lens * ld
    pieces = -(-cnt // max_len)
    rep_off = np.repeat(off, pieces) + max_len * (np.arange(pieces.sum()) - np.repeat(np.cumsum(pieces) - pieces, pieces))
    end = np.repeat(off + cnt, pieces)
    return np.stack([rep_off, np.minimum(end - rep_off, max_len)], axis=1)


def group_pairs(pd):
    """Stable grouping of one class's primitive pairs by (i_ao_start, j_ao_start); within a shell pair
    the diffuse pairs (p < DIFFUSE_PAIR_EXPONENT) come first, each part in list order (descending bound)."""
    order = np.lexsort((pd.P_p[:, 3] >= DIFFUSE_PAIR_EXPONENT, pd.j_ao_start, pd.i_ao_start))
    ki, kj = pd.i_ao_start[order], pd.j_ao_start[order]
    new = np.ones(len(order), dtype=bool)
    new[1:] = (ki[1:] != ki[:-1]) | (kj[1:] != kj[:-1])
    starts = np.flatnonzero(new)
    off = np.append(starts, len(order)).astype(np.int32)
    return order, ki[starts].astype(np.int32), kj[starts].astype(np.int32), off


def group_aux(pk):
    order = np.argsort(pk.i_ao_start, kind="stable")
    k = pk.i_ao_start[order]
    new = np.ones(len(order), dtype=bool)
    new[1:] = k[1:] != k[:-1]
    starts = np.flatnonzero(new)
    return order, k[starts].astype(np.int32), np.append(starts, len(order)).astype(np.int32)


def mixed_split(sp_pp_off, p, ub, q, ks_prim_off, threshold, aux_order="natural"):
    """Float64/float32 split of one (pair class, auxiliary class); host logic of setup_mixed.

    sp_pp_off: CSR offsets of the shell pairs' primitive pairs (diffuse pairs first, then descending
    ub, as group_pairs orders them); p, ub: exponent sum and reference bound per primitive pair;
    q: Q_P per auxiliary shell; ks_prim_off: CSR primitive offsets of the auxiliary shells.
    aux_order "qp" groups the auxiliary shells by descending Q_P (perm = that order), "natural" keeps
    their order (perm = identity; a group's bound is then the largest Q_P among its 64 shells).
    Returns (s, perm, gprim): s[sp, g] = number of leading primitive pairs of shell pair sp evaluated in
    float64 for auxiliary group g (shells perm[64 g : 64 g + 64]), and gprim[g] = auxiliary primitives
    in group g.  A primitive pair is float64 for group g when it is diffuse or
    ub * max_{P in g} Q_P >= threshold.
    """
    return mixed_split_groups(mixed_split_pairs(sp_pp_off, p, ub, threshold), q, ks_prim_off, aux_order)


def mixed_split_pairs(sp_pp_off, p, ub, threshold):
    """The auxiliary-independent half of mixed_split for one pair class (computed once per class)."""
    sp_pp_off = np.asarray(sp_pp_off, dtype=np.int64)
    n_sp = len(sp_pp_off) - 1
    sp_of_pp = np.repeat(np.arange(n_sp, dtype=np.int64), np.diff(sp_pp_off))
    diffuse = np.asarray(p) < DIFFUSE_PAIR_EXPONENT
    nd = ~diffuse
    with np.errstate(divide="ignore"):
        need = threshold / np.asarray(ub, dtype=np.float64)[nd]
    return {"n_sp": n_sp, "n_diff": np.bincount(sp_of_pp[diffuse], minlength=n_sp),
            "sp_nd": sp_of_pp[nd], "neg_need": -need}


def mixed_split_groups(pre, q, ks_prim_off, aux_order="natural"):
    """mixed_split for one auxiliary class, from the pair class's mixed_split_pairs."""
    if aux_order not in ("qp", "natural"):
        raise ValueError("aux_order is 'qp' or 'natural'")
    n_sp = pre["n_sp"]
    q = np.asarray(q, dtype=np.float64)
    perm = (np.argsort(-q, kind="stable") if aux_order == "qp" else np.arange(len(q))).astype(np.int32)
    G = (len(q) + 63) // 64
    starts = np.arange(G) * 64
    qmax = np.maximum.reduceat(q[perm], starts)
    kprim = np.diff(np.asarray(ks_prim_off, dtype=np.int64))[perm]
    gprim = np.add.reduceat(kprim, starts)
    # The count for group g depends only on qmax[g]: work on the groups sorted by descending qmax.
    # g_min: the pair is float64 for the sorted groups h < g_min (qmax_h >= need there).
    by_q = np.argsort(-qmax, kind="stable")
    g_min = np.searchsorted(-qmax[by_q], pre["neg_need"], side="right")
    hist = np.bincount(pre["sp_nd"] * (G + 1) + g_min, minlength=n_sp * (G + 1)).reshape(n_sp, G + 1)
    at_least = np.cumsum(hist[:, ::-1], axis=1)[:, ::-1]        # at_least[:, h] = #(g_min >= h)
    s = np.empty((n_sp, G), dtype=np.int64)
    s[:, by_q] = pre["n_diff"][:, None] + at_least[:, 1:]
    return s, perm, gprim


def mixed_split_pairs_t(sp_pp_off, p, ub, threshold, device):
    """mixed_split_pairs with the per-primitive-pair arrays on ``device`` (torch; same values)."""
    import torch
    off = torch.as_tensor(np.asarray(sp_pp_off, dtype=np.int64), device=device)
    n_sp = off.shape[0] - 1
    sp_of_pp = torch.arange(n_sp, dtype=torch.int64, device=device).repeat_interleave(off[1:] - off[:-1])
    diffuse = torch.as_tensor(np.asarray(p, dtype=np.float64), device=device) < DIFFUSE_PAIR_EXPONENT
    nd = ~diffuse
    ub_nd = torch.as_tensor(np.asarray(ub, dtype=np.float64), device=device)[nd]
    # a zero bound needs Q_P = inf, or nothing at threshold 0 (numpy's 0/0 = nan sorts last: float64)
    need = torch.where(ub_nd > 0, threshold / ub_nd, torch.full_like(ub_nd, float("inf") if threshold > 0 else 0.0))
    return {"n_sp": n_sp, "n_diff": torch.bincount(sp_of_pp[diffuse], minlength=n_sp),
            "sp_nd": sp_of_pp[nd], "neg_need": -need}


def mixed_split_groups_t(pre, q, ks_prim_off, aux_order="natural"):
    """mixed_split_groups from mixed_split_pairs_t: s as an (n_sp, G) int64 tensor on the pairs'
    device; perm and gprim (one entry per auxiliary shell or group) stay numpy."""
    import torch
    if aux_order not in ("qp", "natural"):
        raise ValueError("aux_order is 'qp' or 'natural'")
    n_sp, device = pre["n_sp"], pre["sp_nd"].device
    q = np.asarray(q, dtype=np.float64)
    perm = (np.argsort(-q, kind="stable") if aux_order == "qp" else np.arange(len(q))).astype(np.int32)
    G = (len(q) + 63) // 64
    starts = np.arange(G) * 64
    qmax = np.maximum.reduceat(q[perm], starts)
    gprim = np.add.reduceat(np.diff(np.asarray(ks_prim_off, dtype=np.int64))[perm], starts)
    by_q = np.argsort(-qmax, kind="stable")
    g_min = torch.searchsorted(torch.as_tensor(-qmax[by_q], device=device), pre["neg_need"], right=True)
    hist = torch.bincount(pre["sp_nd"] * (G + 1) + g_min, minlength=n_sp * (G + 1)).reshape(n_sp, G + 1)
    at_least = torch.flip(torch.cumsum(torch.flip(hist, [1]), 1), [1])
    s = torch.empty((n_sp, G), dtype=torch.int64, device=device)
    s[:, torch.as_tensor(by_q, device=device)] = pre["n_diff"][:, None] + at_least[:, 1:]
    return s, perm, gprim


_TUNING_APPLIED = False


def _ensure_tuning():
    """Apply the shipped chunk-plan variant table once per process."""
    global _TUNING_APPLIED
    if not _TUNING_APPLIED:
        _TUNING_APPLIED = True
        apply_tuning()


class ThreeCenterPlan:
    """All classes of one molecule's (ij|P), resident on one GPU."""

    def __init__(self, molecule, pair_data, aux_data, omega=0.0, device="cuda"):
        import torch
        self.torch = torch
        self.lib = library()
        _ensure_tuning()
        self.n_ao, self.n_aux = molecule.n_ao, molecule.n_aux
        self.omega = float(omega)
        self.device = torch.device(device)
        if not molecule.spherical_basis:
            raise ValueError("the device path produces spherical output")
        lib = self.lib
        T = lambda a, dt: torch.as_tensor(np.ascontiguousarray(a), dtype=dt, device=self.device)
        self.pairs: List[PairClass] = []
        self._pair_orders = []
        self._pair_p = []
        stream = _stream_handle(None)
        for (la, lb), pd in pair_data.items():
            nnz = lib.eip_pair_matrix_nnz(la, lb)
            if nnz < 0:
                raise ValueError(f"no specialised kernels for pair class ({la},{lb})")
            order, si, sj, off = group_pairs(pd)
            pc = PairClass(la, lb, len(order), len(si), si, sj, off)
            self._pair_orders.append(order)
            self._pair_p.append(pd.P_p[order, 3].copy())
            d = pc.dev
            d["P"] = T(pd.P_p[order], torch.float64)
            d["A"] = T(pd.A_a[order], torch.float64)
            d["B"] = T(pd.B_b[order], torch.float64)
            d["c"] = T(pd.coefficient[order], torch.float64)
            d["off"] = T(off, torch.int32)
            d["si"] = T(si, torch.int32)
            d["sj"] = T(sj, torch.int32)
            d["E"] = torch.empty((pc.n_pp, nnz), dtype=torch.float64, device=self.device)
            d["E32"] = torch.empty((pc.n_pp, nnz), dtype=torch.float32, device=self.device)
            rc = lib.eip_pair_matrix(la, lb, _ptr(d["P"]), _ptr(d["A"]), _ptr(d["B"]), _ptr(d["c"]), pc.n_pp,
                                     _ptr(d["E"]), _ptr(d["E32"]), stream)
            if rc != 0:
                raise RuntimeError(f"eip_pair_matrix({la},{lb}) failed with {rc}")
            self.pairs.append(pc)
        self.aux: List[AuxClass] = []
        for lc, pk in aux_data.items():
            order, ks, off = group_aux(pk)
            ac = AuxClass(lc, len(ks), len(order), ks, off)
            d = ac.dev
            d["C"] = T(pk.A_a[order], torch.float64)
            d["c"] = T(pk.coefficient[order], torch.float64)
            d["off"] = T(off, torch.int32)
            d["col"] = T(ks, torch.int32)
            self.aux.append(ac)
        # work items: every (shell pair, block of 64 auxiliary shells), in shell-pair order (sliced by
        # shell-pair ranges when streaming), built on the device (a 250-atom molecule has ~2e8 items);
        # whole-class launches read them stored longest-first (items_lpt, built on first use)
        self.items: Dict[Tuple[int, int], object] = {}
        self.items_lpt: Dict[Tuple[int, int], object] = {}
        for pi, pc in enumerate(self.pairs):
            for ai, ac in enumerate(self.aux):
                kb = (ac.n_ks + 63) // 64
                sp = torch.arange(pc.n_sp, dtype=torch.int32, device=self.device).repeat_interleave(kb)
                k0 = (torch.arange(kb, dtype=torch.int32, device=self.device) * 64).repeat(pc.n_sp)
                self.items[(pi, ai)] = torch.stack([sp, k0], 1)
        self._zero = None
        self._job_tables = {}
        self._dense = {}
        self._bf16 = {}
        self._flags = {}
        # dense pair matrices of the classes whose float64 kernels run on the matrix cores (and split
        # bfloat16 ones where the float32 kernels do), built here (on the caller's stream) rather than at
        # a first launch
        for pi in range(len(self.pairs)):
            for ai in range(len(self.aux)):
                self.kernel_E(pi, ai)
                self.kernel_E32(pi, ai)
        torch.cuda.current_stream().synchronize()

    def dense_E(self, pi):
        """Dense float64 pair matrix of pair class pi (the matrix-core kernels' layout), built on first
        use and kept."""
        d = self._dense.get(pi)
        if d is None:
            d = _dense_pair_matrix(self.lib, self.pairs[pi].la, self.pairs[pi].lb, self.pairs[pi].dev["E"], self.device)
            self._dense[pi] = d
        return d

    def kernel_E(self, pi, ai):
        """The float64 pair matrix the current kernels of class (pi, ai) read: dense when that class runs
        on the matrix cores, packed otherwise."""
        pc, ac = self.pairs[pi], self.aux[ai]
        if self.lib.eip_three_center_uses_mfma(pc.la, pc.lb, ac.lc) == 1:
            return self.dense_E(pi)
        return pc.dev["E"]

    def bf16_E(self, pi):
        """Split bfloat16 pair matrix of pair class pi (the float32-tier matrix-core kernels' layout),
        built on first use and kept."""
        d = self._bf16.get(pi)
        if d is None:
            pc = self.pairs[pi]
            d = _bf16_pair_matrix(self.lib, pc.la, pc.lb, pc.dev["E"], self.device)
            self._bf16[pi] = d
        return d

    def kernel_E32(self, pi, ai):
        """The pair matrix the current float32 kernels of class (pi, ai) read: the split bfloat16 one for
        bf16 matrix-core variants, the packed float32 one otherwise."""
        pc, ac = self.pairs[pi], self.aux[ai]
        if self.lib.eip_three_center_f32_uses_bf16(pc.la, pc.lb, ac.lc) == 1:
            return self.bf16_E(pi)
        return pc.dev["E32"]

    def _zero_tables(self):
        """(rows, spans) of the packed rows no kernel writes (shell pairs screened out entirely), which
        run() zeroes; built on first use, since the compact block outputs never need them."""
        if self._zero is None:
            torch = self.torch
            written = np.zeros(self.n_packed, dtype=bool)
            for pc in self.pairs:
                ni, nj = 2 * pc.la + 1, 2 * pc.lb + 1
                i = (pc.sp_i[:, None, None] + np.arange(ni)[None, :, None]).astype(np.int64)
                j = (pc.sp_j[:, None, None] + np.arange(nj)[None, None, :]).astype(np.int64)
                hi, lo = np.maximum(i, j), np.minimum(i, j)
                written[(hi * (hi + 1) // 2 + lo).ravel()] = True
            rows = np.flatnonzero(~written)
            self._zero = (torch.as_tensor(rows.astype(np.int32), device=self.device),
                          torch.as_tensor(zero_spans(rows, self.n_aux), dtype=torch.int64, device=self.device))
        return self._zero

    @property
    def zero_rows(self):
        return self._zero_tables()[0]

    @property
    def zero_spans(self):
        return self._zero_tables()[1]

    # -- sizes -----------------------------------------------------------------------------------
    @property
    def n_packed(self):
        return self.n_ao * (self.n_ao + 1) // 2

    def n_integrals(self):
        return self.n_aux * self.n_packed

    def primitive_triplets(self):
        """{(la, lb, lc): number of primitive triplets}."""
        return {(pc.la, pc.lb, ac.lc): pc.n_pp * ac.n_prim for pc in self.pairs for ac in self.aux}

    def model_flops(self):
        return sum(n * three_center_flops_per_triplet(*k) for k, n in self.primitive_triplets().items())

    def device_bytes(self):
        """Device memory held by the plan (pair and auxiliary data, pair matrices, work items)."""
        def size(v):
            if hasattr(v, "numel"):
                return v.numel() * v.element_size()
            if isinstance(v, dict):
                return sum(size(x) for x in v.values())
            if isinstance(v, (list, tuple)):
                return sum(size(x) for x in v)
            return 0
        total = sum(size(pc.dev) for pc in self.pairs) + sum(size(ac.dev) for ac in self.aux) + size(self._dense)
        total += size(self.items) + size(self.items_lpt) + size(getattr(self, "mixed", {})) + size(self._zero or ())
        return int(total)

    # -- execution -------------------------------------------------------------------------------
    def new_output(self):
        return self.torch.empty((self.n_packed, self.n_aux), dtype=self.torch.float64, device=self.device)

    def class_items(self, pi, ai):
        """Items of a whole-class launch, stored longest-first; built on first use and kept (the first
        call synchronises the current stream, since the launch may go to another stream)."""
        it = self.items_lpt.get((pi, ai))
        if it is None:
            torch = self.torch
            pc, ac = self.pairs[pi], self.aux[ai]
            kb = (ac.n_ks + 63) // 64
            # cost of an item: primitive pairs x the block's largest auxiliary primitive count
            n_pp_sp = np.diff(np.asarray(pc.sp_pp_off, dtype=np.int64))
            nkp = np.diff(ac.ks_prim_off.astype(np.int64))
            kmax = np.maximum.reduceat(nkp, np.arange(kb) * 64) if len(nkp) else np.ones(kb, np.int64)
            cost = (torch.as_tensor(n_pp_sp, device=self.device).repeat_interleave(kb)
                    * torch.as_tensor(kmax, dtype=torch.int64, device=self.device).repeat(pc.n_sp))
            it = self.items[(pi, ai)].index_select(0, longest_first_t(cost))
            if self.device.type == "cuda":
                torch.cuda.current_stream(self.device).synchronize()
            self.items_lpt[(pi, ai)] = it
        return it

    def launch_class(self, pi, ai, out, stream=None, use_f32=False):
        pc, ac = self.pairs[pi], self.aux[ai]
        it = self.class_items(pi, ai)
        dp, da = pc.dev, ac.dev
        E = dp["E"] if use_f32 else self.kernel_E(pi, ai)
        rc = self.lib.eip_three_center_items(
            pc.la, pc.lb, ac.lc, int(use_f32), _ptr(it), _vp(0), it.shape[0], _ptr(dp["off"]), _ptr(dp["si"]),
            _ptr(dp["sj"]), _ptr(dp["P"]), _ptr(E), _ptr(self.kernel_E32(pi, ai)), _vp(0), _ptr(da["off"]),
            _ptr(da["col"]), _ptr(da["C"]), _ptr(da["c"]), ac.n_ks, _ptr(out), _c_ll(self.n_aux), 0,
            _c_double(self.omega), _stream_handle(stream))
        if rc != 0:
            raise RuntimeError(f"eip_three_center_items({pc.la},{pc.lb},{ac.lc}) failed with {rc}")

    def zero_unwritten(self, out, stream=None, max_blocks=0):
        """Zero the packed rows of shell pairs that are screened out (every other element is written
        exactly once by the class kernels); ``max_blocks`` caps the grid (0: one block per span)."""
        rc = self.lib.eip_zero_spans(_ptr(self.zero_spans), int(self.zero_spans.shape[0]), _ptr(out),
                                     int(max_blocks), _stream_handle(stream))
        if rc != 0:
            raise RuntimeError(f"eip_zero_spans failed with {rc}")

    def run(self, out=None, stream=None, zero=True, concurrent=True, metric_out=None):
        """All classes into ``out`` (packed layout); rows of screened-out shell pairs are zero.  With
        ``metric_out`` (a zeroed or reused (n_aux, n_aux) float64 tensor) the metric (P|Q) is computed
        in the same call, on one of the streams.

        With ``concurrent`` the zeroing (HBM-write-bound) runs on a side stream beside the class kernels
        (FP64-VALU-bound), and the class launches are spread over N_SIDE_STREAMS + 1 streams by
        estimated cost: every launch writes output elements no other launch writes (its pair class's
        rows, its auxiliary class's columns; zeroed rows belong to no class).  ``stream`` waits for all
        of them before the call returns, so later work on it is ordered as after a single-stream run."""
        if out is None:
            out = self.new_output()
        zero = zero and self.zero_spans.shape[0] > 0
        for pi in range(len(self.pairs)):                 # dispatch tables exist before any launch
            for ai in range(len(self.aux)):
                self.class_items(pi, ai)
        if not concurrent:
            if zero:
                self.zero_unwritten(out, stream)
            if metric_out is not None:
                self.two_center(metric_out, stream)
            for pi in range(len(self.pairs)):
                for ai in range(len(self.aux)):
                    self.launch_class(pi, ai, out, stream)
            return out
        keys = [(pi, ai) for pi in range(len(self.pairs)) for ai in range(len(self.aux))]
        costs = [self.pairs[pi].n_pp * self.aux[ai].n_prim *
                 three_center_flops_per_triplet(self.pairs[pi].la, self.pairs[pi].lb, self.aux[ai].lc)
                 for pi, ai in keys]
        with _fork(stream, N_SIDE_STREAMS + int(zero)) as sides:
            if zero:
                self.zero_unwritten(out, sides[0], ZERO_BLOCKS_BESIDE_COMPUTE)
            lanes = [stream] + sides[int(zero):]
            if metric_out is not None:
                # the metric goes first on the last lane (LPT gives it the least class work)
                self.two_center(metric_out, lanes[-1])
            for k, idx in enumerate(_assign_lpt(costs, len(lanes))):
                for i in idx:
                    self.launch_class(*keys[i], out, lanes[k])
        return out

    # -- (P|Q) ------------------------------------------------------------------------------------
    def two_center(self, out=None, stream=None):
        """Dense (n_aux, n_aux) float64 (P|Q) on the device (always FP64: the metric is factorised)."""
        torch = self.torch
        if out is None:
            out = torch.zeros((self.n_aux, self.n_aux), dtype=torch.float64, device=self.device)
        for a in self.aux:
            for b in self.aux:
                if a.lc > b.lc:
                    continue
                rc = self.lib.eip_two_center_device(
                    a.lc, b.lc, _ptr(a.dev["off"]), _ptr(a.dev["C"]), _ptr(a.dev["c"]), _ptr(a.dev["col"]), a.n_ks,
                    _ptr(b.dev["off"]), _ptr(b.dev["C"]), _ptr(b.dev["c"]), _ptr(b.dev["col"]), b.n_ks,
                    int(a.lc == b.lc), 1, _ptr(out), _c_ll(self.n_aux), _c_double(self.omega), _stream_handle(stream))
                if rc != 0:
                    raise RuntimeError(f"eip_two_center_device({a.lc},{b.lc}) failed with {rc}")
        return out

    # -- generic work launches (job table) ---------------------------------------------------------
    def mean_pairs_per_shell_pair(self, pi):
        """Mean primitive pairs per shell pair of pair class pi (a work item's relative cost)."""
        pc = self.pairs[pi]
        return float(pc.n_pp) / max(1, pc.n_sp)

    def job_words(self, pi, ai, out_ptr, ld, perm=None, vector=False):
        """The 16 words of one eip_j3c_job entry: pair class pi, auxiliary class ai, output.  vector: the
        packed pair matrices the vector kernels (and the persistent queue kernels) read."""
        dp, da = self.pairs[pi].dev, self.aux[ai].dev
        # E: what the float64 kernels of the class read (dense for matrix-core variants); E32: what the
        # float32 kernels read (split bfloat16 for bf16 matrix-core variants)
        E = dp["E"] if vector else self.kernel_E(pi, ai)
        E32 = dp["E32"] if vector else self.kernel_E32(pi, ai)
        return [dp["off"].data_ptr(), dp["si"].data_ptr(), dp["sj"].data_ptr(), dp["P"].data_ptr(),
                E.data_ptr(), E32.data_ptr(),
                perm.data_ptr() if perm is not None else 0,
                da["off"].data_ptr(), da["col"].data_ptr(), da["C"].data_ptr(), da["c"].data_ptr(),
                self.aux[ai].n_ks, int(out_ptr), int(ld), 0, 0]

    def job_table(self, entries):
        """Device table (int64 [n][16]) of eip_j3c_job entries; kept (and reused) for the plan's lifetime,
        so a table is never freed while launches that read it may be in flight."""
        key = tuple(int(w) for e in entries for w in e)
        t = self._job_tables.get(key)
        if t is None:
            t = self.torch.tensor(np.array(key, dtype=np.int64).reshape(-1, 16), dtype=self.torch.int64,
                                  device=self.device)
            self._job_tables[key] = t
        return t

    def launch_work(self, pi, ai, items, jobs, job=None, pp=None, ks_end=None, layout=0, accumulate=0,
                    use_f32=False, stream=None):
        if items is None or items.shape[0] == 0:
            return
        pc, ac = self.pairs[pi], self.aux[ai]
        rc = self.lib.eip_three_center_work(
            pc.la, pc.lb, ac.lc, int(use_f32), _ptr(items), _ptr(ks_end), _ptr(pp), _ptr(job), int(items.shape[0]),
            _ptr(jobs), int(layout), int(accumulate), _c_double(self.omega), _stream_handle(stream))
        if rc != 0:
            raise RuntimeError(f"eip_three_center_work({pc.la},{pc.lb},{ac.lc}, f32={use_f32}) failed with {rc}")

    def flag_buffer(self, n, stream=None):
        """Item flags of checked mixed precision: at least n int32 per stream, grown on demand and kept
        (a launch that reads them may still be in flight when the next one is queued)."""
        key = _stream_handle(stream).value
        f = self._flags.get(key)
        if f is None or f.numel() < n:
            f = self.torch.empty(max(n, 1 << 20), dtype=self.torch.int32, device=self.device)
            self._flags[key] = f
        return f

    def launch_checked(self, pi, ai, items, jobs, layout, stream=None):
        """Checked mixed precision over ``items``: the float32 kernels store every item and flag the items
        where some |value| reached the threshold; the float64 kernels then recompute exactly those."""
        if items is None or items.shape[0] == 0:
            return
        pc, ac = self.pairs[pi], self.aux[ai]
        if self.lib.eip_three_center_f32_uses_bf16(pc.la, pc.lb, ac.lc) != 1:
            # the flags come from the bf16 matrix-core kernels: a class whose float32 tier runs on the
            # vector kernels is evaluated in float64
            self.launch_work(pi, ai, items, jobs, layout=layout, stream=stream)
            return
        stride = self.lib.eip_three_center_flag_stride(pc.la, pc.lb, ac.lc)
        n = int(items.shape[0])
        flags = self.flag_buffer(n * stride, stream)
        for pass_ in (0, 1):
            rc = self.lib.eip_three_center_work_checked(
                pc.la, pc.lb, ac.lc, pass_, _ptr(items), _vp(0), _vp(0), n, _ptr(jobs), int(layout), _ptr(flags),
                int(stride), _c_double(self.mixed_stats["threshold"]), _c_double(self.omega), _stream_handle(stream))
            if rc != 0:
                raise RuntimeError(f"eip_three_center_work_checked({pc.la},{pc.lb},{ac.lc}, pass {pass_}) failed with {rc}")
        tally = self.__dict__.get("checked_tally")
        if tally is not None:
            # diagnostics (off unless a dict is installed): triplets of all and of the flagged items
            torch = self.torch
            f = flags[:n * stride].view(n, stride).ne(0).any(1)
            n_pp_sp = self.__dict__.setdefault("_n_pp_sp", {})
            if pi not in n_pp_sp:
                n_pp_sp[pi] = torch.as_tensor(np.diff(np.asarray(pc.sp_pp_off, dtype=np.int64)), device=self.device)
            it = items.long()
            w = n_pp_sp[pi][it[:, 0]] * (ac.n_ks - it[:, 1]).clamp(max=64)
            key = (pc.la, pc.lb, ac.lc)
            t = tally.setdefault(key, [torch.zeros((), dtype=torch.int64, device=self.device),
                                       torch.zeros((), dtype=torch.int64, device=self.device)])
            t[0] += w.sum()
            t[1] += (w * f).sum()

    def checked_fraction(self, pi, ai, items, stream=None):
        """Fraction of ``items`` the last checked launch of class (pi, ai) on ``stream`` flagged for
        float64 (reads the flag buffer; synchronises)."""
        pc, ac = self.pairs[pi], self.aux[ai]
        stride = self.lib.eip_three_center_flag_stride(pc.la, pc.lb, ac.lc)
        n = int(items.shape[0])
        f = self.flag_buffer(n * stride, stream)[:n * stride].view(n, stride)
        return float((f != 0).any(1).float().mean().item()) if n else 0.0

    # -- mixed precision ---------------------------------------------------------------------------
    def pair_esum(self, pi):
        """[n_pp][la + lb + 1] float32 of pair class pi: Esum[pp][n] = max_ij sum_{|t| = n} |E_pp[ij][t]|,
        the pair matrix's row sums per Hermite level (from the float64 matrix; built on first use)."""
        cache = self.__dict__.setdefault("_esum", {})
        if pi in cache:
            return cache[pi]
        from .codegen.gen_kernels import pair_layout
        torch = self.torch
        pc = self.pairs[pi]
        lay = pair_layout(pc.la, pc.lb, True)
        nlev, nij = pc.la + pc.lb + 1, (2 * pc.la + 1) * (2 * pc.lb + 1)
        group = torch.as_tensor([row * nlev + sum(t) for t, row in lay], dtype=torch.int64, device=self.device)
        E = pc.dev["E"]
        out = torch.empty((pc.n_pp, nlev), dtype=torch.float32, device=self.device)
        step = 1 << 16
        for c0 in range(0, pc.n_pp, step):
            e = E[c0:c0 + step, :len(lay)].abs()
            S = torch.zeros((e.shape[0], nij * nlev), dtype=torch.float64, device=self.device).index_add_(1, group, e)
            out[c0:c0 + step] = S.view(-1, nij, nlev).amax(1).float()
        cache[pi] = out
        return out

    def item_scales(self, pi, ai, items=None, stream=None):
        """A-priori term scale (eip_three_center_scale) of the work items of class (pi, ai) -- by default
        every item, in the order of ``self.items`` -- as a float32 tensor."""
        torch = self.torch
        pc, ac = self.pairs[pi], self.aux[ai]
        items = self.items[(pi, ai)] if items is None else items
        out = torch.empty(items.shape[0], dtype=torch.float32, device=self.device)
        rc = self.lib.eip_three_center_scale(
            pc.la, pc.lb, ac.lc, _ptr(items), int(items.shape[0]), _ptr(pc.dev["off"]), _ptr(pc.dev["P"]),
            _ptr(self.pair_esum(pi)), _ptr(ac.dev["off"]), _ptr(ac.dev["C"]), _ptr(ac.dev["c"]), ac.n_ks,
            _c_double(aux_transform_cmax(ac.lc)), _c_double(self.omega), _ptr(out), _stream_handle(stream))
        if rc != 0:
            raise RuntimeError(f"eip_three_center_scale({pc.la},{pc.lb},{ac.lc}) failed with {rc}")
        return out

    def setup_mixed(self, pair_list, q_aux, threshold=1e-6, granularity="item", aux_order="natural",
                    estimate="schwarz"):
        """Split every primitive triplet between the float64 and float32 kernels.

        Bound of a triplet (primitive pair pp, auxiliary shell P): ub_pp Q_P, with ub_pp the reference's
        Schwarz-type bound of the primitive pair (pair_utility.py:95-114) and Q_P = sqrt(max_m (P_m|P_m))
        from the float64 metric (``q_aux``: one array per auxiliary class, in shell order).  Triplets
        with ub_pp Q_P < threshold run in float32 (Boys, R and accumulation); pairs with p <
        DIFFUSE_PAIR_EXPONENT always run in float64.  Auxiliary shells are ordered by descending Q_P and
        taken 64 at a time; for a (shell pair, 64 auxiliary shells) item the bound uses the largest Q_P
        of the 64.  Within a shell pair the diffuse pairs come first and the rest by descending ub, so
        the float64 pairs of an item are a prefix [pp0, pp0 + s) and the float32 pairs the suffix: the
        float64 launch stores its prefix sums (zeros when s = 0), the float32 launch adds the rest.
        ``pair_list``: per class ``PairArrays`` or ``list[PrimitivePair]`` (the list the pair data was
        packed from).  Returns triplet counts.

        granularity "pair" splits inside an item as above (two launches over the item, the second adding
        into the output); "item" (default) evaluates an item in float32 only when all its primitive pairs
        qualify (s = 0) and in float64 otherwise, so every item is one wave that stores its result once.

        aux_order "natural" (default) keeps the auxiliary shells' order, so an item's 64 lanes write
        adjacent output columns as in the plain path and streamed blocks dispatch their items
        longest-first; "qp" groups them by descending Q_P (more float32 items, scattered columns).
        Measured on the 200-atom workload: natural 1.06x the float64 stream, qp 0.88x.
        """
        if granularity not in ("item", "pair", "checked"):
            raise ValueError("granularity is 'item', 'pair' or 'checked'")
        if granularity == "checked":
            return self.setup_checked(threshold)
        if estimate not in ("schwarz", "envelope"):
            raise ValueError("estimate is 'schwarz' or 'envelope'")
        if estimate == "envelope":
            if granularity != "item" or aux_order != "natural":
                raise ValueError("the envelope estimate splits whole items in the natural auxiliary order")
            return self.setup_envelope(threshold)
        from .pair_utility import _as_arrays
        torch = self.torch
        T = lambda a, dt: torch.as_tensor(np.ascontiguousarray(a), dtype=dt, device=self.device)
        self.mixed = {}
        self.mixed_aux_order = aux_order
        # block tables of an earlier split must not be reused: a new split is a new generation
        self._mixed_generation = self.__dict__.get("_mixed_generation", 0) + 1
        self.mixed_stats = {"f64_triplets": 0, "f32_triplets": 0, "threshold": threshold,
                            "aux_order": aux_order, "granularity": granularity}
        dev = self.device
        f64_total = torch.zeros((), dtype=torch.int64, device=dev)
        all_total = torch.zeros((), dtype=torch.int64, device=dev)
        # Item tables are built on the device: a 250-atom job has ~2e8 (shell pair, 64 auxiliary
        # shells) items, whose host construction took ~20 s.  off64/off32[sp] = items of shell pairs
        # before sp in each table (tables are in shell-pair order), so block slices need no search.
        for pi, pc in enumerate(self.pairs):
            ub = _as_arrays(pair_list[(pc.la, pc.lb)]).upper_bound[self._pair_orders[pi]]
            pre = mixed_split_pairs_t(pc.sp_pp_off, self._pair_p[pi], ub, threshold, dev)
            off = torch.as_tensor(np.asarray(pc.sp_pp_off, dtype=np.int64), device=dev)
            n_pp_sp = off[1:] - off[:-1]
            for ai, ac in enumerate(self.aux):
                s, perm, gprim = mixed_split_groups_t(pre, q_aux[ai], ac.ks_prim_off, aux_order)
                G = len(gprim)
                gp = torch.as_tensor(gprim, dtype=torch.int64, device=dev)
                sp = torch.arange(pc.n_sp, dtype=torch.int32, device=dev).repeat_interleave(G)
                k0 = (torch.arange(G, dtype=torch.int32, device=dev) * 64).repeat(pc.n_sp)
                items = torch.stack([sp, k0], 1)
                npp = n_pp_sp.repeat_interleave(G)                      # pp1 - pp0 per item
                srow = s.reshape(-1)
                trip = npp * gp.repeat(pc.n_sp)                         # triplets per item
                all_total += trip.sum()

                def offsets(mask):
                    per_sp = mask.reshape(pc.n_sp, G).sum(1) if G else torch.zeros(pc.n_sp, dtype=torch.int64, device=dev)
                    return np.concatenate([[0], np.cumsum(per_sp.cpu().numpy())]).astype(np.int64)

                if granularity == "pair":
                    f32 = srow < npp
                    pp0 = off[:-1].repeat_interleave(G)
                    split = pp0 + srow
                    f64_total += (s * gp[None, :]).sum()
                    m = dict(it64=items, pp64=torch.stack([pp0, split], 1).to(torch.int32),
                             off64=np.arange(pc.n_sp + 1, dtype=np.int64) * G, it32=items[f32],
                             pp32=torch.stack([split[f32], (pp0 + npp)[f32]], 1).to(torch.int32),
                             off32=offsets(f32), acc32=1, cost64=srow, cost32=(npp - srow)[f32])
                else:
                    f32 = srow == 0
                    f64_total += trip[~f32].sum()
                    m = dict(it64=items[~f32], pp64=None, off64=offsets(~f32), it32=items[f32], pp32=None,
                             off32=offsets(f32), acc32=0, cost64=npp[~f32], cost32=npp[f32])
                self.mixed[(pi, ai)] = dict(perm=T(perm, torch.int32), G=G, **m)
        f64_trip, all_trip = int(f64_total.item()), int(all_total.item())
        self.mixed_stats["f64_triplets"] = f64_trip
        self.mixed_stats["f32_triplets"] = all_trip - f64_trip
        return self.mixed_stats

    def setup_envelope(self, threshold=1e-6):
        """Mixed precision split on the a-priori term scale of every work item (``item_scales``): an
        item runs in float32 when its scale -- a bound of the magnitude of every term summed into its
        outputs, distance decay and the cancellation between a contracted shell pair's primitive pairs
        included -- is below ``threshold`` and none of its shell pair's primitive pairs is diffuse
        (p < DIFFUSE_PAIR_EXPONENT); otherwise in float64.  The float32 error of an item is then a few
        float32 roundings of the threshold.  Items keep the plain path's order (natural auxiliary order),
        so the split plugs into the item-granularity launches."""
        torch = self.torch
        dev = self.device
        self.mixed = {}
        self.mixed_aux_order = "natural"
        self._mixed_generation = self.__dict__.get("_mixed_generation", 0) + 1
        f64_total = torch.zeros((), dtype=torch.int64, device=dev)
        all_total = torch.zeros((), dtype=torch.int64, device=dev)
        t0 = _time.perf_counter()
        for pi, pc in enumerate(self.pairs):
            off_np = np.asarray(pc.sp_pp_off, dtype=np.int64)
            diffuse_sp = np.zeros(pc.n_sp, dtype=bool)
            dpp = np.flatnonzero(self._pair_p[pi] < DIFFUSE_PAIR_EXPONENT)
            if dpp.size:
                diffuse_sp[np.searchsorted(off_np, dpp, side="right") - 1] = True
            diffuse_t = torch.as_tensor(diffuse_sp, device=dev)
            n_pp_sp = torch.as_tensor(np.diff(off_np), device=dev)
            for ai, ac in enumerate(self.aux):
                items = self.items[(pi, ai)]
                G = (ac.n_ks + 63) // 64
                scale = self.item_scales(pi, ai, items)
                sp = items[:, 0].long()
                f32 = (scale < threshold) & ~diffuse_t[sp]
                npp = n_pp_sp[sp]
                kprim = torch.as_tensor(np.diff(ac.ks_prim_off.astype(np.int64)), device=dev)
                gprim = torch.zeros(G, dtype=torch.int64, device=dev).index_add_(
                    0, torch.arange(ac.n_ks, device=dev) // 64, kprim)
                trip = npp * gprim[items[:, 1].long() // 64]
                all_total += trip.sum()
                f64_total += trip[~f32].sum()

                def offsets(mask):
                    per_sp = mask.reshape(pc.n_sp, G).sum(1)
                    return np.concatenate([[0], np.cumsum(per_sp.cpu().numpy())]).astype(np.int64)

                self.mixed[(pi, ai)] = dict(perm=None, G=G, it64=items[~f32], pp64=None, off64=offsets(~f32),
                                            it32=items[f32], pp32=None, off32=offsets(f32), acc32=0,
                                            cost64=npp[~f32], cost32=npp[f32])
        f64_trip, all_trip = int(f64_total.item()), int(all_total.item())
        self.mixed_stats = {"f64_triplets": f64_trip, "f32_triplets": all_trip - f64_trip, "threshold": threshold,
                            "aux_order": "natural", "granularity": "item", "estimate": "envelope",
                            "estimate_seconds": _time.perf_counter() - t0}
        return self.mixed_stats

    def setup_checked(self, threshold=1e-6):
        """Checked mixed precision: every work item of a class whose float32 tier runs on the bf16
        matrix cores is first evaluated there; the kernels also accumulate, per output, the term scale
        sum_m sum_k |E_k| |V_k,m| (pair-matrix entries times the auxiliary side's Hermite sums, over every
        primitive pair of the item) and flag each item where some output's scale reaches ``threshold``;
        the float64 kernels then recompute exactly the flagged items (and every item of the other
        classes).  The split follows the computed terms, with their distance decay, instead of a
        distance-free bound, and a float32 item's error stays a few float32 roundings of the threshold
        even where the primitive pairs of a contracted shell pair cancel."""
        self.mixed = {(pi, ai): {"checked": True, "perm": None} for pi in range(len(self.pairs))
                      for ai in range(len(self.aux))}
        self.mixed_aux_order = "natural"
        self._mixed_generation = self.__dict__.get("_mixed_generation", 0) + 1
        trip = sum(self.primitive_triplets().values())
        self.mixed_stats = {"f64_triplets": None, "f32_triplets": trip, "threshold": threshold,
                            "aux_order": "natural", "granularity": "checked"}
        return self.mixed_stats

    @property
    def checked(self):
        return getattr(self, "mixed_stats", {}).get("granularity") == "checked"

    def run_mixed(self, out=None, stream=None, zero=True):
        """Mixed-precision (ij|P) into the packed float64 output (requires ``setup_mixed``)."""
        if out is None:
            out = self.new_output()
        if self.checked:
            keys = list(self.mixed)
            jobs = self.job_table([self.job_words(pi, ai, out.data_ptr(), self.n_aux) for pi, ai in keys])
            with _fork(stream, int(zero and self.zero_spans.shape[0] > 0)) as sides:
                if sides:
                    self.zero_unwritten(out, sides[0], ZERO_BLOCKS_BESIDE_COMPUTE)
                for j, (pi, ai) in enumerate(keys):
                    self.launch_checked(pi, ai, self.class_items(pi, ai), jobs[j], layout=0, stream=stream)
            return out
        keys = list(self.mixed)
        jobs = self.job_table([self.job_words(pi, ai, out.data_ptr(), self.n_aux, self.mixed[(pi, ai)]["perm"])
                               for pi, ai in keys])
        with _fork(stream, int(zero and self.zero_spans.shape[0] > 0)) as sides:
            if sides:
                self.zero_unwritten(out, sides[0], ZERO_BLOCKS_BESIDE_COMPUTE)
            for j, (pi, ai) in enumerate(keys):
                m = self.mixed[(pi, ai)]
                self.launch_work(pi, ai, m["it64"], jobs[j], pp=m["pp64"], layout=0, accumulate=0, stream=stream)
                self.launch_work(pi, ai, m["it32"], jobs[j], pp=m["pp32"], layout=0, accumulate=m["acc32"],
                                 use_f32=True, stream=stream)
        return out

    def mixed_items(self, pi, ai, sp0, sp1):
        """(items64, pp64, items32, pp32, accumulate32) of shell pairs [sp0, sp1) (views; pp may be None)."""
        m = self.mixed[(pi, ai)]
        a, b = int(m["off64"][sp0]), int(m["off64"][sp1])
        c, d = int(m["off32"][sp0]), int(m["off32"][sp1])
        pp64 = m["pp64"][a:b] if m["pp64"] is not None else None
        pp32 = m["pp32"][c:d] if m["pp32"] is not None else None
        return m["it64"][a:b], pp64, m["it32"][c:d], pp32, m["acc32"]

    # -- streaming by shell-pair blocks (outputs larger than device memory) ---------------------------
    def row_blocks(self, max_bytes):
        """Blocks (pi, sp0, sp1) of consecutive shell pairs of one class whose compact output
        ((sp1 - sp0) * nij rows of n_aux float64) fits in max_bytes."""
        blocks = []
        for pi, pc in enumerate(self.pairs):
            nij = (2 * pc.la + 1) * (2 * pc.lb + 1)
            per = max(1, int(max_bytes // (nij * self.n_aux * 8)))
            for sp0 in range(0, pc.n_sp, per):
                blocks.append((pi, sp0, min(pc.n_sp, sp0 + per)))
        return blocks

    def block_rows(self, pi, sp0, sp1):
        pc = self.pairs[pi]
        return (sp1 - sp0) * (2 * pc.la + 1) * (2 * pc.lb + 1)

    def run_block(self, pi, sp0, sp1, out, mixed=False, stream=None, aux_classes=None, concurrent=False,
                  lanes=None):
        """(ij|P) of shell pairs [sp0, sp1) of class pi, all auxiliary shells, into the compact block
        out[(sp - sp0) * nij + ij, P] (every row written; no zeroing needed).  ``aux_classes`` (indices
        into self.aux) restricts the launches to those auxiliary classes' columns.  ``concurrent`` spreads
        the launches over N_SIDE_STREAMS + 1 streams by estimated cost (each auxiliary class writes its own
        columns, and an item-granularity mixed split its own items in each tier); ``stream`` waits for all
        of them before the call returns.  ``lanes`` (a list of streams) spreads them over those streams
        instead, with no fork or join: the caller orders the lanes against other work (e.g. double-
        buffered blocks, where block k + 2 waits only for block k)."""
        pc = self.pairs[pi]
        nij = (2 * pc.la + 1) * (2 * pc.lb + 1)
        # rows of n_aux columns at the tensor's row stride (a padded buffer may have a longer one)
        if out.stride(1) != 1 or out.shape[1] < self.n_aux or out.stride(0) < self.n_aux:
            raise ValueError("run_block: out must hold rows of n_aux contiguous columns")
        ld = out.stride(0)
        base = out.data_ptr() - 8 * sp0 * nij * ld
        entries = [self.job_words(pi, ai, base, ld, self.mixed[(pi, ai)]["perm"] if mixed else None)
                   for ai in range(len(self.aux))]
        jobs = self.job_table(entries)
        groups = []          # (cost, [launch(stream) in order]): groups are independent of each other
        for ai, ac in enumerate(self.aux):
            if aux_classes is not None and ai not in aux_classes:
                continue
            fl = three_center_flops_per_triplet(pc.la, pc.lb, ac.lc) * max(1, ac.n_prim)
            if not mixed:
                it = self._block_items(pi, ai, sp0, sp1)[0]
                groups.append((fl * it.shape[0], [lambda s, ai=ai, it=it: self.launch_work(
                    pi, ai, it, jobs[ai], layout=1, stream=s)]))
                continue
            if self.checked:
                it = self._block_items(pi, ai, sp0, sp1)[0]
                groups.append((fl * it.shape[0], [lambda s, ai=ai, it=it: self.launch_checked(
                    pi, ai, it, jobs[ai], layout=1, stream=s)]))
                continue
            if self.mixed_aux_order == "natural":
                it64, pp64 = self._block_items(pi, ai, sp0, sp1, "f64")
                it32, pp32 = self._block_items(pi, ai, sp0, sp1, "f32")
                acc32 = self.mixed[(pi, ai)]["acc32"]
            else:
                # Q_P-ordered splits stay in shell-pair order: longest-first dispatch measured slower
                # there (output columns are permuted as well)
                it64, pp64, it32, pp32, acc32 = self.mixed_items(pi, ai, sp0, sp1)
            f64 = lambda s, ai=ai, it=it64, pp=pp64: self.launch_work(pi, ai, it, jobs[ai], pp=pp, layout=1,
                                                                       accumulate=0, stream=s)
            f32 = lambda s, ai=ai, it=it32, pp=pp32, acc=acc32: self.launch_work(
                pi, ai, it, jobs[ai], pp=pp, layout=1, accumulate=acc, use_f32=True, stream=s)
            n64 = it64.shape[0] if it64 is not None else 0
            n32 = it32.shape[0] if it32 is not None else 0
            if acc32:
                groups.append((fl * (n64 + n32), [f64, f32]))       # the float32 part adds onto the float64
            else:
                groups.append((fl * n64, [f64]))
                groups.append((0.6 * fl * n32, [f32]))
        if lanes is not None:
            # the job table (and any item table) may have just been written on the current stream
            cur = self.torch.cuda.current_stream(self.device)
            for ln in lanes:
                ln.wait_stream(cur)
            for k, idx in enumerate(_assign_lpt([c for c, _ in groups], len(lanes))):
                for i in idx:
                    for fn in groups[i][1]:
                        fn(lanes[k])
            return out
        if not concurrent or len(groups) < 2:
            for _, fns in groups:
                for fn in fns:
                    fn(stream)
            return out
        with _fork(stream, min(N_SIDE_STREAMS, len(groups) - 1)) as sides:
            lanes = [stream] + sides
            for k, idx in enumerate(_assign_lpt([c for c, _ in groups], len(lanes))):
                for i in idx:
                    for fn in groups[i][1]:
                        fn(lanes[k])
        return out

    def _block_items(self, pi, ai, sp0, sp1, kind="plain"):
        """(items, pp ranges or None) of shell pairs [sp0, sp1) for one launch -- "plain" (float64,
        every item) or "f64"/"f32" (the current mixed split) -- stored for longest-first dispatch; built
        on first use and kept (a table is never freed while a launch that reads it may be in flight)."""
        cache = self.__dict__.setdefault("_block_cache", {})
        gen = 0 if kind == "plain" else self.__dict__.get("_mixed_generation", 0)
        key = (pi, ai, sp0, sp1, kind, gen)
        hit = cache.get(key)
        if hit is not None:
            return hit
        torch = self.torch
        if kind == "plain":
            G = (self.aux[ai].n_ks + 63) // 64
            it, pp = self.items[(pi, ai)][sp0 * G:sp1 * G], None
            n_pp_sp = np.diff(np.asarray(self.pairs[pi].sp_pp_off, dtype=np.int64))[sp0:sp1]
            cost = torch.as_tensor(n_pp_sp, device=self.device).repeat_interleave(G)
        else:
            m, k = self.mixed[(pi, ai)], kind[1:]
            a, b = int(m["off" + k][sp0]), int(m["off" + k][sp1])
            it = m["it" + k][a:b]
            pp = m["pp" + k][a:b] if m["pp" + k] is not None else None
            cost = m["cost" + k][a:b]
        if cost.shape[0] > 1:
            order = longest_first_t(cost)
            it = it.index_select(0, order)
            pp = pp.index_select(0, order) if pp is not None else None
            # the launch may go to another stream than the one that reordered the tables
            if self.device.type == "cuda":
                torch.cuda.current_stream(self.device).synchronize()
        cache[key] = (it, pp)
        return it, pp

    def prepare_blocks(self, blocks, mixed=False):
        """Build the dispatch tables ``run_block`` uses for every (pi, sp0, sp1) in ``blocks`` ahead of a
        timed or captured sequence (a first use builds them on the host and synchronises)."""
        kinds = (("f64", "f32") if mixed and getattr(self, "mixed_aux_order", None) == "natural" and not self.checked
                 else ())
        for pi, sp0, sp1 in blocks:
            for ai in range(len(self.aux)):
                for kind in ("plain",) + kinds:
                    self._block_items(pi, ai, sp0, sp1, kind)


def metric_cholesky(J2c):
    """Lower Cholesky factor L (L L^T = (P|Q)) of the float64 metric on the device (the solver library
    through torch.linalg), the factorisation a density-fitting consumer applies to (P|ij).  Raises when
    the metric is not numerically positive definite."""
    import torch
    L, info = torch.linalg.cholesky_ex(J2c, upper=False)
    if int(info.item()) != 0:
        raise ValueError(f"metric is not positive definite (leading minor {int(info.item())})")
    return L


INVERSE_BLOCK = 512      # right-hand sides per triangular solve when CoulombFit forms L^-1


class CoulombFit:
    """Density-fitted Coulomb matrix from (ij|P) consumed on the device as it is produced:

        gamma_P = sum_ij (P|ij) D_ij,   (P|Q) d = gamma,   J_ij = sum_P (ij|P) d_P

    (the standard RI-J contraction; the reference computes only the integrals).  The metric (P|Q) is
    factorised once in float64 (metric_cholesky).  ``from_packed`` consumes a packed (ij|P) output (two
    matrix-vector products over it); ``streamed`` consumes compact blocks of shell pairs streamed through
    one device buffer, evaluating the integrals twice (gamma in the first pass, J in the second) so that
    nothing larger than one block is ever materialised -- outputs of terabytes need no more than the
    block.  The two contractions are the library's streaming kernels (eip_weighted_column_sums,
    eip_row_dots: one pass over the output in HBM, fixed summation order)."""

    def __init__(self, plan, J2c=None, L=None, inverse=True):
        """inverse: also form L^-1 once (a triangular solve against the identity), so each metric solve is
        two matrix-vector products over it instead of two latency-bound triangular vector solves."""
        self.plan = plan
        torch = plan.torch
        if L is None:
            if J2c is None:
                J2c = plan.two_center()
            L = metric_cholesky(J2c)
        self.L = L
        self.Linv = None
        if inverse:
            # L^-1 column block by column block (the solver library's workspace stays small): block
            # [c0, c1) solves L X = I[:, c0:c1], whose rows above c0 are zero
            n = L.shape[0]
            self.Linv = torch.zeros((n, n), dtype=L.dtype, device=L.device)
            for c0 in range(0, n, INVERSE_BLOCK):
                c1 = min(n, c0 + INVERSE_BLOCK)
                rhs = torch.zeros((n - c0, c1 - c0), dtype=L.dtype, device=L.device)
                rhs[:c1 - c0].fill_diagonal_(1.0)
                self.Linv[c0:, c0:c1] = torch.linalg.solve_triangular(L[c0:, c0:], rhs, upper=False)
        n = plan.n_ao
        i, j = np.tril_indices(n)
        self.tri_i = torch.as_tensor(i, device=plan.device)
        self.tri_j = torch.as_tensor(j, device=plan.device)

    def solve(self, gamma):
        """d = (P|Q)^-1 gamma through the float64 Cholesky factor (d = L^-T (L^-1 gamma))."""
        torch = self.plan.torch
        if self.Linv is not None:
            return torch.mv(self.Linv.t(), torch.mv(self.Linv, gamma))
        return torch.cholesky_solve(gamma.view(-1, 1), self.L).view(-1)

    def _weights(self, D):
        torch = self.plan.torch
        return D[self.tri_i, self.tri_j] * torch.where(self.tri_i == self.tri_j, 1.0, 2.0).to(D.dtype)

    def _col_sums(self, A, w, gamma, accumulate, stream=None):
        """gamma (+)= A^T w by the library's streaming column-sum kernels (A row-major, float64)."""
        torch = self.plan.torch
        lib = self.plan.lib
        n_rows, n_cols = A.shape
        need = int(lib.eip_weighted_column_sums_workspace(_c_ll(n_rows), n_cols))
        work = self.__dict__.get("_work")
        if work is None or work.numel() < need:
            work = self._work = torch.empty(max(need, 1), dtype=torch.float64, device=A.device)
        rc = lib.eip_weighted_column_sums(_ptr(A), _c_ll(n_rows), n_cols, _c_ll(A.stride(0)), _ptr(w), _ptr(work),
                                          _c_ll(work.numel()), int(accumulate), _ptr(gamma), _stream_handle(stream))
        if rc != 0:
            raise RuntimeError(f"eip_weighted_column_sums failed with {rc}")
        return gamma

    def _row_dots(self, A, x, out, stream=None):
        n_rows, n_cols = A.shape
        rc = self.plan.lib.eip_row_dots(_ptr(A), _c_ll(n_rows), n_cols, _c_ll(A.stride(0)), _ptr(x), _ptr(out),
                                        _stream_handle(stream))
        if rc != 0:
            raise RuntimeError(f"eip_row_dots failed with {rc}")
        return out

    def gamma_packed(self, packed, D):
        """gamma_P = sum_ij (P|ij) D_ij from the packed output J3c[tri(i, j), P] (i >= j)."""
        torch = self.plan.torch
        gamma = torch.empty(packed.shape[1], dtype=torch.float64, device=packed.device)
        return self._col_sums(packed, self._weights(D).contiguous(), gamma, False)

    def contract_packed(self, packed, d):
        """J_ij = sum_P (ij|P) d_P from the packed output (symmetric (n_ao, n_ao))."""
        torch = self.plan.torch
        jp = self._row_dots(packed, d.contiguous(), torch.empty(packed.shape[0], dtype=torch.float64,
                                                                 device=packed.device))
        n = self.plan.n_ao
        J = torch.zeros((n, n), dtype=jp.dtype, device=jp.device)
        J[self.tri_i, self.tri_j] = jp
        J[self.tri_j, self.tri_i] = jp
        return J

    def from_packed(self, packed, D):
        """J (n_ao, n_ao) from the packed output and a symmetric density D."""
        return self.contract_packed(packed, self.solve(self.gamma_packed(packed, D)))

    def _block_index(self, pi, sp0, sp1):
        """AO indices (I, J) of the compact rows of shell pairs [sp0, sp1) and each row's weight in gamma
        (1 inside a diagonal shell pair, whose block holds both (i, j) and (j, i); 2 otherwise)."""
        cache = self.__dict__.setdefault("_idx", {})
        key = (pi, sp0, sp1)
        if key not in cache:
            torch = self.plan.torch
            pc = self.plan.pairs[pi]
            ni, nj = 2 * pc.la + 1, 2 * pc.lb + 1
            sp = np.arange(sp0, sp1)
            si, sj = pc.sp_i[sp].astype(np.int64), pc.sp_j[sp].astype(np.int64)
            I = np.repeat(si[:, None] + np.arange(ni)[None, :], nj, axis=1).ravel()
            J = np.tile(sj[:, None] + np.arange(nj)[None, :], (1, ni)).ravel()
            w = np.repeat(np.where(si == sj, 1.0, 2.0), ni * nj)
            cache[key] = tuple(torch.as_tensor(a, device=self.plan.device) for a in (I, J, w))
        return cache[key]

    def _blocks(self, block_bytes):
        plan = self.plan
        torch = plan.torch
        blocks = plan.row_blocks(block_bytes)
        rows = max(plan.block_rows(*b) for b in blocks)
        buf = self.__dict__.get("_buf")
        if buf is None or buf.shape[0] < rows:
            buf = self._buf = torch.empty((rows, plan.n_aux), dtype=torch.float64, device=plan.device)
        return blocks, buf

    def gamma_streamed(self, D, block_bytes, mixed=False, stream=None):
        """gamma from one streamed pass of compact blocks (one reused buffer of at most block_bytes)."""
        plan = self.plan
        blocks, buf = self._blocks(block_bytes)
        gamma = plan.torch.zeros(plan.n_aux, dtype=plan.torch.float64, device=plan.device)
        for pi, sp0, sp1 in blocks:
            n = plan.block_rows(pi, sp0, sp1)
            plan.run_block(pi, sp0, sp1, buf, mixed=mixed, stream=stream)
            I, J, w = self._block_index(pi, sp0, sp1)
            self._col_sums(buf[:n], (D[I, J] * w).contiguous(), gamma, True, stream)
        return gamma

    def contract_streamed(self, d, block_bytes, mixed=False, stream=None):
        """J from one streamed pass of compact blocks."""
        plan = self.plan
        torch = plan.torch
        blocks, buf = self._blocks(block_bytes)
        out = torch.zeros((plan.n_ao, plan.n_ao), dtype=torch.float64, device=plan.device)
        for pi, sp0, sp1 in blocks:
            n = plan.block_rows(pi, sp0, sp1)
            plan.run_block(pi, sp0, sp1, buf, mixed=mixed, stream=stream)
            I, J, _ = self._block_index(pi, sp0, sp1)
            jr = self._row_dots(buf[:n], d.contiguous(), torch.empty(n, dtype=torch.float64, device=d.device), stream)
            out[I, J] = jr
            out[J, I] = jr
        return out

    def streamed(self, D, block_bytes, mixed=False, stream=None):
        """J from two streamed passes (gamma, then J): nothing larger than one block is materialised."""
        d = self.solve(self.gamma_streamed(D, block_bytes, mixed, stream))
        return self.contract_streamed(d, block_bytes, mixed, stream)


def packed_to_dense(packed, n_ao):
    """(n_ao(n_ao+1)/2, n_aux) packed -> (n_aux, n_ao, n_ao) dense (reference layout), on the device."""
    import torch
    idx_i, idx_j = torch.tril_indices(n_ao, n_ao, device=packed.device)
    n_aux = packed.shape[1]
    dense = torch.zeros((n_aux, n_ao, n_ao), dtype=packed.dtype, device=packed.device)
    dense[:, idx_i, idx_j] = packed.t()
    dense[:, idx_j, idx_i] = packed.t()
    return dense


# ------------------------------------------------------------------------------------------------
# Persistent queue kernels: work items of many classes (and molecules or jobs) in one device queue per
# register tier, run by one persistent kernel per tier.
# ------------------------------------------------------------------------------------------------

QUEUE_PARTS = 8          # one queue part per XCD
QUEUE_RUN = 16           # consecutive entries dealt to one part (items of one shell pair share its L2)
# Cheap items are grouped into one queue entry until the entry holds about this much work (per-lane
# flops: primitive pairs x flops per triplet of the chunk, ~25 us of one wave's time) but at most
# QUEUE_MAX_GROUP items: one atomic fetch per entry, not per item.
QUEUE_ENTRY_COST = 3.0e4
QUEUE_MAX_GROUP = 64
# Tuned queue parameters per precision ({"f64": {"waves_per_simd": w, "entry_cost": c}}), loaded from
# tuning/queue.json when it exists (autotune_queue writes it).
QUEUE_TUNING_FILE = __import__("pathlib").Path(__file__).resolve().parent / "tuning" / "queue.json"
QUEUE_TUNING = __import__("json").loads(QUEUE_TUNING_FILE.read_text()) if QUEUE_TUNING_FILE.exists() else {}
# Control block words (one 128-byte line per part counter; see eip_three_center_queue_run)
_CTL_FIRST = QUEUE_PARTS * 32
_CTL_COUNT = _CTL_FIRST + QUEUE_PARTS


class WorkQueue:
    """Three-centre work items of many classes for the persistent tier kernels.

    ``add`` registers one class's item list (through a job table, so items of several molecules or jobs
    can share it) as one slot per component chunk of the class's queue plan; ``build`` orders every tier's
    entries -- classes by decreasing mean item cost, items longest-first within a class ("sorted by class
    and primitive count") -- and deals them in runs of QUEUE_RUN entries over the 8 XCD parts; ``upload``
    writes slots, entries and control blocks to the device (into existing buffers when they are given and
    large enough, so a captured graph can be replayed over new work); ``run`` launches one kernel per tier,
    the tiers concurrently on side streams."""

    def __init__(self, lib=None, device="cuda", use_f32=False, waves_per_simd=None, entry_cost=None):
        """waves_per_simd: cap of the tier kernels' grid occupancy (0: as many waves as fit); entry_cost:
        work per queue entry cheap items are grouped up to.  None: the tuned values (QUEUE_TUNING, written by
        autotune_queue) or the defaults."""
        import torch
        self.torch = torch
        self.lib = lib or library()
        self.device = torch.device(device)
        self.f32 = int(bool(use_f32))
        tuned = QUEUE_TUNING.get("f32" if use_f32 else "f64", {})
        self.waves_per_simd = int(tuned.get("waves_per_simd", 0) if waves_per_simd is None else waves_per_simd)
        self.entry_cost = float(tuned.get("entry_cost", QUEUE_ENTRY_COST) if entry_cost is None else entry_cost)
        self.slot_bytes = int(self.lib.eip_three_center_queue_slot_bytes())
        self.ctl_words = int(self.lib.eip_three_center_queue_ctl_words())
        self.ntiers = int(self.lib.eip_three_center_queue_tiers(self.f32))
        self._slots = []
        self._tier = {t: [] for t in range(self.ntiers)}
        self.n_triplet_items = 0
        self.host = None
        self.dev = None

    def supports(self, la, lb, lc):
        """Whether the class runs in the tier kernels: it has bodies there (classes whose bodies need more
        registers than a persistent loop leaves -- the highest (df|g) and (ff|x) chunks -- do not), and
        they are its current variant's kernels for this precision (a class tuned to the matrix cores, or
        to a vector plan the queue could not hold, keeps its own, faster kernels)."""
        key = (la, lb, lc, self.f32, self.lib.eip_three_center_variant(la, lb, lc, self.f32))
        hit = _QUEUE_SUPPORT.get(key)
        if hit is None:
            m0, m1, tier = _c_int(), _c_int(), _c_int()
            queue_plan, c = [], 0
            while self.lib.eip_three_center_queue_chunk(la, lb, lc, self.f32, c, ctypes.byref(m0), ctypes.byref(m1),
                                                        ctypes.byref(tier)) == 0:
                queue_plan.append((m0.value, m1.value))
                c += 1
            own_plan, c = [], 0
            while self.lib.eip_three_center_chunk(la, lb, lc, key[4], c, ctypes.byref(m0), ctypes.byref(m1)) == 0:
                own_plan.append((m0.value, m1.value))
                c += 1
            # in the queue only when its bodies are the class's current (tuned) vector kernels
            hit = bool(queue_plan) and queue_plan == own_plan and \
                self.lib.eip_three_center_variant_mfma(la, lb, lc, key[4]) == 0
            _QUEUE_SUPPORT[key] = hit
        return hit

    def add(self, la, lb, lc, items, n_items, jobs_ptr, costs, ks_end=None, pp=None, job=None, layout=0,
            accumulate=0, omega=0.0):
        """items / ks_end / pp / job: device int32 tensors of eip_three_center_work's item list (kept alive by
        the caller); jobs_ptr: device address of the job table; costs: host array of each item's relative
        cost (primitive pairs x auxiliary primitives), or one number for every item (the items then run in
        their stored order, and the queue is built on the device: ``build_device``)."""
        if n_items <= 0:
            return
        m0, m1, tier = _c_int(), _c_int(), _c_int()
        per_comp = three_center_flops_per_triplet(la, lb, lc) / (2 * lc + 1)
        costs = float(costs) if np.isscalar(costs) else np.asarray(costs, dtype=np.float64)
        c = 0
        while self.lib.eip_three_center_queue_chunk(la, lb, lc, self.f32, c, ctypes.byref(m0), ctypes.byref(m1),
                                                    ctypes.byref(tier)) == 0:
            buf = ctypes.create_string_buffer(self.slot_bytes)
            t = self.lib.eip_three_center_queue_slot(la, lb, lc, self.f32, c, _ptr(items), _ptr(ks_end), _ptr(pp),
                                                     _ptr(job), int(n_items), _vp(int(jobs_ptr)), int(layout),
                                                     int(accumulate), _c_double(omega), buf)
            if t < 0:
                raise RuntimeError(f"eip_three_center_queue_slot({la},{lb},{lc}, chunk {c}) failed with {t}")
            self._tier[t].append((len(self._slots), costs * per_comp * (m1.value - m0.value), int(n_items)))
            self._slots.append(buf.raw)
            c += 1
        if c == 0:
            raise RuntimeError(f"no queue kernels for class ({la}{lb}|{lc})")
        self.n_triplet_items += int(n_items)

    def build(self):
        """Host arrays: slots (bytes), entries [n][4] int32 (slot, first item, item count, 0) and the
        control block of every tier.  A slot's items are grouped in their stored order into runs of about
        entry_cost work; runs are ordered longest-first within a slot, slots by decreasing mean item
        cost, and the ordered runs are dealt QUEUE_RUN at a time over the 8 parts."""
        entries, ctls, tot_cost = [], [], []
        for t in range(self.ntiers):
            lst = sorted(self._tier[t], key=lambda e: -float(np.mean(e[1])))
            rows = []
            for s, cst, _ in lst:
                n = len(cst)
                g = int(np.clip(np.ceil(self.entry_cost / max(float(cst.mean()), 1e-30)), 1, QUEUE_MAX_GROUP))
                first = np.arange(0, n, g, dtype=np.int64)
                count = np.minimum(g, n - first)
                run_cost = np.add.reduceat(cst, first)
                o = np.argsort(-run_cost, kind="stable")
                rows.append(np.stack([np.full(len(first), s), first[o], count[o], np.zeros(len(first), np.int64)], 1))
            ent = np.concatenate(rows).astype(np.int32) if rows else np.zeros((0, 4), np.int32)
            n = len(ent)
            part = (np.arange(n) // QUEUE_RUN) % QUEUE_PARTS
            order = np.argsort(part, kind="stable")
            counts = np.bincount(part, minlength=QUEUE_PARTS).astype(np.int32)
            ctl = np.zeros(self.ctl_words, np.int32)
            ctl[_CTL_FIRST:_CTL_FIRST + QUEUE_PARTS] = np.concatenate([[0], np.cumsum(counts)[:-1]])
            ctl[_CTL_COUNT:_CTL_COUNT + QUEUE_PARTS] = counts
            entries.append(np.ascontiguousarray(ent[order]))
            ctls.append(ctl)
            tot_cost.append(sum(float(np.sum(c)) for _, c, _ in lst))
        # heaviest tier first: launched first, it takes the device while the lighter ones fill its tail
        self.tier_order = [int(t) for t in np.argsort(-np.array(tot_cost), kind="stable")]
        self.host = {"slots": b"".join(self._slots), "entries": entries, "ctl": ctls}
        return self

    def build_device(self, stream=None):
        """Build and upload the queue with the entries formed on the device (for item lists that live on
        the device and may hold 10^8 items: only per-slot numbers are handled on the host).  Items run in
        their stored order, grouped into runs by the slot's mean cost; slots by decreasing mean cost; runs
        dealt QUEUE_RUN at a time over the 8 parts.  Returns the device buffers."""
        torch = self.torch
        dev = self.device
        s = stream if stream is not None else torch.cuda.current_stream()
        ents, ctls, tot_cost = [], [], []
        self._staging = []

        def h2d(a, dtype=None):
            # pinned staging and an asynchronous copy on s: the host does not wait for work already
            # queued on s (a serving step in flight); the staging buffers live as long as the queue
            h = torch.from_numpy(np.ascontiguousarray(a)).pin_memory()
            self._staging.append(h)
            return h.to(dev, non_blocking=True)

        with torch.cuda.stream(s):
            for t in range(self.ntiers):
                lst = sorted(self._tier[t], key=lambda e: -float(np.mean(e[1])))
                slot_ids, n_it, grp, n_run = [], [], [], []
                for sl, cst, n in lst:
                    g = int(np.clip(np.ceil(self.entry_cost / max(float(np.mean(cst)), 1e-30)), 1, QUEUE_MAX_GROUP))
                    slot_ids.append(sl); n_it.append(n); grp.append(g); n_run.append((n + g - 1) // g)
                N = int(sum(n_run))
                if N:
                    per_slot = h2d(np.array([n_run, slot_ids, grp, n_it], dtype=np.int64))
                    nr = per_slot[0]
                    sid = torch.repeat_interleave(per_slot[1], nr, output_size=N)
                    start = torch.repeat_interleave(torch.cumsum(nr, 0) - nr, nr, output_size=N)
                    g = torch.repeat_interleave(per_slot[2], nr, output_size=N)
                    n = torch.repeat_interleave(per_slot[3], nr, output_size=N)
                    j = torch.arange(N, dtype=torch.int64, device=dev)
                    first = (j - start) * g
                    count = torch.minimum(g, n - first)
                    ent = torch.stack([sid, first, count, torch.zeros_like(first)], 1).to(torch.int32)
                    # deal: ordered entry j goes to part p = (j // RUN) % 8, after the runs dealt to p before
                    p = (j // QUEUE_RUN) % QUEUE_PARTS
                    counts = np.bincount((np.arange(N) // QUEUE_RUN) % QUEUE_PARTS, minlength=QUEUE_PARTS)
                    offs = np.concatenate([[0], np.cumsum(counts)[:-1]])
                    dest = (h2d(offs.astype(np.int64))[p] + (j // (QUEUE_RUN * QUEUE_PARTS)) * QUEUE_RUN
                            + j % QUEUE_RUN)
                    out = torch.empty_like(ent)
                    out[dest] = ent
                else:
                    counts = np.zeros(QUEUE_PARTS, np.int64)
                    offs = counts
                    out = torch.zeros((1, 4), dtype=torch.int32, device=dev)
                ctl = np.zeros(self.ctl_words, np.int32)
                ctl[_CTL_FIRST:_CTL_FIRST + QUEUE_PARTS] = offs
                ctl[_CTL_COUNT:_CTL_COUNT + QUEUE_PARTS] = counts
                ents.append(out.reshape(-1))
                ctls.append(ctl)
                tot_cost.append(sum(float(np.mean(c)) * nn for _, c, nn in lst))
            self.tier_order = [int(t) for t in np.argsort(-np.array(tot_cost), kind="stable")]
            slots = np.frombuffer(bytearray(b"".join(self._slots) or bytes(8)), dtype=np.uint8)
            self.dev = {"slots": h2d(slots), "entries": ents, "ctl": h2d(np.concatenate(ctls))}
        return self.dev

    def upload(self, into=None, stream=None):
        """Copy the built queue to the device (into ``into``, a previous ``self.dev``-like dict of device
        buffers, when every buffer there is large enough); returns the buffers."""
        torch = self.torch
        if self.host is None:
            self.build()
        h = self.host
        slots = torch.frombuffer(bytearray(h["slots"]) or bytearray(8), dtype=torch.uint8)
        ent = [torch.from_numpy(e.reshape(-1)) for e in h["entries"]]
        ctl = torch.from_numpy(np.concatenate(h["ctl"]))
        fits = (into is not None and into["slots"].numel() >= slots.numel()
                and all(d.numel() >= max(1, e.numel()) for d, e in zip(into["entries"], ent)))
        if not fits:
            into = {"slots": torch.empty(slots.numel(), dtype=torch.uint8, device=self.device),
                    "entries": [torch.empty(max(1, e.numel()), dtype=torch.int32, device=self.device) for e in ent],
                    "ctl": torch.empty(ctl.numel(), dtype=torch.int32, device=self.device)}
        s = stream if stream is not None else torch.cuda.current_stream()
        with torch.cuda.stream(s):
            into["slots"][:slots.numel()].copy_(slots.to(self.device, non_blocking=False))
            for d, e in zip(into["entries"], ent):
                if e.numel():
                    d[:e.numel()].copy_(e.to(self.device))
            into["ctl"].copy_(ctl.to(self.device))
        self.dev = into
        return into

    @staticmethod
    def buffers_for(queues):
        """Device buffers large enough for every one of the (built) queues: one captured launch sequence
        can then be replayed over each of them in turn after ``upload(into=...)``."""
        q0 = queues[0]
        torch = q0.torch
        nslot = max(len(q.host["slots"]) for q in queues)
        nent = [max(max(1, q.host["entries"][t].size) for q in queues) for t in range(q0.ntiers)]
        return {"slots": torch.empty(max(8, nslot), dtype=torch.uint8, device=q0.device),
                "entries": [torch.empty(n, dtype=torch.int32, device=q0.device) for n in nent],
                "ctl": torch.empty(q0.ctl_words * q0.ntiers, dtype=torch.int32, device=q0.device)}

    def grid(self, t):
        key = (self.f32, t, self.waves_per_simd)
        g = _QUEUE_GRIDS.get(key)
        if g is None:
            g = int(self.lib.eip_three_center_queue_grid(self.f32, t, self.waves_per_simd))
            if g <= 0:
                raise RuntimeError(f"eip_three_center_queue_grid({self.f32}, {t}) failed with {g}")
            _QUEUE_GRIDS[key] = g
        return g

    def fits(self, bufs):
        h = self.host
        return (bufs is not None and bufs["slots"].numel() >= len(h["slots"])
                and all(d.numel() >= e.size for d, e in zip(bufs["entries"], h["entries"])))

    def run(self, stream=None, concurrent=True):
        """One launch per tier (all tiers, empty ones exit at once: the launch sequence does not depend on
        the work, so it can be captured once and replayed after new uploads)."""
        if self.dev is None:
            self.upload(stream=stream)
        d = self.dev
        order = getattr(self, "tier_order", list(range(self.ntiers)))
        with _fork(stream, len(order) - 1 if concurrent else 0) as sides:
            streams = [stream] + list(sides)
            for k, t in enumerate(order):
                st = streams[k] if concurrent else stream
                rc = self.lib.eip_three_center_queue_run(
                    self.f32, t, _ptr(d["slots"]), _ptr(d["entries"][t]),
                    _vp(d["ctl"].data_ptr() + 4 * self.ctl_words * t), self.grid(t), _stream_handle(st))
                if rc != 0:
                    raise RuntimeError(f"eip_three_center_queue_run(tier {t}) failed with {rc}")


_QUEUE_GRIDS = {}
_QUEUE_SUPPORT = {}


class QueueGraph:
    """One captured HIP graph of the persistent tier kernels, replayed over successive work queues: each
    queue is uploaded into the same device buffers (slots, entries, control blocks), which the launches
    read, so the graph is captured once and replayed for every batch.  The buffers grow with headroom --
    and the graph is captured again -- only when a queue does not fit (``captures`` counts this)."""

    def __init__(self, headroom=1.5):
        self.headroom = headroom
        self.bufs = None
        self.graph = None
        self.order = None
        self.captures = 0

    def _grow(self, q):
        torch = q.torch
        h = q.host
        hr = self.headroom
        old = self.bufs
        self.bufs = {"slots": torch.empty(int(max(8, len(h["slots"])) * hr) + 8, dtype=torch.uint8, device=q.device),
                     "entries": [torch.empty(int(max(1, e.size) * hr) + 2, dtype=torch.int32, device=q.device)
                                 for e in h["entries"]],
                     "ctl": torch.zeros(q.ctl_words * q.ntiers, dtype=torch.int32, device=q.device)}
        if old is not None:
            for k in ("slots", "entries"):
                v = self.bufs[k]
                if isinstance(v, list):
                    for i, (a, b) in enumerate(zip(v, old[k])):
                        if a.numel() < b.numel():
                            v[i] = torch.empty(b.numel(), dtype=b.dtype, device=b.device)
                elif v.numel() < old[k].numel():
                    self.bufs[k] = torch.empty(old[k].numel(), dtype=old[k].dtype, device=old[k].device)
        self.graph = None

    def run(self, q, stream=None):
        """Upload queue q into the graph's buffers and replay the graph on ``stream`` (captured first if
        needed)."""
        torch = q.torch
        s = stream if stream is not None else torch.cuda.current_stream()
        self.ensure(q, s)
        q.tier_order = self.order
        q.upload(into=self.bufs, stream=s)
        with torch.cuda.stream(s):
            self.graph.replay()

    def ensure(self, q, stream=None):
        """Buffers that hold q and a captured graph over them (the capture is preceded by one launch over
        an empty queue, which does no work).  Call it before forking other streams from ``stream``."""
        torch = q.torch
        s = stream if stream is not None else torch.cuda.current_stream()
        if not q.fits(self.bufs):
            self._grow(q)
        if self.graph is None:
            for t in range(q.ntiers):
                q.grid(t)
            self.order = list(getattr(q, "tier_order", range(q.ntiers)))
            empty = WorkQueue(q.lib, q.device, use_f32=bool(q.f32), waves_per_simd=q.waves_per_simd,
                              entry_cost=q.entry_cost).build()
            empty.tier_order = self.order
            empty.upload(into=self.bufs, stream=s)
            empty.run(s)                                   # warm-up outside capture: no work
            s.synchronize()
            self.graph = torch.cuda.CUDAGraph()
            with torch.cuda.stream(s):
                with torch.cuda.graph(self.graph, stream=s):
                    empty.run(s)
            self.captures += 1


class BatchPlan:
    """(ij|P) of many molecules with one launch per class (and auxiliary-component chunk) per batch.

    Shell pairs and auxiliary shells of all molecules are concatenated per angular class; every work
    item pairs a shell pair with up to 64 auxiliary shells of the *same* molecule, and each shell pair
    carries the offset and row stride of its molecule's packed output inside one device buffer.  The
    launch sequence is fixed for a batch, so it can be captured once into a HIP graph and replayed.
    """

    def __init__(self, molecules, omega=0.0, device="cuda"):
        """molecules: list of (molecule, pair_data, aux_data)."""
        import torch
        self.torch = torch
        self.lib = library()
        _ensure_tuning()
        self.omega = float(omega)
        self.device = torch.device(device)
        T = lambda a, dt: torch.as_tensor(np.ascontiguousarray(a), dtype=dt, device=self.device)
        self.sizes = [(m.n_ao, m.n_aux) for m, _, _ in molecules]
        npk = [n_ao * (n_ao + 1) // 2 * n_aux for n_ao, n_aux in self.sizes]
        self.out_base = np.concatenate([[0], np.cumsum(npk)]).astype(np.int64)
        self.n_out = int(self.out_base[-1])
        pair_keys = sorted({k for _, pd, _ in molecules for k in pd})
        aux_keys = sorted({k for _, _, ad in molecules for k in ad})
        stream = _stream_handle(None)
        self.pairs, self.aux = [], []
        pair_n_pp = []  # primitive pairs per shell pair, per pair class
        mol_sp = {}   # (pair key, molecule) -> (sp_lo, sp_hi)
        for key in pair_keys:
            la, lb = key
            nnz = self.lib.eip_pair_matrix_nnz(la, lb)
            if nnz < 0:
                raise ValueError(f"no specialised kernels for pair class {key}")
            P, A, B, c, si, sj, off, base, ld = [], [], [], [], [], [], [0], [], []
            n_sp = 0
            for m, (mol, pd, _) in enumerate(molecules):
                if key not in pd:
                    continue
                order, msi, msj, moff = group_pairs(pd[key])
                P.append(pd[key].P_p[order]); A.append(pd[key].A_a[order]); B.append(pd[key].B_b[order])
                c.append(pd[key].coefficient[order]); si.append(msi); sj.append(msj)
                off.extend((moff[1:] + off[-1]).tolist())
                base.append(np.full(len(msi), self.out_base[m], np.int64)); ld.append(np.full(len(msi), mol.n_aux, np.int32))
                mol_sp[(key, m)] = (n_sp, n_sp + len(msi))
                n_sp += len(msi)
            n_pp = off[-1]
            d = {"P": T(np.concatenate(P), torch.float64), "A": T(np.concatenate(A), torch.float64),
                 "B": T(np.concatenate(B), torch.float64), "c": T(np.concatenate(c), torch.float64),
                 "off": T(np.array(off, np.int32), torch.int32), "si": T(np.concatenate(si), torch.int32),
                 "sj": T(np.concatenate(sj), torch.int32), "base": T(np.concatenate(base), torch.int64),
                 "ld": T(np.concatenate(ld), torch.int32)}
            d["E"] = torch.empty((n_pp, nnz), dtype=torch.float64, device=self.device)
            d["E32"] = torch.empty((n_pp, nnz), dtype=torch.float32, device=self.device)
            rc = self.lib.eip_pair_matrix(la, lb, _ptr(d["P"]), _ptr(d["A"]), _ptr(d["B"]), _ptr(d["c"]), n_pp,
                                          _ptr(d["E"]), _ptr(d["E32"]), stream)
            if rc != 0:
                raise RuntimeError(f"eip_pair_matrix{key} failed with {rc}")
            self.pairs.append((key, n_sp, n_pp, d))
            pair_n_pp.append(np.diff(np.array(off, np.int64)))
        mol_ks = {}
        for lc in aux_keys:
            C, c, off, col = [], [], [0], []
            n_ks = 0
            for m, (mol, _, ad) in enumerate(molecules):
                if lc not in ad:
                    continue
                order, ks, moff = group_aux(ad[lc])
                C.append(ad[lc].A_a[order]); c.append(ad[lc].coefficient[order]); col.append(ks)
                off.extend((moff[1:] + off[-1]).tolist())
                mol_ks[(lc, m)] = (n_ks, n_ks + len(ks))
                n_ks += len(ks)
            d = {"C": T(np.concatenate(C), torch.float64), "c": T(np.concatenate(c), torch.float64),
                 "off": T(np.array(off, np.int32), torch.int32), "col": T(np.concatenate(col), torch.int32)}
            self.aux.append((lc, n_ks, d))
        # items per (pair class, aux class), molecule by molecule
        self.items = {}
        self.item_costs = {}
        self.n_triplet_items = 0
        for pi, (key, _, _, _) in enumerate(self.pairs):
            for ai, (lc, _, _) in enumerate(self.aux):
                it_sp, it_k0, it_end = [], [], []
                for m in range(len(molecules)):
                    if (key, m) not in mol_sp or (lc, m) not in mol_ks:
                        continue
                    s0, s1 = mol_sp[(key, m)]
                    k0, k1 = mol_ks[(lc, m)]
                    nb = (k1 - k0 + 63) // 64
                    it_sp.append(np.repeat(np.arange(s0, s1, dtype=np.int32), nb))
                    it_k0.append(np.tile(k0 + 64 * np.arange(nb, dtype=np.int32), s1 - s0))
                    it_end.append(np.full((s1 - s0) * nb, k1, np.int32))
                if it_sp:
                    sp = np.concatenate(it_sp)
                    items, end = np.stack([sp, np.concatenate(it_k0)], axis=1), np.concatenate(it_end)
                    # longest-first dispatch (cost: primitive pairs of the item's shell pair)
                    order = longest_first(pair_n_pp[pi][sp])
                    self.items[(pi, ai)] = (T(items[order], torch.int32), T(end[order], torch.int32))
                    self.item_costs[(pi, ai)] = pair_n_pp[pi][sp][order].astype(np.float64)
                    self.n_triplet_items += len(sp)
        self.mol_ks = mol_ks
        # per molecule: packed rows that no kernel writes (shell pairs screened out entirely)
        self.zero_rows = {}
        for m, (mol, pd, _) in enumerate(molecules):
            n_ao = mol.n_ao
            written = np.zeros(n_ao * (n_ao + 1) // 2, dtype=bool)
            for (la, lb), p in pd.items():
                _, si, sj, _ = group_pairs(p)
                i = (si[:, None, None] + np.arange(2 * la + 1)[None, :, None]).astype(np.int64)
                j = (sj[:, None, None] + np.arange(2 * lb + 1)[None, None, :]).astype(np.int64)
                hi, lo = np.maximum(i, j), np.minimum(i, j)
                written[(hi * (hi + 1) // 2 + lo).ravel()] = True
            rows = np.flatnonzero(~written)
            if len(rows):
                sp = zero_spans(rows, mol.n_aux)
                sp[:, 0] += int(self.out_base[m])
                self.zero_rows[m] = sp
        self.zero_spans = T(np.concatenate(list(self.zero_rows.values())) if self.zero_rows
                            else np.zeros((0, 2), np.int64), torch.int64)
        # (P|Q) of every molecule: dense (n_aux, n_aux) blocks, one after another
        n2 = [n_aux * n_aux for _, n_aux in self.sizes]
        self.out2c_base = np.concatenate([[0], np.cumsum(n2)]).astype(np.int64)
        torch.cuda.current_stream().synchronize()

    def kernel_E(self, pi, ai):
        """Float64 pair matrix of the current kernels of class (pi, ai): dense (built on first use, which
        must happen outside graph capture -- BatchPlan.prepare does it) for matrix-core variants."""
        (la, lb), _, _, dp = self.pairs[pi]
        if self.lib.eip_three_center_uses_mfma(la, lb, self.aux[ai][0]) == 1:
            if "ED" not in dp:
                dp["ED"] = _dense_pair_matrix(self.lib, la, lb, dp["E"], self.device)
            return dp["ED"]
        return dp["E"]

    def prepare(self):
        """Build every table the launches read (dense pair matrices of matrix-core classes) ahead of a
        captured sequence."""
        for pi, ai in self.items:
            self.kernel_E(pi, ai)

    def n_integrals(self):
        return int(sum(n_aux * n_ao * (n_ao + 1) // 2 for n_ao, n_aux in self.sizes))

    def new_output(self):
        return self.torch.empty(self.n_out, dtype=self.torch.float64, device=self.device)

    def molecule_output(self, out, m):
        n_ao, n_aux = self.sizes[m]
        return out[self.out_base[m]:self.out_base[m + 1]].view(n_ao * (n_ao + 1) // 2, n_aux)

    def new_two_center_output(self):
        return self.torch.zeros(int(self.out2c_base[-1]), dtype=self.torch.float64, device=self.device)

    def molecule_two_center(self, out2c, m):
        n_aux = self.sizes[m][1]
        return out2c[self.out2c_base[m]:self.out2c_base[m + 1]].view(n_aux, n_aux)

    def two_center(self, out2c, stream=None):
        """Float64 (P|Q) of every molecule into its block of out2c (all blocks of the matrix are written)."""
        h = _stream_handle(stream)
        for m, (_, n_aux) in enumerate(self.sizes):
            base = out2c.data_ptr() + 8 * int(self.out2c_base[m])
            for a, (la, _, da) in enumerate(self.aux):
                for b, (lb, _, db) in enumerate(self.aux):
                    if la > lb or (la, m) not in self.mol_ks or (lb, m) not in self.mol_ks:
                        continue
                    a0, a1 = self.mol_ks[(la, m)]
                    b0, b1 = self.mol_ks[(lb, m)]
                    rc = self.lib.eip_two_center_device(
                        la, lb, _vp(da["off"].data_ptr() + 4 * a0), _ptr(da["C"]), _ptr(da["c"]),
                        _vp(da["col"].data_ptr() + 4 * a0), a1 - a0, _vp(db["off"].data_ptr() + 4 * b0), _ptr(db["C"]),
                        _ptr(db["c"]), _vp(db["col"].data_ptr() + 4 * b0), b1 - b0, int(la == lb), 1, _vp(base),
                        _c_ll(n_aux), _c_double(self.omega), h)
                    if rc != 0:
                        raise RuntimeError(f"eip_two_center_device({la},{lb}) failed with {rc}")
        return out2c

    def job_table(self, out):
        """Device table (int64 [n_classes][16], eip_j3c_job) of every (pair class, auxiliary class) of the
        batch writing into ``out``, and the row of each class; the vector pair matrices (the queue kernels'
        bodies are the vector kernels)."""
        keys = sorted(self.items)
        rows = []
        for pi, ai in keys:
            dp, da = self.pairs[pi][3], self.aux[ai][2]
            rows.append([dp["off"].data_ptr(), dp["si"].data_ptr(), dp["sj"].data_ptr(), dp["P"].data_ptr(),
                         dp["E"].data_ptr(), dp["E32"].data_ptr(), 0, da["off"].data_ptr(), da["col"].data_ptr(),
                         da["C"].data_ptr(), da["c"].data_ptr(), self.aux[ai][1], out.data_ptr(), 0,
                         dp["base"].data_ptr(), dp["ld"].data_ptr()])
        table = self.torch.tensor(np.array(rows, dtype=np.int64).reshape(-1, 16), dtype=self.torch.int64,
                                  device=self.device)
        return table, {k: i for i, k in enumerate(keys)}

    def work_queue(self, out, waves_per_simd=None, entry_cost=None):
        """The batch's (ij|P) as one WorkQueue (persistent tier kernels) writing into ``out``."""
        table, row = self.job_table(out)
        q = WorkQueue(self.lib, self.device, waves_per_simd=waves_per_simd, entry_cost=entry_cost)
        q.job_table = table                      # kept alive with the queue
        q.fallback = []                          # classes without queue bodies: their own kernels
        for (pi, ai), (it, end) in self.items.items():
            (la, lb), _, _, _ = self.pairs[pi]
            lc = self.aux[ai][0]
            if not q.supports(la, lb, lc):
                q.fallback.append((pi, ai))
                continue
            q.add(la, lb, lc, it, int(it.shape[0]), table.data_ptr() + 128 * row[(pi, ai)], self.item_costs[(pi, ai)],
                  ks_end=end, layout=0, omega=self.omega)
        return q.build()

    def zero_unwritten(self, out, stream=None):
        """Zero the packed rows of every molecule's screened-out shell pairs (no kernel writes them)."""
        if self.zero_spans.shape[0] > 0:
            rc = self.lib.eip_zero_spans(_ptr(self.zero_spans), int(self.zero_spans.shape[0]), _ptr(out), 0,
                                         _stream_handle(stream))
            if rc != 0:
                raise RuntimeError(f"eip_zero_spans failed with {rc}")

    def run_persistent(self, out, queue, graph, out2c=None, stream=None, zero=True):
        """The batch through the persistent tier kernels: ``queue`` (this batch's WorkQueue writing into
        out) replayed from ``graph`` (a QueueGraph shared by every batch), the zeroing of screened rows and
        -- when out2c is given -- every molecule's (P|Q) beside it on side streams."""
        torch = self.torch
        graph.ensure(queue, stream)
        with _fork(stream, 2) as sides:
            if zero and self.zero_spans.shape[0] > 0:
                rc = self.lib.eip_zero_spans(_ptr(self.zero_spans), int(self.zero_spans.shape[0]), _ptr(out),
                                             ZERO_BLOCKS_BESIDE_COMPUTE, _stream_handle(sides[0]))
                if rc != 0:
                    raise RuntimeError(f"eip_zero_spans failed with {rc}")
            if out2c is not None:
                self.two_center(out2c, stream=sides[1])
            for pi, ai in getattr(queue, "fallback", []):
                self._launch_class(pi, ai, out, sides[0])
            graph.run(queue, stream if stream is not None else torch.cuda.current_stream())
        return out

    def run(self, out, stream=None, zero=True, queue=None):
        """(ij|P) of every molecule into its packed block of out; rows of screened-out shell pairs are
        zeroed by a kernel (no memset: the launch sequence can be captured in a graph).  With ``queue`` (a
        WorkQueue of this batch writing into out, uploaded), the class launches are replaced by the
        persistent tier kernels over that queue."""
        with _fork(stream, int(zero and self.zero_spans.shape[0] > 0)) as sides:
            if sides:
                # concurrent with the class kernels: the zeroed rows are disjoint from the written ones
                rc = self.lib.eip_zero_spans(_ptr(self.zero_spans), int(self.zero_spans.shape[0]), _ptr(out),
                                             ZERO_BLOCKS_BESIDE_COMPUTE, _stream_handle(sides[0]))
                if rc != 0:
                    raise RuntimeError(f"eip_zero_spans failed with {rc}")
            for pi, ai in (self.items if queue is None else queue.fallback):
                self._launch_class(pi, ai, out, stream)
            if queue is not None:
                queue.run(stream)
        return out

    def _launch_class(self, pi, ai, out, stream):
        """One class of the batch through its own kernels (current variant)."""
        it, end = self.items[(pi, ai)]
        (la, lb), _, _, dp = self.pairs[pi]
        lc, n_ks, da = self.aux[ai]
        E = self.kernel_E(pi, ai)
        rc = self.lib.eip_three_center_items_batched(
            la, lb, lc, 0, _ptr(it), _ptr(end), it.shape[0], _ptr(dp["off"]), _ptr(dp["si"]), _ptr(dp["sj"]),
            _ptr(dp["base"]), _ptr(dp["ld"]), _ptr(dp["P"]), _ptr(E), _ptr(dp["E32"]), _vp(0),
            _ptr(da["off"]), _ptr(da["col"]), _ptr(da["C"]), _ptr(da["c"]), n_ks, _ptr(out),
            _c_double(self.omega), _stream_handle(stream))
        if rc != 0:
            raise RuntimeError(f"eip_three_center_items_batched({la},{lb},{lc}) failed with {rc}")


# ------------------------------------------------------------------------------------------------
# Autotuner: per three-centre class, pick the chunk-plan variant (register budget / occupancy trade-off)
# that runs fastest on the actual workload, skipping variants whose kernels spill to scratch memory.
# ------------------------------------------------------------------------------------------------

import json as _json
import pathlib as _pathlib

TUNING_FILE = _pathlib.Path(__file__).resolve().parent / "tuning" / "three_center_variants.json"
RESOURCES_FILE = _pathlib.Path(__file__).resolve().parent / "lib" / "kernel_resources.json"


def _variant_kernels(lib, la, lb, lc, v, f32):
    m0, m1 = ctypes.c_int(), ctypes.c_int()
    names, c = [], 0
    kind = lib.eip_three_center_variant_mfma(la, lb, lc, v)
    prefix = {1: "j3cm", 2: "j3cb"}.get(kind, "j3c")
    suffix = "_f32" if f32 and kind == 0 else ""
    while lib.eip_three_center_chunk(la, lb, lc, v, c, ctypes.byref(m0), ctypes.byref(m1)) == 0:
        names.append(f"{prefix}_{la}{lb}{lc}_{m0.value}_{m1.value}{suffix}")
        c += 1
    return names


def usable_variants(lib, la, lb, lc, f32=False):
    """Variants of a class none of whose kernels spill to scratch memory (variant 0 is always kept) of
    one tier: float64 matrix-core variants are float64 only, bf16 matrix-core variants float32 only."""
    from .build import spills
    res = _json.loads(RESOURCES_FILE.read_text()) if RESOURCES_FILE.exists() else {}
    out = []
    for v in range(max(0, lib.eip_three_center_nvariants(la, lb, lc))):
        if lib.eip_three_center_variant_mfma(la, lb, lc, v) == (1 if f32 else 2):
            continue
        names = _variant_kernels(lib, la, lb, lc, v, f32)
        if v == 0 or all(n in res and not spills(res[n]) for n in names):
            out.append(v)
    return out


def apply_tuning(path=TUNING_FILE):
    """Set every class's variant from the tuning file (if it exists); returns the table.  Any explicit
    call counts as the process's tuning: plans created afterwards do not apply the shipped table over it."""
    global _TUNING_APPLIED
    _TUNING_APPLIED = True
    lib = library()
    if not _pathlib.Path(path).exists():
        return {}
    table = _json.loads(_pathlib.Path(path).read_text())
    for key, v in table.get("variants", {}).items():
        cls, prec = key.split("_")
        lib.eip_set_three_center_variant(int(cls[0]), int(cls[1]), int(cls[2]), int(prec == "f32"), int(v))
    return table


def autotune(plan, repeats=3, use_f32=False, save=None, log=None, block_bytes=None):
    """Time every usable variant of every class of ``plan`` and keep the fastest.

    Classes are timed on the packed output, or -- with ``block_bytes`` (molecules whose output exceeds
    device memory) -- on the first block of shell pairs of each pair class (compact rows)."""
    torch = plan.torch
    lib = plan.lib
    chosen = {}
    stream = torch.cuda.current_stream()
    if block_bytes is None:
        out = plan.new_output()
        launch = lambda pi, ai: plan.launch_class(pi, ai, out, use_f32=use_f32)
    else:
        firsts = {}
        for b in plan.row_blocks(block_bytes):
            firsts.setdefault(b[0], b)
        rows = max(plan.block_rows(*b) for b in firsts.values())
        buf = torch.empty((rows, plan.n_aux), dtype=torch.float64, device=plan.device)

        def launch(pi, ai):
            _, sp0, sp1 = firsts[pi]
            pc = plan.pairs[pi]
            nij = (2 * pc.la + 1) * (2 * pc.lb + 1)
            jobs = plan.job_table([plan.job_words(pi, ai, buf.data_ptr() - 8 * sp0 * nij * plan.n_aux, plan.n_aux)])
            G = (plan.aux[ai].n_ks + 63) // 64
            plan.launch_work(pi, ai, plan.items[(pi, ai)][sp0 * G:sp1 * G], jobs, layout=1, use_f32=use_f32)
    for pi, pc in enumerate(plan.pairs):
        for ai, ac in enumerate(plan.aux):
            key = f"{pc.la}{pc.lb}{ac.lc}_{'f32' if use_f32 else 'f64'}"
            best = None
            for v in usable_variants(lib, pc.la, pc.lb, ac.lc, use_f32):
                lib.eip_set_three_center_variant(pc.la, pc.lb, ac.lc, int(use_f32), v)
                launch(pi, ai)                                        # warm-up
                times = []
                for _ in range(repeats):
                    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                    e0.record(stream)
                    launch(pi, ai)
                    e1.record(stream)
                    e1.synchronize()
                    times.append(e0.elapsed_time(e1))
                t = min(times)
                if log:
                    log(f"  autotune {key} variant {v}: {t:.3f} ms")
                if best is None or t < best[1]:
                    best = (v, t)
            lib.eip_set_three_center_variant(pc.la, pc.lb, ac.lc, int(use_f32), best[0])
            chosen[key] = best[0]
    if save:
        path = _pathlib.Path(save)
        path.parent.mkdir(parents=True, exist_ok=True)
        old = _json.loads(path.read_text()) if path.exists() else {"variants": {}}
        old["variants"].update(chosen)
        path.write_text(_json.dumps(old, indent=1, sort_keys=True))
    return chosen


def autotune_in_step(lib, classes, step, use_f32=False, repeats=2, save=None, log=None):
    """Variant choice by the time of the whole step the classes run in, not of each class alone.

    ``step()`` runs one full step (every class, on however many streams it uses) and returns when it is
    done; ``classes`` are (la, lb, lc) in the order to tune (the costliest first).  Coordinate descent:
    for each class in turn, every usable variant of the class (the others at their current choice) is
    timed over ``repeats`` steps (the fastest counts), and the fastest kept.  Timed alone, a class often
    prefers the variant with the most waves per SIMD; inside a step whose streams already fill the
    device with other classes' waves, the variant with the least work can win instead.  Returns
    {"<la><lb><lc>_<f32|f64>": variant}."""
    import time
    chosen = {}
    prec = "f32" if use_f32 else "f64"

    def timed():
        best = None
        for _ in range(repeats):
            t0 = time.perf_counter()
            step()
            t = time.perf_counter() - t0
            best = t if best is None else min(best, t)
        return best

    for la, lb, lc in classes:
        key = f"{la}{lb}{lc}_{prec}"
        cur = lib.eip_three_center_variant(la, lb, lc, int(use_f32))
        # the current choice stays a candidate even where a kernel of it reaches scratch memory (a table
        # entry measured faster than the variants that do not)
        cands = sorted(set(usable_variants(lib, la, lb, lc, use_f32)) | ({cur} if cur >= 0 else set()))
        if len(cands) < 2:
            continue
        res = {}
        for v in cands:
            lib.eip_set_three_center_variant(la, lb, lc, int(use_f32), v)
            res[v] = timed()
        best = min(res, key=res.get)
        # keep the current choice unless another is clearly (1 %) faster: step times vary by that much
        if cur in res and res[cur] <= 1.01 * res[best]:
            best = cur
        lib.eip_set_three_center_variant(la, lb, lc, int(use_f32), best)
        chosen[key] = best
        if log:
            log(f"  step-tune {key}: " + ", ".join(f"v{v} {t * 1e3:.1f} ms" for v, t in res.items()) + f" -> v{best}")
    if save:
        path = _pathlib.Path(save)
        path.parent.mkdir(parents=True, exist_ok=True)
        old = _json.loads(path.read_text()) if path.exists() else {"variants": {}}
        old["variants"].update(chosen)
        path.write_text(_json.dumps(old, indent=1, sort_keys=True))
    return chosen


def autotune_queue(batch, out, waves=(0, 2, 1), entry_costs=(1.0e4, 3.0e4, 1.0e5), repeats=3, save=None, log=None):
    """Pick the persistent queue's grid occupancy cap (waves per SIMD; 0 = as many as fit) and entry
    grouping (work per entry) for a BatchPlan ``batch`` writing into ``out``: every combination is built,
    uploaded and timed (HIP events, fastest of ``repeats`` after a warm-up); the fastest is kept in
    QUEUE_TUNING (and written to ``save``).  Returns {(waves, entry_cost): ms}."""
    import torch
    stream = torch.cuda.current_stream()
    times = {}
    for w in waves:
        for c in entry_costs:
            q = batch.work_queue(out, waves_per_simd=w, entry_cost=c)
            q.upload()
            batch.run(out, queue=q, zero=False)
            best = None
            for _ in range(repeats):
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record(stream)
                batch.run(out, queue=q, zero=False)
                e1.record(stream)
                e1.synchronize()
                best = e0.elapsed_time(e1) if best is None else min(best, e0.elapsed_time(e1))
            times[(w, c)] = best
            if log:
                log(f"  autotune queue waves_per_simd {w} entry_cost {c:.0e}: {best:.2f} ms")
    (w, c), _ = min(times.items(), key=lambda kv: kv[1])
    QUEUE_TUNING["f64"] = {"waves_per_simd": int(w), "entry_cost": float(c)}
    if save:
        path = _pathlib.Path(save)
        path.parent.mkdir(parents=True, exist_ok=True)
        path.write_text(_json.dumps(QUEUE_TUNING, indent=1, sort_keys=True))
    return times
