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


def group_pairs(pd):
    """Stable grouping of one class's primitive pairs by (i_ao_start, j_ao_start)."""
    order = np.lexsort((pd.j_ao_start, pd.i_ao_start))
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


class ThreeCenterPlan:
    """All classes of one molecule's (ij|P), resident on one GPU."""

    def __init__(self, molecule, pair_data, aux_data, omega=0.0, device="cuda"):
        import torch
        self.torch = torch
        self.lib = library()
        self.n_ao, self.n_aux = molecule.n_ao, molecule.n_aux
        self.omega = float(omega)
        self.device = torch.device(device)
        if not molecule.spherical_basis:
            raise ValueError("the device path produces spherical output")
        lib = self.lib
        T = lambda a, dt: torch.as_tensor(np.ascontiguousarray(a), dtype=dt, device=self.device)
        self.pairs: List[PairClass] = []
        stream = _stream_handle(None)
        for (la, lb), pd in pair_data.items():
            nnz = lib.eip_pair_matrix_nnz(la, lb)
            if nnz < 0:
                raise ValueError(f"no specialised kernels for pair class ({la},{lb})")
            order, si, sj, off = group_pairs(pd)
            pc = PairClass(la, lb, len(order), len(si), si, sj, off)
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
        # work items: every (shell pair, block of 64 auxiliary shells)
        self.items: Dict[Tuple[int, int], object] = {}
        for pi, pc in enumerate(self.pairs):
            for ai, ac in enumerate(self.aux):
                kb = (ac.n_ks + 63) // 64
                sp = np.repeat(np.arange(pc.n_sp, dtype=np.int32), kb)
                k0 = np.tile(np.arange(kb, dtype=np.int32) * 64, pc.n_sp)
                self.items[(pi, ai)] = T(np.stack([sp, k0], axis=1), torch.int32)
        torch.cuda.current_stream().synchronize()

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

    # -- execution -------------------------------------------------------------------------------
    def new_output(self):
        return self.torch.empty((self.n_packed, self.n_aux), dtype=self.torch.float64, device=self.device)

    def launch_class(self, pi, ai, out, stream=None, use_f32=False):
        pc, ac = self.pairs[pi], self.aux[ai]
        it = self.items[(pi, ai)]
        dp, da = pc.dev, ac.dev
        rc = self.lib.eip_three_center_items(
            pc.la, pc.lb, ac.lc, int(use_f32), _ptr(it), _vp(0), it.shape[0], _ptr(dp["off"]), _ptr(dp["si"]),
            _ptr(dp["sj"]), _ptr(dp["P"]), _ptr(dp["E"]), _ptr(dp["E32"]), _vp(0), _ptr(da["off"]),
            _ptr(da["col"]), _ptr(da["C"]), _ptr(da["c"]), ac.n_ks, _ptr(out), _c_ll(self.n_aux), 0,
            _c_double(self.omega), _stream_handle(stream))
        if rc != 0:
            raise RuntimeError(f"eip_three_center_items({pc.la},{pc.lb},{ac.lc}) failed with {rc}")

    def run(self, out=None, stream=None, zero=True):
        """All classes into ``out`` (packed layout); rows of screened-out shell pairs are zero."""
        if out is None:
            out = self.new_output()
        if zero:
            out.zero_()
        for pi in range(len(self.pairs)):
            for ai in range(len(self.aux)):
                self.launch_class(pi, ai, out, stream)
        return out


def packed_to_dense(packed, n_ao):
    """(n_ao(n_ao+1)/2, n_aux) packed -> (n_aux, n_ao, n_ao) dense (reference layout), on the device."""
    import torch
    idx_i, idx_j = torch.tril_indices(n_ao, n_ao, device=packed.device)
    n_aux = packed.shape[1]
    dense = torch.zeros((n_aux, n_ao, n_ao), dtype=packed.dtype, device=packed.device)
    dense[:, idx_i, idx_j] = packed.t()
    dense[:, idx_j, idx_i] = packed.t()
    return dense
