"""Primitive-pair screening and structure-of-arrays packing.

Public functions keep the reference's names, arguments and results (pair_utility.py:6-229):
``form_primitive_pair_list``, ``form_primitive_pair_data``, ``form_pair_data``,
``form_primitive_auxiliary_pair_data``, ``form_auxiliary_pair_data``; ``form_pair_arrays`` is the same
list as arrays (no per-pair Python objects), which the device path uses.

Screening decisions must be bit-identical to the reference.  The O(N_shell^2 N_prim^2) loop
(pair_utility.py:78-124) runs natively in the library (``eip_screen_primitive_pairs``): shell-pair
admission, the cumulative distance rule and a conservative pre-screen of the bound.  The exact bound of
every surviving candidate is then evaluated here with the same numpy operations, in the same
association order, as the reference, and the exact keep test and ordering are applied to those values.
Two reference behaviours are reproduced on purpose:

* the distance used in the bound is shortened *cumulatively* inside the primitive loops and clamped at
  zero after every inner step (pair_utility.py:85, 99, 107, 109), so a pair's bound depends on the
  order of the primitives before it;
* the per-class list is ordered by a *stable* descending sort of the bound (pair_utility.py:131),
  with ties kept in loop order (i_shell, j_shell, i_primitive, j_primitive).

``form_pair_data`` ignores its threshold argument and screens at 1e-14, as the reference does
(pair_utility.py:185-188).
"""

import ctypes
from collections import OrderedDict
from dataclasses import dataclass

import numpy as np

from . import library
from .data_structure import Molecule, PrimitivePair, PrimitivePairData

REFERENCE_SCREENING_THRESHOLD = 1e-14
# relative slack of the native pre-screen (its libm exp/pow differ from numpy's by a few ulp at most)
PRESCREEN_MARGIN = 1e-9


@dataclass
class PairArrays:
    """One angular class of the screened list, in the reference's order (descending bound)."""
    i_shell: np.ndarray       # int64
    j_shell: np.ndarray
    i_primitive: np.ndarray
    j_primitive: np.ndarray
    upper_bound: np.ndarray   # float64

    def __len__(self):
        return len(self.upper_bound)


def _atom_pair_distance(geometry: np.ndarray):
    """|A-B| and |A-B|^2 per atom pair, each formed by the same numpy call the reference makes."""
    n = geometry.shape[0]
    norm = np.empty((n, n))
    sq = np.empty((n, n))
    for a in range(n):
        for b in range(n):
            diff = geometry[a, :] - geometry[b, :]
            norm[a, b] = np.linalg.norm(diff)
            sq[a, b] = diff @ diff
    return norm, sq


def _shell_arrays(shells):
    l = np.array([s.angular for s in shells], dtype=np.int32)
    ao = np.array([s.i_ao_start for s in shells], dtype=np.int32)
    atom = np.array([s.i_atom for s in shells], dtype=np.int32)
    n_prim = np.array([len(s.primitive_exponents) for s in shells], dtype=np.int64)
    off = np.concatenate([[0], np.cumsum(n_prim)]).astype(np.int32)
    exps = np.concatenate([np.asarray(s.primitive_exponents, dtype=np.float64) for s in shells])
    coefs = np.concatenate([np.asarray(s.primitive_coefficients, dtype=np.float64) for s in shells])
    return l, ao, atom, off, exps, coefs


def _native_candidates(shell_arrays, norm, threshold):
    """(i_shell, j_shell, i_prim, j_prim, clamped distance) of every pre-screened candidate, loop order."""
    lib = library()
    l, ao, atom, off, exps, coefs = shell_arrays
    norm = np.ascontiguousarray(norm, dtype=np.float64)
    ip_ = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
    dp_ = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    fn = lib.eip_screen_primitive_pairs
    fn.restype = ctypes.c_longlong
    args = (len(l), ip_(l), ip_(ao), ip_(atom), ip_(off), dp_(exps), dp_(coefs), dp_(norm), norm.shape[0],
            ctypes.c_double(threshold), ctypes.c_double(PRESCREEN_MARGIN))
    n = fn(*args, ctypes.c_longlong(0), None, None, None, None, None)
    if n < 0:
        raise RuntimeError("eip_screen_primitive_pairs rejected its arguments")
    out = [np.empty(n, dtype=np.int32) for _ in range(4)] + [np.empty(n, dtype=np.float64)]
    m = fn(*args, ctypes.c_longlong(n), *(ip_(a) for a in out[:4]), dp_(out[4]))
    if m != n:
        raise RuntimeError(f"eip_screen_primitive_pairs returned {m} candidates, expected {n}")
    return out


def _exact_bound(li, lj, a, ca, b, cb, d):
    """The reference's bound (pair_utility.py:95-114), same numpy operations and association order."""
    cA = np.abs(ca)
    if li > 0:
        cA = cA * np.power(li / (2.0 * np.e * a), li / 2.0)
    cB = np.abs(cb)
    if lj > 0:
        cB = cB * np.power(lj / (2.0 * np.e * b), lj / 2.0)
    gaussian = np.exp(-a * b / (a + b) * d * d)
    c_ab = 2.0 * cA * cB
    overlap_factor = np.power(np.pi / (a + b), 1.5)
    schwarz_factor = np.power(np.pi / (a + b), 1.25) * np.power(2, 0.25)
    return c_ab * gaussian * np.maximum(overlap_factor, schwarz_factor)


def form_pair_arrays(molecule: Molecule, schwarz_upper_bound: float):
    """Screened primitive pairs per angular class ``(li, lj)`` as ``PairArrays``, in the reference's key
    order (first appearance in the loop) and element order (stable descending bound)."""
    shells = molecule.basis_shells
    assert shells is not None and len(shells) > 0 and molecule.n_ao > 0
    norm, _ = _atom_pair_distance(molecule.geometry)
    sa = _shell_arrays(shells)
    l, _, _, off, exps, coefs = sa
    ci, cj, cip, cjp, cd = _native_candidates(sa, norm, schwarz_upper_bound)
    li_all, lj_all = l[ci], l[cj]
    ka = off[ci].astype(np.int64) + cip
    kb = off[cj].astype(np.int64) + cjp
    keep = np.zeros(len(ci), dtype=bool)
    ub = np.empty(len(ci))
    code = li_all.astype(np.int64) * 64 + lj_all
    for c in np.unique(code):
        li, lj = int(c) // 64, int(c) % 64
        sel = np.flatnonzero(code == c)
        v = _exact_bound(li, lj, exps[ka[sel]], coefs[ka[sel]], exps[kb[sel]], coefs[kb[sel]], cd[sel])
        ub[sel] = v
        keep[sel] = ~(v < schwarz_upper_bound)
    kept = np.flatnonzero(keep)
    # class order = order of each class's first kept pair in the loop
    cls = code[kept]
    _, first = np.unique(cls, return_index=True)
    out = OrderedDict()
    for f in np.sort(first):
        key = (int(li_all[kept[f]]), int(lj_all[kept[f]]))
        rows = kept[cls == key[0] * 64 + key[1]]
        rows = rows[np.argsort(-ub[rows], kind="stable")]
        out[key] = PairArrays(ci[rows].astype(np.int64), cj[rows].astype(np.int64), cip[rows].astype(np.int64),
                              cjp[rows].astype(np.int64), ub[rows])
    return out


def form_primitive_pair_list(molecule: Molecule, schwarz_upper_bound: float):
    """Screened primitive pairs per angular class ``(li, lj)`` with ``li <= lj``.

    Returns ``dict[(li, lj)] -> list[PrimitivePair]`` in the reference's key and element order.
    """
    out = OrderedDict()
    for key, pa in form_pair_arrays(molecule, schwarz_upper_bound).items():
        out[key] = [PrimitivePair(int(i), int(j), int(ip), int(jp), ub) for i, j, ip, jp, ub in
                    zip(pa.i_shell, pa.j_shell, pa.i_primitive, pa.j_primitive, pa.upper_bound)]
    return out


def _as_arrays(pairs):
    if isinstance(pairs, PairArrays):
        return pairs
    return PairArrays(np.array([p.i_shell for p in pairs], dtype=np.int64),
                      np.array([p.j_shell for p in pairs], dtype=np.int64),
                      np.array([p.i_primitive for p in pairs], dtype=np.int64),
                      np.array([p.j_primitive for p in pairs], dtype=np.int64),
                      np.array([p.upper_bound for p in pairs], dtype=np.float64))


def form_primitive_pair_data(molecule: Molecule, pair_list):
    """Pack each class's pairs into ``PrimitivePairData`` (reference pair_utility.py:135-183).

    ``pair_list`` holds ``list[PrimitivePair]`` or ``PairArrays`` per class.
    ``P = (aA + bB)/p`` and ``coefficient = c_a c_b exp(-ab/p |AB|^2)``.
    """
    shells = molecule.basis_shells
    assert shells is not None and len(shells) > 0 and molecule.n_ao > 0
    _, sq = _atom_pair_distance(molecule.geometry)
    geometry = molecule.geometry
    _, ao, atom, off, exps, coefs = _shell_arrays(shells)
    out = OrderedDict()
    for key, pairs in pair_list.items():
        pa = _as_arrays(pairs)
        n = len(pa)
        atom_i = atom[pa.i_shell].astype(np.int64)
        atom_j = atom[pa.j_shell].astype(np.int64)
        a = exps[off[pa.i_shell] + pa.i_primitive]
        b = exps[off[pa.j_shell] + pa.j_primitive]
        ca = coefs[off[pa.i_shell] + pa.i_primitive]
        cb = coefs[off[pa.j_shell] + pa.j_primitive]
        A = geometry[atom_i, :]
        B = geometry[atom_j, :]
        p = a + b
        P = (a[:, None] * A + b[:, None] * B) / p[:, None]
        gaussian = np.exp(-a * b / p * sq[atom_i, atom_j])
        data = PrimitivePairData(
            P_p=np.empty((n, 4), dtype=np.float64, order="C"),
            A_a=np.empty((n, 4), dtype=np.float64, order="C"),
            B_b=np.empty((n, 4), dtype=np.float64, order="C"),
            coefficient=ca * cb * gaussian,
            i_ao_start=ao[pa.i_shell].astype(np.int32),
            j_ao_start=ao[pa.j_shell].astype(np.int32),
            i_atom=atom_i.astype(np.int32),
            j_atom=atom_j.astype(np.int32),
        )
        data.P_p[:, 0:3] = P
        data.P_p[:, 3] = p
        data.A_a[:, 0:3] = A
        data.A_a[:, 3] = a
        data.B_b[:, 0:3] = B
        data.B_b[:, 3] = b
        out[key] = data
    return out


def form_pair_data(molecule: Molecule, schwarz_upper_bound: float):
    """Screen at the reference's fixed 1e-14 (the argument is ignored, pair_utility.py:185-188)."""
    del schwarz_upper_bound
    return form_primitive_pair_data(molecule, form_pair_arrays(molecule, REFERENCE_SCREENING_THRESHOLD))


def form_primitive_auxiliary_pair_data(molecule: Molecule):
    """Auxiliary primitives per angular value (reference pair_utility.py:190-226).

    Keys follow iteration over the ``set`` of angular values, as in the reference; each entry holds
    ``A_a = [C; c]``, ``coefficient = c`` and ``i_ao_start`` of the owning shell.
    """
    shells = molecule.auxiliary_basis_shells
    assert shells is not None and len(shells) > 0 and molecule.n_aux > 0
    out = OrderedDict()
    for l in set(s.angular for s in shells):
        same = [s for s in shells if s.angular == l]
        A_a = np.concatenate([
            np.column_stack([np.repeat(molecule.geometry[s.i_atom:s.i_atom + 1, :],
                                       len(s.primitive_exponents), axis=0),
                             s.primitive_exponents]) for s in same])
        out[l] = PrimitivePairData(
            P_p=None,
            A_a=np.ascontiguousarray(A_a, dtype=np.float64),
            B_b=None,
            coefficient=np.concatenate([np.asarray(s.primitive_coefficients, dtype=np.float64)
                                        for s in same]),
            i_ao_start=np.concatenate([np.full(len(s.primitive_exponents), s.i_ao_start,
                                               dtype=np.int32) for s in same]),
            j_ao_start=None,
            i_atom=np.concatenate([np.full(len(s.primitive_exponents), s.i_atom, dtype=np.int32)
                                   for s in same]),
            j_atom=None,
        )
    return out


def form_auxiliary_pair_data(molecule: Molecule):
    return form_primitive_auxiliary_pair_data(molecule)
