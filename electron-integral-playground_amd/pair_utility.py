"""Primitive-pair screening and structure-of-arrays packing, vectorised over shell pairs.

Public functions keep the reference's names, arguments and results (pair_utility.py:6-229):
``form_primitive_pair_list``, ``form_primitive_pair_data``, ``form_pair_data``,
``form_primitive_auxiliary_pair_data``, ``form_auxiliary_pair_data``.

Screening decisions must be bit-identical to the reference, so every floating-point value is formed
with the same numpy operations and the same association order as there; only the loop over shell
pairs is replaced by array operations over all shell pairs that share a primitive-count signature.
Two reference behaviours are reproduced on purpose:

* the distance used in the bound is shortened *cumulatively* inside the primitive loops and clamped at
  zero after every inner step (pair_utility.py:85, 99, 107, 109), so a pair's bound depends on the
  order of the primitives before it;
* the per-class list is ordered by a *stable* descending sort of the bound (pair_utility.py:131),
  with ties kept in loop order (i_shell, j_shell, i_primitive, j_primitive).

``form_pair_data`` ignores its threshold argument and screens at 1e-14, as the reference does
(pair_utility.py:185-188).
"""

from collections import OrderedDict

import numpy as np

from .data_structure import Molecule, PrimitivePair, PrimitivePairData

REFERENCE_SCREENING_THRESHOLD = 1e-14


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


def _screen_group(li, lj, exp_i, coef_i, exp_j, coef_j, dist, threshold):
    """Bounds for a group of shell pairs with equal primitive counts.

    exp_i/coef_i: (n_group, ni); exp_j/coef_j: (n_group, nj); dist: (n_group,).
    Returns ``bound`` (n_group, ni, nj) and ``keep`` mask.
    """
    n_group, ni = exp_i.shape
    nj = exp_j.shape[1]
    cA = np.abs(coef_i)
    if li > 0:
        cA = cA * np.power(li / (2.0 * np.e * exp_i), li / 2.0)
    cB = np.abs(coef_j)
    if lj > 0:
        cB = cB * np.power(lj / (2.0 * np.e * exp_j), lj / 2.0)
    shrink_i = np.sqrt(li / 2.0 / exp_i)
    shrink_j = np.sqrt(lj / 2.0 / exp_j)
    d = dist.copy()
    bound = np.empty((n_group, ni, nj))
    for ip in range(ni):
        a = exp_i[:, ip]
        d = d - shrink_i[:, ip]
        for jp in range(nj):
            b = exp_j[:, jp]
            d = d - shrink_j[:, jp]
            d = np.where(0.0 > d, 0.0, d)          # Python max(d, 0.0)
            gaussian = np.exp(-a * b / (a + b) * d * d)
            c_ab = 2.0 * cA[:, ip] * cB[:, jp]
            overlap_factor = np.power(np.pi / (a + b), 1.5)
            schwarz_factor = np.power(np.pi / (a + b), 1.25) * np.power(2, 0.25)
            bound[:, ip, jp] = c_ab * gaussian * np.maximum(overlap_factor, schwarz_factor)
    return bound, ~(bound < threshold)


def form_primitive_pair_list(molecule: Molecule, schwarz_upper_bound: float):
    """Screened primitive pairs per angular class ``(li, lj)`` with ``li <= lj``.

    Returns ``dict[(li, lj)] -> list[PrimitivePair]`` in the reference's key and element order.
    """
    shells = molecule.basis_shells
    assert shells is not None and len(shells) > 0 and molecule.n_ao > 0
    norm, _ = _atom_pair_distance(molecule.geometry)

    # Shell pairs admitted by the class rule (pair_utility.py:87-91), bucketed by primitive counts.
    buckets = {}
    for i_shell, si in enumerate(shells):
        for j_shell, sj in enumerate(shells):
            if si.angular > sj.angular:
                continue
            if si.angular == sj.angular and si.i_ao_start > sj.i_ao_start:
                continue
            key = (si.angular, sj.angular, len(si.primitive_exponents), len(sj.primitive_exponents))
            buckets.setdefault(key, []).append((i_shell, j_shell))

    kept = []   # rows: (i_shell, j_shell, ip, jp, bound, li, lj)
    for (li, lj, ni, nj), members in buckets.items():
        idx = np.array(members, dtype=np.int64)
        exp_i = np.stack([shells[s].primitive_exponents for s in idx[:, 0]])
        coef_i = np.stack([shells[s].primitive_coefficients for s in idx[:, 0]])
        exp_j = np.stack([shells[s].primitive_exponents for s in idx[:, 1]])
        coef_j = np.stack([shells[s].primitive_coefficients for s in idx[:, 1]])
        atoms_i = np.array([shells[s].i_atom for s in idx[:, 0]])
        atoms_j = np.array([shells[s].i_atom for s in idx[:, 1]])
        bound, keep = _screen_group(li, lj, exp_i, coef_i, exp_j, coef_j,
                                    norm[atoms_i, atoms_j], schwarz_upper_bound)
        g, ip, jp = np.nonzero(keep)
        for row in zip(idx[g, 0], idx[g, 1], ip, jp, bound[g, ip, jp]):
            kept.append((int(row[0]), int(row[1]), int(row[2]), int(row[3]), row[4], li, lj))

    # Reference append order is lexicographic in (i_shell, j_shell, i_primitive, j_primitive).
    kept.sort(key=lambda r: (r[0], r[1], r[2], r[3]))
    pair_list = OrderedDict()
    for i_shell, j_shell, ip, jp, ub, li, lj in kept:
        pair_list.setdefault((li, lj), []).append(PrimitivePair(i_shell, j_shell, ip, jp, ub))
    for key in pair_list:
        pair_list[key].sort(key=lambda p: p.upper_bound, reverse=True)   # stable, larger first
    return pair_list


def form_primitive_pair_data(molecule: Molecule, pair_list):
    """Pack each class's pairs into ``PrimitivePairData`` (reference pair_utility.py:135-183).

    ``P = (aA + bB)/p`` and ``coefficient = c_a c_b exp(-ab/p |AB|^2)``.
    """
    shells = molecule.basis_shells
    assert shells is not None and len(shells) > 0 and molecule.n_ao > 0
    _, sq = _atom_pair_distance(molecule.geometry)
    geometry = molecule.geometry
    out = OrderedDict()
    for key, pairs in pair_list.items():
        n = len(pairs)
        si = [shells[p.i_shell] for p in pairs]
        sj = [shells[p.j_shell] for p in pairs]
        atom_i = np.array([s.i_atom for s in si], dtype=np.int64)
        atom_j = np.array([s.i_atom for s in sj], dtype=np.int64)
        a = np.array([s.primitive_exponents[p.i_primitive] for s, p in zip(si, pairs)])
        b = np.array([s.primitive_exponents[p.j_primitive] for s, p in zip(sj, pairs)])
        ca = np.array([s.primitive_coefficients[p.i_primitive] for s, p in zip(si, pairs)])
        cb = np.array([s.primitive_coefficients[p.j_primitive] for s, p in zip(sj, pairs)])
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
            i_ao_start=np.array([s.i_ao_start for s in si], dtype=np.int32),
            j_ao_start=np.array([s.i_ao_start for s in sj], dtype=np.int32),
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
    return form_primitive_pair_data(
        molecule, form_primitive_pair_list(molecule, REFERENCE_SCREENING_THRESHOLD))


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
