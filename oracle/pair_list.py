"""Scalar restatement of the reference's primitive-pair screening and packing.

TEST INFRASTRUCTURE ONLY: the checker for the package's vectorised pair_utility (and any native pair
builder).  Only tests/ may import it; it is never used by the product.

Follows reference pair_utility.py line by line in loop structure and floating-point operation order:
  form_primitive_pair_list      <- pair_utility.py:6-133  (cumulative distance shortening :85, 99, 107,
                                   clamp :109, bound :110-114, keep test :116, stable descending sort :131)
  form_primitive_pair_data      <- pair_utility.py:135-183
  form_auxiliary_pair_data      <- pair_utility.py:190-226
The screening arithmetic is numpy scalar arithmetic, as in the reference, so decisions are the
reference's bit for bit.
"""

import numpy as np


class Pair:
    __slots__ = ("i_shell", "j_shell", "i_primitive", "j_primitive", "upper_bound")

    def __init__(self, i_shell, j_shell, i_primitive, j_primitive, upper_bound):
        self.i_shell, self.j_shell = i_shell, j_shell
        self.i_primitive, self.j_primitive = i_primitive, j_primitive
        self.upper_bound = upper_bound


def form_primitive_pair_list(molecule, threshold):
    pair_list = {}
    for i_shell, shell_i in enumerate(molecule.basis_shells):
        li = shell_i.angular
        A = molecule.geometry[shell_i.i_atom, :]
        for j_shell, shell_j in enumerate(molecule.basis_shells):
            lj = shell_j.angular
            B = molecule.geometry[shell_j.i_atom, :]
            d = np.linalg.norm(A - B)
            if li > lj:
                continue
            if li == lj and shell_i.i_ao_start > shell_j.i_ao_start:
                continue
            for ip in range(len(shell_i.primitive_exponents)):
                a = shell_i.primitive_exponents[ip]
                cA = np.abs(shell_i.primitive_coefficients[ip])
                if li > 0:
                    cA *= np.power(li / (2.0 * np.e * a), li / 2.0)
                d -= np.sqrt(li / 2.0 / a)
                for jp in range(len(shell_j.primitive_exponents)):
                    b = shell_j.primitive_exponents[jp]
                    cB = np.abs(shell_j.primitive_coefficients[jp])
                    if lj > 0:
                        cB *= np.power(lj / (2.0 * np.e * b), lj / 2.0)
                    d -= np.sqrt(lj / 2.0 / b)
                    d = max(d, 0.0)
                    g = np.exp(-a * b / (a + b) * d * d)
                    cAB = 2.0 * cA * cB
                    overlap_factor = np.power(np.pi / (a + b), 1.5)
                    schwarz_factor = np.power(np.pi / (a + b), 1.25) * np.power(2, 0.25)
                    ub = cAB * g * max(overlap_factor, schwarz_factor)
                    if ub < threshold:
                        continue
                    pair_list.setdefault((li, lj), []).append(Pair(i_shell, j_shell, ip, jp, ub))
    for key in pair_list:
        pair_list[key].sort(key=lambda p: p.upper_bound, reverse=True)
    return pair_list


def form_primitive_pair_data(molecule, pair_list):
    out = {}
    for key, pairs in pair_list.items():
        n = len(pairs)
        P_p = np.empty((n, 4)); A_a = np.empty((n, 4)); B_b = np.empty((n, 4))
        coef = np.empty(n)
        i_start = np.empty(n, dtype=np.int32); j_start = np.empty(n, dtype=np.int32)
        for k, pr in enumerate(pairs):
            si = molecule.basis_shells[pr.i_shell]
            sj = molecule.basis_shells[pr.j_shell]
            A = molecule.geometry[si.i_atom, :]
            B = molecule.geometry[sj.i_atom, :]
            a = si.primitive_exponents[pr.i_primitive]
            b = sj.primitive_exponents[pr.j_primitive]
            ca = si.primitive_coefficients[pr.i_primitive]
            cb = sj.primitive_coefficients[pr.j_primitive]
            p = a + b
            P = (a * A + b * B) / p
            d2 = (A - B) @ (A - B)
            g = np.exp(-a * b / p * d2)
            A_a[k, 0:3] = A; A_a[k, 3] = a
            B_b[k, 0:3] = B; B_b[k, 3] = b
            P_p[k, 0:3] = P; P_p[k, 3] = p
            coef[k] = ca * cb * g
            i_start[k] = si.i_ao_start
            j_start[k] = sj.i_ao_start
        out[key] = dict(P_p=P_p, A_a=A_a, B_b=B_b, coefficient=coef, i_ao_start=i_start, j_ao_start=j_start)
    return out
