"""Contracted-shell normalisation and AO index assignment.

Numbers produced here feed the integral coefficients, so the floating-point operation sequence of
``normalize_shell`` is kept identical to the reference's (basis_utility.py:131-140): the same numpy
calls on the same array shapes, in the same order.  That is what makes coefficients, and through them
every integral, bit-identical to the reference's host layer.
"""

import copy
from enum import Enum
from math import factorial

import numpy as np

from .data_structure import BasisSet, GaussianShell, Molecule


def n_ao_of_angular(l: int, spherical: bool = True) -> int:
    """Functions per shell: 2l+1 spherical or (l+1)(l+2)/2 Cartesian (basis_utility.py:18-22)."""
    return 2 * l + 1 if spherical else (l + 1) * (l + 2) // 2


class AtomicOrbitalOrder(Enum):
    ATOM_LEADING = 1
    ANGULAR_LEADING = 2


def normalize_shell(shell: GaussianShell) -> None:
    """Scale ``shell.primitive_coefficients`` so the pure-x Cartesian component has unit norm.

    Each primitive first gets its exponent-dependent factor a^(L/2+3/4), then the contraction is
    normalised with S = pi^(3/2) (2L)! / (4^L L!) * sum_mn c_m c_n (a_m+a_n)^(-L-3/2).
    Not idempotent (reference basis_utility.py:85-140).
    """
    exps = shell.primitive_exponents
    assert isinstance(exps, np.ndarray) and exps.ndim == 1 and exps.shape[0] > 0
    assert isinstance(shell.primitive_coefficients, np.ndarray)
    assert shell.primitive_coefficients.shape == exps.shape
    L = shell.angular
    assert isinstance(L, int) and L >= 0
    coef = shell.primitive_coefficients * np.power(exps, L * 0.5 + 0.75)
    prefactor = np.power(np.pi, 1.5) * factorial(2 * L) / (4 ** L * factorial(L))
    pair_sum = np.power(exps[:, None] + exps[None, :], -L - 1.5)
    norm = prefactor * np.sum(np.outer(coef, coef) * pair_sum)
    coef *= np.power(norm, -0.5)
    shell.primitive_coefficients = coef


def _assign(shells, order: AtomicOrbitalOrder, spherical: bool) -> int:
    assert len(shells) > 0
    if order == AtomicOrbitalOrder.ATOM_LEADING:
        shells.sort(key=lambda s: s.i_atom)          # stable: keeps basis-file order per atom
    elif order == AtomicOrbitalOrder.ANGULAR_LEADING:
        shells.sort(key=lambda s: s.angular)
    else:
        raise NotImplementedError(f"Unsupported atomic orbital order {order}")
    offset = 0
    for shell in shells:
        shell.i_ao_start = offset
        offset += n_ao_of_angular(shell.angular, spherical)
    return offset


def assign_ao_index_to_shell(molecule: Molecule,
                             ao_order: AtomicOrbitalOrder = AtomicOrbitalOrder.ATOM_LEADING) -> None:
    """Give every orbital and auxiliary shell its AO offset (reference basis_utility.py:61-82)."""
    done = False
    if molecule.basis_shells:
        molecule.n_ao = _assign(molecule.basis_shells, ao_order, molecule.spherical_basis)
        done = True
    if molecule.auxiliary_basis_shells:
        molecule.n_aux = _assign(molecule.auxiliary_basis_shells, ao_order, molecule.spherical_basis)
        done = True
    assert done


def attach_basis_to_molecule(molecule: Molecule, basis_set: BasisSet,
                             if_auxiliary_basis: bool = False) -> None:
    """Normalise a per-element basis once, replicate it on every atom, assign AO offsets.

    Reference basis_utility.py:24-52.
    """
    used = {}
    for element in molecule.elements:
        assert element in basis_set, f"no basis for element {element}"
        if element not in used:
            used[element] = copy.deepcopy(basis_set[element])
    for shells in used.values():
        for shell in shells:
            normalize_shell(shell)
    target = []
    for i_atom, element in enumerate(molecule.elements):
        for shell in copy.deepcopy(used[element]):
            shell.i_atom = i_atom
            target.append(shell)
    if if_auxiliary_basis:
        assert not molecule.auxiliary_basis_shells
        molecule.auxiliary_basis_shells = target
    else:
        assert not molecule.basis_shells
        molecule.basis_shells = target
    assign_ao_index_to_shell(molecule)


def attach_auxiliary_basis_to_molecule(molecule: Molecule, basis_set: BasisSet) -> None:
    attach_basis_to_molecule(molecule, basis_set, True)
