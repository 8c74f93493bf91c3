"""Plain data types shared by the host layer and the integral entry points.

Field names, meanings and dtypes follow the reference's dataclasses
(electron_integral_playground/data_structure.py:6-47) so that arrays built here can be handed to
either library unchanged.  The reference spells its container aliases with the Python 3.12 ``type``
statement; this module targets Python 3.10 and uses ordinary aliases instead.
"""

from dataclasses import dataclass
from typing import Dict, List, Optional, Tuple

import numpy as np


@dataclass
class GaussianShell:
    """One contracted shell: angular momentum, AO offset, owning atom, primitives.

    Reference: data_structure.py:6-12.
    """
    angular: int
    i_ao_start: int
    i_atom: int
    primitive_exponents: np.ndarray
    primitive_coefficients: np.ndarray


BasisSet = Dict[str, List[GaussianShell]]


@dataclass
class Molecule:
    """Geometry in Bohr plus the orbital and auxiliary shells attached to it.

    Reference: data_structure.py:16-24.
    """
    elements: List[str]
    geometry: np.ndarray
    basis_shells: Optional[List[GaussianShell]]
    n_ao: int
    auxiliary_basis_shells: Optional[List[GaussianShell]]
    n_aux: int
    spherical_basis: bool


@dataclass
class PrimitivePair:
    """One screened primitive pair.  Reference: data_structure.py:26-32."""
    i_shell: int
    j_shell: int
    i_primitive: int
    j_primitive: int
    upper_bound: float


PrimitivePairList = Dict[Tuple[int, int], List[PrimitivePair]]


@dataclass
class PrimitivePairData:
    """Structure-of-arrays packing of the primitive pairs of one angular class.

    ``P_p``, ``A_a``, ``B_b`` are C-contiguous float64 (n, 4) rows ``[x, y, z, exponent]``;
    ``coefficient`` is float64 (n,); the AO offsets and atom indices are int32 (n,).
    Auxiliary "pair" data leaves ``P_p``, ``B_b``, ``j_ao_start`` and ``j_atom`` as ``None``.
    Reference: data_structure.py:36-45, pair_utility.py:142-151, 199-208.
    """
    P_p: Optional[np.ndarray]
    A_a: np.ndarray
    B_b: Optional[np.ndarray]
    coefficient: np.ndarray
    i_ao_start: np.ndarray
    j_ao_start: Optional[np.ndarray]
    i_atom: np.ndarray
    j_atom: Optional[np.ndarray]


PrimitivePairDataAngularList = Dict[object, PrimitivePairData]
