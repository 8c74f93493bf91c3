"""Overlap matrix S (reference molecular_integral/overlap.py:9-32)."""

import ctypes

import numpy as np

from .. import library
from ._ctypes import dptr, iptr


def overlap(pair_data, molecule) -> np.ndarray:
    n_ao = molecule.n_ao
    assert n_ao > 0
    S = np.zeros((n_ao, n_ao), dtype=np.float64, order="C")
    lib = library()
    for (i_L, j_L), pd in pair_data.items():
        n_pair = pd.P_p.shape[0]
        assert n_pair > 0
        rc = lib.overlap(ctypes.c_int(i_L), ctypes.c_int(j_L), dptr(pd.A_a), dptr(pd.B_b), dptr(pd.coefficient),
                         iptr(pd.i_ao_start), iptr(pd.j_ao_start), ctypes.c_int(n_pair), dptr(S),
                         ctypes.c_int(n_ao), ctypes.c_bool(molecule.spherical_basis))
        assert rc == 0, f"overlap({i_L},{j_L}) returned {rc}"
    return S
