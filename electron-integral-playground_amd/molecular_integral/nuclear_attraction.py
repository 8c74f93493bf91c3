"""Nuclear-attraction tensor V[mu, nu, C] = int mu nu / |r - C| per point C, no charge, positive sign
(reference molecular_integral/nuclear_attraction.py:9-44)."""

import ctypes

import numpy as np

from .. import library
from ._ctypes import dptr, iptr


def nuclear_attraction(pair_data, molecule, charge_position: np.ndarray, omega: float = 0.0) -> np.ndarray:
    assert omega >= 0.0
    n_ao = molecule.n_ao
    assert n_ao > 0
    assert charge_position.ndim == 2 and charge_position.shape[1] == 3
    n_charge = charge_position.shape[0]
    assert n_charge > 0
    charge_position = np.ascontiguousarray(charge_position, dtype=np.float64)
    V = np.zeros((n_ao, n_ao, n_charge), dtype=np.float64, order="C")
    lib = library()
    for (i_L, j_L), pd in pair_data.items():
        n_pair = pd.P_p.shape[0]
        assert n_pair > 0
        rc = lib.nuclear_attraction(ctypes.c_int(i_L), ctypes.c_int(j_L), dptr(pd.P_p), dptr(pd.A_a), dptr(pd.B_b),
                                    dptr(pd.coefficient), iptr(pd.i_ao_start), iptr(pd.j_ao_start),
                                    ctypes.c_int(n_pair), dptr(charge_position), ctypes.c_int(n_charge), dptr(V),
                                    ctypes.c_int(n_ao), ctypes.c_bool(molecule.spherical_basis),
                                    ctypes.c_double(omega))
        assert rc == 0, f"nuclear_attraction({i_L},{j_L}) returned {rc}"
    return V
