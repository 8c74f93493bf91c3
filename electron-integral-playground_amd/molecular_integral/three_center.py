"""Three-centre Coulomb integrals (P|ij) as a dense (n_aux, n_ao, n_ao) array, optionally
erf-attenuated (reference molecular_integral/three_center.py:9-51)."""

import ctypes

import numpy as np

from .. import library
from ._ctypes import dptr, iptr


def three_center(pair_data, auxiliary_pair_data, molecule, omega: float = 0.0) -> np.ndarray:
    assert omega >= 0.0
    n_ao, n_aux = molecule.n_ao, molecule.n_aux
    assert n_ao > 0 and n_aux > 0
    J3c = np.zeros((n_aux, n_ao, n_ao), dtype=np.float64, order="C")
    lib = library()
    for (i_L, j_L), pd in pair_data.items():
        n_pair = pd.P_p.shape[0]
        assert n_pair > 0
        for k_L, pk in auxiliary_pair_data.items():
            n_k = pk.A_a.shape[0]
            assert n_k > 0
            rc = lib.three_center(ctypes.c_int(i_L), ctypes.c_int(j_L), ctypes.c_int(k_L), dptr(pd.P_p),
                                  dptr(pd.A_a), dptr(pd.B_b), dptr(pd.coefficient), iptr(pd.i_ao_start),
                                  iptr(pd.j_ao_start), ctypes.c_int(n_pair), dptr(pk.A_a), dptr(pk.coefficient),
                                  iptr(pk.i_ao_start), ctypes.c_int(n_k), dptr(J3c), ctypes.c_int(n_ao),
                                  ctypes.c_int(n_aux), ctypes.c_bool(molecule.spherical_basis),
                                  ctypes.c_double(omega))
            assert rc == 0, f"three_center({i_L},{j_L},{k_L}) returned {rc}"
    return J3c
