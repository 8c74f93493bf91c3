"""Two-centre Coulomb metric (P|Q), optionally erf-attenuated
(reference molecular_integral/two_center.py:9-43)."""

import ctypes

import numpy as np

from .. import library
from ._ctypes import dptr, iptr


def two_center(pair_data, molecule, omega: float = 0.0) -> np.ndarray:
    assert omega >= 0.0
    n_aux = molecule.n_aux
    assert n_aux > 0
    J2c = np.zeros((n_aux, n_aux), dtype=np.float64, order="C")
    lib = library()
    for i_L, pi in pair_data.items():
        assert pi.A_a.shape[0] > 0
        for j_L, pj in pair_data.items():
            assert pj.A_a.shape[0] > 0
            if i_L > j_L:
                continue
            rc = lib.two_center(ctypes.c_int(i_L), ctypes.c_int(j_L), dptr(pi.A_a), dptr(pi.coefficient),
                                iptr(pi.i_ao_start), ctypes.c_int(pi.A_a.shape[0]), dptr(pj.A_a),
                                dptr(pj.coefficient), iptr(pj.i_ao_start), ctypes.c_int(pj.A_a.shape[0]),
                                dptr(J2c), ctypes.c_int(n_aux), ctypes.c_bool(molecule.spherical_basis),
                                ctypes.c_double(omega))
            assert rc == 0, f"two_center({i_L},{j_L}) returned {rc}"
    return J2c
