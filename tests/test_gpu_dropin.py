"""Drop-in C ABI on the GPU against the oracle: every class family, spherical and Cartesian, omega.

The HIP library is called exactly as the reference's Python entry points call
libmolecular_integral.so; the oracle runs the same loops on the CPU.
"""

import numpy as np
import pytest

from tests.helpers import (ATOL, RTOL, oracle_lib, random_molecule, run_nuclear_oracle, run_overlap_oracle,
                           run_three_center_oracle, run_two_center_oracle)
from electron_integral_playground_amd import pair_utility
from electron_integral_playground_amd.molecular_integral import (nuclear_attraction, overlap, three_center,
                                                                 two_center)

pytestmark = pytest.mark.gpu

CASES = [
    # (orbital l, aux l, spherical, omega, seed); seeds >= 100 use two atoms and at most two primitives
    ((0, 1), (0, 1, 2), True, 0.0, 1),
    ((0, 1, 2), (0, 1, 2, 3), True, 0.0, 2),
    ((0, 1, 2, 3), (0, 2, 4), True, 0.0, 3),
    ((0, 1, 2), (0, 1, 3), True, 0.3, 4),
    ((0, 1, 2), (0, 1, 2), False, 0.0, 5),
    ((4, 5), (5, 6), True, 0.0, 106),
    ((1, 6), (7, 9), True, 0.2, 107),
    ((3, 4), (2, 8), False, 0.0, 108),
]


def _mol(orb, aux, sph, seed):
    print(f"case orbital={orb} aux={aux} spherical={sph} seed={seed}", flush=True)
    small = seed >= 100
    return random_molecule(n_atom=2 if small else 3, seed=seed, orbital_l=orb, aux_l=aux, spherical=sph,
                           nprim=(1, 2) if small else (1, 3), aux_nprim=(1, 1) if small else (1, 2))


@pytest.mark.parametrize("orb,aux,sph,omega,seed", CASES)
def test_three_center_vs_oracle(orb, aux, sph, omega, seed):
    mol = _mol(orb, aux, sph, seed)
    pairs = pair_utility.form_pair_data(mol, 1e-14)
    auxd = pair_utility.form_auxiliary_pair_data(mol)
    got = three_center(pairs, auxd, mol, omega)
    ref = run_three_center_oracle(oracle_lib(), pairs, auxd, mol, omega)
    np.testing.assert_allclose(got, ref, rtol=RTOL, atol=ATOL)


@pytest.mark.parametrize("orb,aux,sph,omega,seed", CASES)
def test_two_center_vs_oracle(orb, aux, sph, omega, seed):
    mol = _mol(orb, aux, sph, seed)
    auxd = pair_utility.form_auxiliary_pair_data(mol)
    got = two_center(auxd, mol, omega)
    ref = run_two_center_oracle(oracle_lib(), auxd, mol, omega)
    np.testing.assert_allclose(got, ref, rtol=RTOL, atol=ATOL)


@pytest.mark.parametrize("orb,aux,sph,omega,seed", CASES)
def test_overlap_and_nuclear_vs_oracle(orb, aux, sph, omega, seed):
    mol = _mol(orb, aux, sph, seed)
    pairs = pair_utility.form_pair_data(mol, 1e-14)
    np.testing.assert_allclose(overlap(pairs, mol), run_overlap_oracle(oracle_lib(), pairs, mol),
                               rtol=RTOL, atol=ATOL)
    charges = np.vstack([mol.geometry, [[30.0, -2.0, 1.0], [0.1, 0.2, 0.3]]])
    np.testing.assert_allclose(nuclear_attraction(pairs, mol, charges, omega),
                               run_nuclear_oracle(oracle_lib(), pairs, mol, charges, omega), rtol=RTOL, atol=ATOL)


def test_accumulates_into_existing_buffer():
    """The C ABI adds into the caller's array (reference three_center.cpp:242-244)."""
    import ctypes
    from tests.helpers import dp, ip
    from electron_integral_playground_amd import library
    mol = _mol((0, 1), (0, 1), True, 11)
    pairs = pair_utility.form_pair_data(mol, 1e-14)
    auxd = pair_utility.form_auxiliary_pair_data(mol)
    base = np.random.default_rng(0).normal(size=(mol.n_aux, mol.n_ao, mol.n_ao))
    J = base.copy()
    lib = library()
    for (i_L, j_L), pd in pairs.items():
        for k_L, pk in auxd.items():
            rc = lib.three_center(i_L, j_L, k_L, dp(pd.P_p), dp(pd.A_a), dp(pd.B_b), dp(pd.coefficient),
                                  ip(pd.i_ao_start), ip(pd.j_ao_start), pd.P_p.shape[0], dp(pk.A_a),
                                  dp(pk.coefficient), ip(pk.i_ao_start), pk.A_a.shape[0], dp(J), mol.n_ao,
                                  mol.n_aux, ctypes.c_bool(True), ctypes.c_double(0.0))
            assert rc == 0
    ref = run_three_center_oracle(oracle_lib(), pairs, auxd, mol)
    np.testing.assert_allclose(J - base, ref, rtol=RTOL, atol=1e-11)
