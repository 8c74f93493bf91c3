"""three_center_noalloc (no device allocation, fixed scratch region) against three_center, through the
C ABI: two calls in a row, the second with a smaller input (the region keeps the first call's
contents), spherical and Cartesian classes; and an input beyond the region's size rejected with rc 5."""

import ctypes

import numpy as np
import pytest

from tests.helpers import dp, ip, random_molecule
from electron_integral_playground_amd import library, pair_utility

pytestmark = pytest.mark.gpu


def _run(fn, pairs, aux, mol):
    J = np.zeros((mol.n_aux, mol.n_ao, mol.n_ao))
    for (i_L, j_L), pd in pairs.items():
        for k_L, pk in aux.items():
            rc = fn(i_L, j_L, k_L, dp(pd.P_p), dp(pd.A_a), dp(pd.B_b), dp(pd.coefficient), ip(pd.i_ao_start),
                    ip(pd.j_ao_start), pd.P_p.shape[0], dp(pk.A_a), dp(pk.coefficient), ip(pk.i_ao_start),
                    pk.A_a.shape[0], dp(J), mol.n_ao, mol.n_aux, ctypes.c_bool(mol.spherical_basis),
                    ctypes.c_double(0.0))
            assert rc == 0, (i_L, j_L, k_L, rc)
    return J


@pytest.mark.parametrize("spherical", [True, False])
def test_noalloc_equals_three_center_on_two_calls(spherical):
    lib = library()
    lib.three_center_noalloc_capacity.restype = ctypes.c_longlong
    print(f"case spherical={spherical} scratch={lib.three_center_noalloc_capacity()} bytes", flush=True)
    big = random_molecule(n_atom=4, seed=71, orbital_l=(0, 1, 2, 3), aux_l=(0, 1, 2, 3, 4), spherical=spherical)
    small = random_molecule(n_atom=2, seed=72, orbital_l=(0, 1), aux_l=(0, 2), spherical=spherical)
    for mol in (big, small):
        pairs = pair_utility.form_pair_data(mol, 1e-14)
        aux = pair_utility.form_auxiliary_pair_data(mol)
        ref = _run(lib.three_center, pairs, aux, mol)
        got = _run(lib.three_center_noalloc, pairs, aux, mol)
        assert np.array_equal(got, ref), float(np.max(np.abs(got - ref)))
        assert np.abs(ref).max() > 0


def test_noalloc_rejects_inputs_beyond_its_region():
    """3.2 million (s|s) primitive pairs need more than the fixed region: rc 5 before any kernel runs."""
    lib = library()
    lib.three_center_noalloc_capacity.restype = ctypes.c_longlong
    cap = lib.three_center_noalloc_capacity()
    n = int(cap // 80) + 1024
    print(f"case oversized n_pair={n} capacity={cap}", flush=True)
    rng = np.random.default_rng(0)
    P = np.ascontiguousarray(np.c_[rng.normal(size=(n, 3)), np.full(n, 2.0)])
    A = P.copy()
    B = P.copy()
    c = np.ones(n)
    st = np.zeros(n, dtype=np.int32)
    K = np.array([[0.0, 0.0, 0.0, 1.0]])
    kc = np.ones(1)
    ks = np.zeros(1, dtype=np.int32)
    J = np.zeros((1, 1, 1))
    rc = lib.three_center_noalloc(0, 0, 0, dp(P), dp(A), dp(B), dp(c), ip(st), ip(st), n, dp(K), dp(kc), ip(ks), 1,
                                  dp(J), 1, 1, ctypes.c_bool(True), ctypes.c_double(0.0))
    assert rc == 5
    assert J[0, 0, 0] == 0.0
