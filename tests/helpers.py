"""Shared test helpers: oracle loading, small random molecules, oracle-side entry points.

The oracle (oracle/build/libmd_oracle.so) is the checker; the functions here call it with exactly the
arrays and loop structure the reference's Python entry points use
(reference molecular_integral/{overlap,nuclear_attraction,two_center,three_center}.py).
"""

import ctypes
import pathlib
import sys

import numpy as np

ROOT = pathlib.Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))
import __graft_entry__  # noqa: E402

__graft_entry__.load_package()

from electron_integral_playground_amd.basis_utility import (  # noqa: E402
    attach_auxiliary_basis_to_molecule, attach_basis_to_molecule)
from electron_integral_playground_amd.data_structure import GaussianShell, Molecule  # noqa: E402

ORACLE_SO = ROOT / "oracle" / "build" / "libmd_oracle.so"
REF_SO = ROOT / "oracle" / "_ref" / "libmolecular_integral_ref.so"

D = ctypes.POINTER(ctypes.c_double)
I = ctypes.POINTER(ctypes.c_int)


def dp(a):
    return a.ctypes.data_as(D)


def ip(a):
    return a.ctypes.data_as(I)


_oracle = None


def oracle_lib():
    global _oracle
    if _oracle is None:
        _oracle = ctypes.CDLL(str(ORACLE_SO))
        _oracle.oracle_cart_to_sph_row.restype = ctypes.c_double
        _oracle.oracle_boys.argtypes = [ctypes.c_int, ctypes.c_double, D]
    return _oracle


def random_basis(rng, ls, nprim=(1, 3), exp_range=(0.1, 5.0)):
    shells = []
    for l in ls:
        n = int(rng.integers(nprim[0], nprim[1] + 1))
        exps = np.exp(rng.uniform(np.log(exp_range[0]), np.log(exp_range[1]), n))
        coefs = rng.normal(size=n)
        shells.append(GaussianShell(l, -1, -1, exps, coefs))
    return shells


def random_molecule(n_atom=3, seed=0, orbital_l=(0, 1), aux_l=(0, 1), box=3.0, spherical=True,
                    nprim=(1, 3), aux_nprim=(1, 2)):
    """Atoms in a box (Bohr) sharing one random element 'X' basis; contracted shells."""
    rng = np.random.default_rng(seed)
    geom = rng.uniform(-box / 2, box / 2, size=(n_atom, 3))
    mol = Molecule(elements=["X"] * n_atom, geometry=geom, basis_shells=None, n_ao=0,
                   auxiliary_basis_shells=None, n_aux=0, spherical_basis=spherical)
    attach_basis_to_molecule(mol, {"X": random_basis(rng, orbital_l, nprim)})
    attach_auxiliary_basis_to_molecule(mol, {"X": random_basis(rng, aux_l, aux_nprim, (0.2, 8.0))})
    return mol


def run_three_center_oracle(lib, pair_data, aux_data, mol, omega=0.0, fn="oracle_three_center"):
    J = np.zeros((mol.n_aux, mol.n_ao, mol.n_ao))
    for (i_L, j_L), pd in pair_data.items():
        for k_L, pk in aux_data.items():
            rc = getattr(lib, fn)(i_L, j_L, k_L, dp(pd.P_p), dp(pd.A_a), dp(pd.B_b), dp(pd.coefficient),
                                  ip(pd.i_ao_start), ip(pd.j_ao_start), pd.P_p.shape[0], dp(pk.A_a),
                                  dp(pk.coefficient), ip(pk.i_ao_start), pk.A_a.shape[0], dp(J), mol.n_ao,
                                  mol.n_aux, ctypes.c_bool(mol.spherical_basis), ctypes.c_double(omega))
            assert rc == 0
    return J


def run_two_center_oracle(lib, aux_data, mol, omega=0.0, fn="oracle_two_center"):
    J = np.zeros((mol.n_aux, mol.n_aux))
    for i_L, pi in aux_data.items():
        for j_L, pj in aux_data.items():
            if i_L > j_L:
                continue
            rc = getattr(lib, fn)(i_L, j_L, dp(pi.A_a), dp(pi.coefficient), ip(pi.i_ao_start), pi.A_a.shape[0],
                                  dp(pj.A_a), dp(pj.coefficient), ip(pj.i_ao_start), pj.A_a.shape[0], dp(J),
                                  mol.n_aux, ctypes.c_bool(mol.spherical_basis), ctypes.c_double(omega))
            assert rc == 0
    return J


def run_overlap_oracle(lib, pair_data, mol, fn="oracle_overlap"):
    S = np.zeros((mol.n_ao, mol.n_ao))
    for (i_L, j_L), pd in pair_data.items():
        rc = getattr(lib, fn)(i_L, j_L, dp(pd.A_a), dp(pd.B_b), dp(pd.coefficient), ip(pd.i_ao_start),
                              ip(pd.j_ao_start), pd.A_a.shape[0], dp(S), mol.n_ao, ctypes.c_bool(mol.spherical_basis))
        assert rc == 0
    return S


def run_nuclear_oracle(lib, pair_data, mol, charges, omega=0.0, fn="oracle_nuclear_attraction"):
    charges = np.ascontiguousarray(charges, dtype=np.float64)
    V = np.zeros((mol.n_ao, mol.n_ao, charges.shape[0]))
    for (i_L, j_L), pd in pair_data.items():
        rc = getattr(lib, fn)(i_L, j_L, dp(pd.P_p), dp(pd.A_a), dp(pd.B_b), dp(pd.coefficient), ip(pd.i_ao_start),
                              ip(pd.j_ao_start), pd.A_a.shape[0], dp(charges), charges.shape[0], dp(V), mol.n_ao,
                              ctypes.c_bool(mol.spherical_basis), ctypes.c_double(omega))
        assert rc == 0
    return V


# Parity bar of the floating-point integrals: |HIP - oracle| <= 1e-12 + 1e-10 |oracle|.
ATOL = 1e-12
RTOL = 1e-10


def oracle_aux_shell_block(lib, pair_data, pk, k_L, start, n_ao, omega=0.0, fn="oracle_three_center"):
    """Dense (2 k_L + 1, n_ao, n_ao) oracle (ij|P) of ONE auxiliary shell (the primitives of class k_L
    whose i_ao_start equals ``start``, rebased to function 0) against every orbital pair: the reference's
    per-class loop restricted to one auxiliary shell, so samples of workloads whose full output exceeds
    host memory can be checked column by column."""
    sel = pk.i_ao_start == start
    A = np.ascontiguousarray(pk.A_a[sel])
    c = np.ascontiguousarray(pk.coefficient[sel])
    st = np.zeros(int(sel.sum()), dtype=np.int32)
    nk = 2 * k_L + 1
    J = np.zeros((nk, n_ao, n_ao))
    for (i_L, j_L), pd in pair_data.items():
        rc = getattr(lib, fn)(i_L, j_L, k_L, dp(pd.P_p), dp(pd.A_a), dp(pd.B_b), dp(pd.coefficient),
                              ip(pd.i_ao_start), ip(pd.j_ao_start), pd.P_p.shape[0], dp(A), dp(c), ip(st),
                              A.shape[0], dp(J), n_ao, nk, ctypes.c_bool(True), ctypes.c_double(omega))
        assert rc == 0
    return J


def sample_aux_shells(aux_data, per_class, seed):
    """{k_L: sorted array of sampled auxiliary shell starts}: ``per_class`` (an int, or {k_L: int})
    distinct shells per angular class, drawn with a seeded generator."""
    rng = np.random.default_rng(seed)
    out = {}
    for k_L, pk in aux_data.items():
        starts = np.unique(pk.i_ao_start)
        n = per_class if isinstance(per_class, int) else per_class.get(k_L, 0)
        out[k_L] = np.sort(rng.choice(starts, size=min(n, len(starts)), replace=False))
    return out


def assert_parity(got, ref, what, atol=ATOL, rtol=RTOL):
    """|got - ref| <= atol + rtol |ref| elementwise, with the worst element named on failure."""
    got, ref = np.asarray(got), np.asarray(ref)
    err = np.abs(got - ref)
    bad = err > atol + rtol * np.abs(ref)
    if bad.any():
        k = int(np.argmax(np.where(bad, err, -1.0)))
        raise AssertionError(f"{what}: {int(bad.sum())} of {bad.size} elements outside {atol:g} + {rtol:g}|ref|; "
                             f"worst at flat index {k}: got {got.flat[k]!r}, oracle {ref.flat[k]!r}")
    return float(err.max()) if err.size else 0.0


ORACLE_LD_SO = ROOT / "oracle" / "build" / "libmd_oracle_ld.so"
_oracle_ld = None


def oracle_ld_lib():
    """The oracle built with long-double internal arithmetic (decides ill-conditioned elements)."""
    global _oracle_ld
    if _oracle_ld is None:
        _oracle_ld = ctypes.CDLL(str(ORACLE_LD_SO))
    return _oracle_ld


def oracle_shell_pair_block(lib, pd, rows, la, lb, pk, k_L, start, omega=0.0, fn="oracle_three_center"):
    """(2 k_L + 1, 2 la + 1, 2 lb + 1) oracle block of ONE contracted shell pair -- the primitive pairs
    ``rows`` of the class's pair data ``pd`` (all with the same AO offsets) -- against ONE auxiliary shell
    (primitives of ``pk`` starting at ``start``), on a rebased AO range so the output stays small."""
    rows = np.asarray(rows)
    I, J = int(pd.i_ao_start[rows[0]]), int(pd.j_ao_start[rows[0]])
    ni, nj = 2 * la + 1, 2 * lb + 1
    i0, j0 = 0, (0 if I == J else ni)
    n = ni if I == J else ni + nj
    take = lambda a: np.ascontiguousarray(a[rows])
    P, A, B, c = take(pd.P_p), take(pd.A_a), take(pd.B_b), take(pd.coefficient)
    si, sj = np.full(len(rows), i0, np.int32), np.full(len(rows), j0, np.int32)
    sel = pk.i_ao_start == start
    K, kc = np.ascontiguousarray(pk.A_a[sel]), np.ascontiguousarray(pk.coefficient[sel])
    ks = np.zeros(int(sel.sum()), dtype=np.int32)
    nk = 2 * k_L + 1
    out = np.zeros((nk, n, n))
    rc = getattr(lib, fn)(la, lb, k_L, dp(P), dp(A), dp(B), dp(c), ip(si), ip(sj), len(rows), dp(K), dp(kc), ip(ks),
                          K.shape[0], dp(out), n, nk, ctypes.c_bool(True), ctypes.c_double(omega))
    assert rc == 0
    return out[:, i0:i0 + ni, j0:j0 + nj]
