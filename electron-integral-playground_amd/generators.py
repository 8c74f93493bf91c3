"""Seeded synthetic molecules and basis sets for the five benchmark workloads.

The reference ships no generators; these follow the workload descriptions:

1. ``small``      12 atoms uniform in a 6 A cube, min separation 1.0 A, Z in {1,6,7,8}; orbital shells per
                  element 1 contracted s (3 primitives) + 1 s + 1 p (2 primitives), exponents log-uniform in
                  [0.1, 50], coefficients N(0,1) then normalised; auxiliary s, p, d with 6 even-tempered
                  exponents each, ratio 2.5, starting at 0.2.
2. ``plain``      80 atoms in a 14 A box; orbital 3s2p1d, 1-6 primitives per shell, exponents log-uniform in
                  [0.05, 5e3]; auxiliary s-f, 8 even-tempered exponents per l, ratio 2.2, start 0.2
                  (n_aux = 80 * 8 * (1+3+5+7) = 10240).
3. ``batched``    256 molecules, atom counts lognormal(median 20, sigma 0.5) clipped to [4, 120], plain
                  generators with box 14 A * (n/80)^(1/3).
4. ``mixed``      200 atoms in a 20 A box; orbital 4s3p2d1f; auxiliary s-g, 8 exponents per l.
5. ``serving``    per-job atom counts lognormal(median 40, sigma 0.5) clipped to [10, 250], mixed generators
                  with box 20 A * (n/200)^(1/3).

Seeds: geometry ``20261203 + workload``, basis ``20261203 + workload + 1000`` (per-molecule offsets for
the multi-molecule workloads).  Bases are drawn once per element, so atoms of one element share shells,
as in a real basis set.  All outputs are spherical.
"""

import numpy as np

from .basis_utility import attach_auxiliary_basis_to_molecule, attach_basis_to_molecule
from .data_structure import GaussianShell, Molecule
from .units import angstrom_to_bohr

SEED_BASE = 20261203
ELEMENTS = {1: "H", 6: "C", 7: "N", 8: "O"}


def random_geometry(rng, n_atom, box_angstrom, min_sep_angstrom=1.0, max_tries=200000):
    pts = []
    tries = 0
    while len(pts) < n_atom:
        tries += 1
        if tries > max_tries:
            raise RuntimeError("could not place atoms with the requested separation")
        x = rng.uniform(0.0, box_angstrom, 3)
        if all(np.linalg.norm(x - p) >= min_sep_angstrom for p in pts):
            pts.append(x)
    return np.array(pts) * angstrom_to_bohr


def _contracted(rng, l, n_prim, lo, hi):
    exps = np.sort(np.exp(rng.uniform(np.log(lo), np.log(hi), n_prim)))[::-1].copy()
    return GaussianShell(l, -1, -1, exps, rng.normal(size=n_prim))


def _even_tempered(l_values, n_exp, ratio, start):
    shells = []
    for l in l_values:
        for k in range(n_exp):
            shells.append(GaussianShell(l, -1, -1, np.array([start * ratio ** k]), np.array([1.0])))
    return shells


def orbital_basis(rng, spec, prim_range, exp_range):
    """spec: list of (l, count) or (l, count, fixed_n_prim)."""
    shells = []
    for item in spec:
        l, count = item[0], item[1]
        for _ in range(count):
            n = item[2] if len(item) > 2 else int(rng.integers(prim_range[0], prim_range[1] + 1))
            shells.append(_contracted(rng, l, n, *exp_range))
    return shells


WORKLOADS = {
    "small": dict(box=6.0, spec=[(0, 1, 3), (0, 1, 1), (1, 1, 2)], prims=(1, 1), exps=(0.1, 50.0),
                  aux_l=(0, 1, 2), aux_n=6, aux_ratio=2.5, aux_start=0.2, id=1),
    "plain": dict(box=14.0, spec=[(0, 3), (1, 2), (2, 1)], prims=(1, 6), exps=(0.05, 5e3),
                  aux_l=(0, 1, 2, 3), aux_n=8, aux_ratio=2.2, aux_start=0.2, id=2),
    "mixed": dict(box=20.0, spec=[(0, 4), (1, 3), (2, 2), (3, 1)], prims=(1, 6), exps=(0.05, 5e3),
                  aux_l=(0, 1, 2, 3, 4), aux_n=8, aux_ratio=2.2, aux_start=0.2, id=4),
}


def make_molecule(workload, n_atom=None, seed_offset=0, box=None):
    """Molecule of one workload with attached orbital and auxiliary bases (spherical)."""
    w = WORKLOADS[workload]
    n_atom = n_atom if n_atom is not None else {"small": 12, "plain": 80, "mixed": 200}[workload]
    box = box if box is not None else w["box"]
    geo_rng = np.random.default_rng(SEED_BASE + w["id"] + seed_offset)
    bas_rng = np.random.default_rng(SEED_BASE + w["id"] + 1000 + seed_offset)
    geometry = random_geometry(geo_rng, n_atom, box)
    z = geo_rng.choice(list(ELEMENTS), size=n_atom)
    elements = [ELEMENTS[int(k)] for k in z]
    basis = {ELEMENTS[k]: orbital_basis(bas_rng, w["spec"], w["prims"], w["exps"]) for k in sorted(ELEMENTS)}
    aux = {ELEMENTS[k]: _even_tempered(w["aux_l"], w["aux_n"], w["aux_ratio"], w["aux_start"])
           for k in sorted(ELEMENTS)}
    mol = Molecule(elements=elements, geometry=geometry, basis_shells=None, n_ao=0,
                   auxiliary_basis_shells=None, n_aux=0, spherical_basis=True)
    attach_basis_to_molecule(mol, basis)
    attach_auxiliary_basis_to_molecule(mol, aux)
    return mol


def batched_sizes(n_mol=256, median=20, sigma=0.5, lo=4, hi=120, seed=SEED_BASE + 3):
    rng = np.random.default_rng(seed)
    return np.clip(np.round(rng.lognormal(np.log(median), sigma, n_mol)), lo, hi).astype(int)


def batched_molecule(index, n_atom):
    """Workload 3: plain generators, box scaled with the atom count (orbital l <= 2, aux l <= 3)."""
    return make_molecule("plain", n_atom=n_atom, seed_offset=10000 + index,
                         box=14.0 * (n_atom / 80.0) ** (1.0 / 3.0))


def serving_jobs(n_jobs=2000, rate=20.0, median=40, sigma=0.5, lo=10, hi=250, seed=SEED_BASE + 5):
    """Workload 5: (arrival time, atom count, seed offset) of a Poisson stream."""
    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, n_jobs))
    sizes = np.clip(np.round(rng.lognormal(np.log(median), sigma, n_jobs)), lo, hi).astype(int)
    return [(float(t), int(n), 20000 + i) for i, (t, n) in enumerate(zip(arrivals, sizes))]


def serving_molecule(n_atom, seed_offset):
    return make_molecule("mixed", n_atom=n_atom, seed_offset=seed_offset,
                         box=20.0 * (n_atom / 200.0) ** (1.0 / 3.0))
