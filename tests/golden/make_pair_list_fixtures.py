"""Write pair_lists.npz: the reference's own primitive-pair lists on seeded workload molecules.

Run in a container that has /root/reference (never at test time, and never on the GPU box):

    python tests/golden/make_pair_list_fixtures.py

The reference's form_primitive_pair_list (electron_integral_playground/pair_utility.py:6-133) is run
as the reference wrote it.  Its modules use Python 3.12 ``type`` alias statements
(data_structure.py:14, 34, 47) that this interpreter (3.10) cannot parse; the modules are therefore
executed from their source text in memory with each ``type NAME = ...`` line replaced by ``NAME =
None`` (the aliases only appear in annotations).  No reference file is copied or modified.

Inputs are this package's seeded generators (the molecules, with normalised shells and AO offsets, are
rebuilt by the test from the same seeds); the fixture holds only the outputs, per angular class and in
the reference's order: i_shell, j_shell, i_primitive, j_primitive (int32) and upper_bound (float64).
These pin what the reference's own tests leave open -- the order among equal bounds and membership
when a bound lies within rounding of the threshold.

Molecules: the small workload (12 atoms), the plain workload (80 atoms) and the first 30 atoms of the
mixed workload's generator (a 30-atom molecule in a box scaled to the same density).
"""

import pathlib
import re
import sys
import time
import types

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
ROOT = HERE.parents[1]
REF = pathlib.Path("/root/reference/electron_integral_playground")
OUT = HERE / "pair_lists.npz"

THRESHOLD = 1e-14
# (fixture key, workload, n_atom, box in Angstrom or None)
MOLECULES = [("small", "small", None, None), ("plain", "plain", None, None), ("mixed30", "mixed", 30, 20.0 * (30 / 200) ** (1 / 3))]


def _load_reference_module(name, path, package):
    text = path.read_text()
    text = re.sub(r"^type\s+(\w+)\s*=.*$", r"\1 = None", text, flags=re.M)
    mod = types.ModuleType(name)
    mod.__file__ = str(path)
    sys.modules[name] = mod
    exec(compile(text, str(path), "exec"), mod.__dict__)
    setattr(package, name.rsplit(".", 1)[1], mod)
    return mod


def reference_pair_list_fn():
    pkg = types.ModuleType("electron_integral_playground")
    pkg.__path__ = [str(REF)]
    sys.modules["electron_integral_playground"] = pkg
    ds = _load_reference_module("electron_integral_playground.data_structure", REF / "data_structure.py", pkg)
    pu = _load_reference_module("electron_integral_playground.pair_utility", REF / "pair_utility.py", pkg)
    return ds, pu.form_primitive_pair_list


def workload_molecule(workload, n_atom, box):
    sys.path.insert(0, str(ROOT))
    import __graft_entry__
    __graft_entry__.load_package()
    from electron_integral_playground_amd import generators
    return generators.make_molecule(workload, n_atom=n_atom, box=box)


def main():
    ds, form_list = reference_pair_list_fn()
    arrays = {}
    for key, workload, n_atom, box in MOLECULES:
        mol = workload_molecule(workload, n_atom, box)
        shells = [ds.GaussianShell(s.angular, int(s.i_ao_start), int(s.i_atom), np.array(s.primitive_exponents),
                                   np.array(s.primitive_coefficients)) for s in mol.basis_shells]
        rmol = ds.Molecule(elements=list(mol.elements), geometry=np.array(mol.geometry), basis_shells=shells,
                           n_ao=mol.n_ao, auxiliary_basis_shells=None, n_aux=0, spherical_basis=True)
        t0 = time.perf_counter()
        pl = form_list(rmol, THRESHOLD)
        n = sum(len(v) for v in pl.values())
        print(f"{key}: {len(mol.elements)} atoms, {n} pairs in {len(pl)} classes, {time.perf_counter() - t0:.1f} s",
              flush=True)
        arrays[f"{key}/classes"] = np.array([list(k) for k in pl], dtype=np.int32)
        for (li, lj), pairs in pl.items():
            p = f"{key}/{li}{lj}/"
            arrays[p + "i_shell"] = np.array([x.i_shell for x in pairs], dtype=np.int32)
            arrays[p + "j_shell"] = np.array([x.j_shell for x in pairs], dtype=np.int32)
            arrays[p + "i_primitive"] = np.array([x.i_primitive for x in pairs], dtype=np.int32)
            arrays[p + "j_primitive"] = np.array([x.j_primitive for x in pairs], dtype=np.int32)
            arrays[p + "upper_bound"] = np.array([x.upper_bound for x in pairs], dtype=np.float64)
    np.savez_compressed(OUT, **arrays)
    print(f"wrote {OUT} ({OUT.stat().st_size / 1e6:.1f} MB)")


if __name__ == "__main__":
    main()
