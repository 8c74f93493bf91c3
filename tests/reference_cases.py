"""The reference's own integral test cases, rebuilt from tests/golden/reference_cases.{json,npz}."""

import json
import pathlib

import numpy as np

from electron_integral_playground_amd import pair_utility
from electron_integral_playground_amd.basis_utility import attach_auxiliary_basis_to_molecule, attach_basis_to_molecule
from electron_integral_playground_amd.readers import (read_basis_content, read_xyz_content,
                                                      read_xyz_content_permissive)
from electron_integral_playground_amd.units import LengthUnits

GOLDEN = pathlib.Path(__file__).resolve().parent / "golden"
CASES = json.loads((GOLDEN / "reference_cases.json").read_text())


def expected(name):
    with np.load(GOLDEN / "reference_cases.npz", allow_pickle=False) as z:
        return z[name]


def build(case):
    reader = read_xyz_content_permissive if case["permissive"] else read_xyz_content
    mol = reader(case["xyz"], unit=LengthUnits[case["unit"]])
    mol.spherical_basis = case["spherical"]
    out = {"mol": mol}
    if "basis" in case:
        attach_basis_to_molecule(mol, read_basis_content(case["basis"]["text"]))
        out["pairs"] = pair_utility.form_pair_data(mol, 1e-14)
    if "aux_basis" in case:
        attach_auxiliary_basis_to_molecule(mol, read_basis_content(case["aux_basis"]["text"]))
        out["aux"] = pair_utility.form_auxiliary_pair_data(mol)
    if case["charges"] == "nuclei":
        out["charges"] = mol.geometry.copy()
    elif case["charges"] is not None:
        out["charges"] = np.array(case["charges"], dtype=np.float64)
    return out


def evaluate(case, b, three_center, two_center, overlap, nuclear_attraction):
    """Run the case through the four entry points given (HIP library or oracle wrappers)."""
    mol, kind, omega = b["mol"], case["kind"], case["omega"]
    if kind == "S":
        return overlap(b["pairs"], mol)
    if kind == "V":
        return nuclear_attraction(b["pairs"], mol, b["charges"], omega)
    if kind == "J2c":
        return two_center(b["aux"], mol, omega)
    return three_center(b["pairs"], b["aux"], mol, omega)
