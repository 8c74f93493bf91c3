"""XYZ geometry and Gaussian-format basis-set readers.

Behaviour follows the reference readers so that molecules built here carry identical numbers:
- ``read_xyz_content``: reference geometry_reader/xyz_reader.py:9-42 (first line = atom count,
  second line = comment, then ``element x y z``; Angstrom input is scaled by ``angstrom_to_bohr``).
- ``read_basis_content``: reference basis_reader/gaussian_basis_reader.py:11-94 (``****``-terminated
  element blocks, ``L nprim 1.00`` headers, Fortran ``D`` exponents, ``SP`` split into an s and a
  p shell sharing exponents, ``!`` comments, ECP rejected).
"""

import pathlib
from typing import List, Union

import numpy as np

from .data_structure import BasisSet, GaussianShell, Molecule
from .units import LengthUnits, angstrom_to_bohr

# Spectroscopic letters, including the J-less convention (I, K, L, M) of the reference
# (basis_utility.py:11).
_L_OF_LETTER = {"S": 0, "P": 1, "D": 2, "F": 3, "G": 4, "H": 5, "I": 6, "K": 7, "L": 8, "M": 9}


def angular_letter_to_l_value(letter: str) -> int:
    if len(letter) != 1 or letter.upper() not in _L_OF_LETTER:
        raise ValueError(f"unknown angular momentum letter {letter!r}")
    return _L_OF_LETTER[letter.upper()]


def _lines(content: Union[str, List[str]]) -> List[str]:
    return content.split("\n") if isinstance(content, str) else list(content)


def read_xyz_content(xyz_file_lines: Union[str, List[str]],
                     unit: LengthUnits = LengthUnits.ANGSTROM) -> Molecule:
    lines = _lines(xyz_file_lines)
    assert len(lines) >= 2
    n_atom = int(lines[0])
    assert n_atom > 0 and len(lines) >= n_atom + 2
    elements = []
    geometry = np.empty((n_atom, 3))
    for i_atom in range(n_atom):
        fields = lines[i_atom + 2].split()
        assert len(fields) == 4, f"bad xyz line {lines[i_atom + 2]!r}"
        elements.append(fields[0])
        geometry[i_atom, :] = [float(fields[1]), float(fields[2]), float(fields[3])]
    if unit == LengthUnits.ANGSTROM:
        geometry *= angstrom_to_bohr
    elif unit != LengthUnits.BOHR:
        raise NotImplementedError(f"Unsupported unit {unit}")
    return Molecule(elements=elements, geometry=geometry, basis_shells=None, n_ao=0,
                    auxiliary_basis_shells=None, n_aux=0, spherical_basis=True)


def read_xyz_file(path, unit: LengthUnits = LengthUnits.ANGSTROM) -> Molecule:
    with open(path) as f:
        return read_xyz_content(f.readlines(), unit)


def _fortran_float(text: str) -> float:
    return float(text.replace("D", "E").replace("d", "e"))


def read_basis_content(basis_file_lines: Union[str, List[str]]) -> BasisSet:
    basis_set: BasisSet = {}
    element = None
    shells = None
    remaining = 0
    letter = None
    rows: list = []
    for line in _lines(basis_file_lines):
        if "ECP" in line:
            raise NotImplementedError("ECP not supported")
        fields = line.split()
        if not fields or fields[0].startswith("!"):
            continue
        if fields[0].startswith("****"):
            assert element is not None and shells is not None
            basis_set[element] = shells
            element, shells = None, None
            continue
        if shells is None:
            assert len(fields) == 2 and fields[1] == "0"
            element = fields[0]
            assert element not in basis_set
            shells = []
            continue
        if remaining <= 0:
            assert len(fields) == 3 and float(fields[2]) == 1.0
            letter = fields[0]
            assert len(letter) == 1 or letter == "SP"
            remaining = int(fields[1])
            assert remaining > 0
            rows = []
            continue
        if letter == "SP":
            assert len(fields) == 3
        else:
            assert len(fields) == 2
        row = [_fortran_float(x) for x in fields]
        assert row[0] > 0.0
        rows.append(row)
        remaining -= 1
        if remaining == 0:
            table = np.array(rows)
            if letter == "SP":
                shells.append(GaussianShell(0, -1, -1, table[:, 0], table[:, 1]))
                shells.append(GaussianShell(1, -1, -1, table[:, 0], table[:, 2]))
            else:
                shells.append(GaussianShell(angular_letter_to_l_value(letter), -1, -1,
                                            table[:, 0], table[:, 1]))
    return basis_set


def read_basis_file(path) -> BasisSet:
    with open(path) as f:
        return read_basis_content(f.readlines())


def read_basis(name: str, directory=None) -> BasisSet:
    """Read ``<name>.gaussian.basis`` from ``directory`` (default: the package's ``data/``)."""
    directory = pathlib.Path(directory) if directory is not None else \
        pathlib.Path(__file__).parent / "data"
    path = directory / (name.lower() + ".gaussian.basis")
    if not path.is_file():
        raise FileNotFoundError(f"basis {name} not found in {directory}")
    return read_basis_file(path)


def read_xyz_content_permissive(xyz_lines: Union[str, List[str]],
                                unit: LengthUnits = LengthUnits.ANGSTROM) -> Molecule:
    """Every non-empty line is ``element x y z``; no count or comment line
    (reference geometry_reader/permissive_xyz_reader.py:9-41)."""
    elements, coords = [], []
    for line in _lines(xyz_lines):
        fields = line.split()
        if not fields:
            continue
        assert len(fields) == 4, f"bad xyz line {line!r}"
        elements.append(fields[0])
        coords.append([float(fields[1]), float(fields[2]), float(fields[3])])
    assert elements
    geometry = np.array(coords)
    if unit == LengthUnits.ANGSTROM:
        geometry *= angstrom_to_bohr
    elif unit != LengthUnits.BOHR:
        raise NotImplementedError(f"Unsupported unit {unit}")
    return Molecule(elements=elements, geometry=geometry, basis_shells=None, n_ao=0,
                    auxiliary_basis_shells=None, n_aux=0, spherical_basis=True)
