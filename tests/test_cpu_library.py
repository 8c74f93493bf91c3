"""The HIP library loads on a CPU-only host, exports every symbol include/molecular_integral.h declares,
and rejects unsupported classes with the reference's return code before touching the GPU."""

import ctypes
import pathlib
import re

import numpy as np

from electron_integral_playground_amd import LIBRARY_PATH, library

HEADER = pathlib.Path(__file__).resolve().parents[1] / "include" / "molecular_integral.h"


def declared_functions():
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    return re.findall(r"^\s*(?:int|double|void|long long)\s+(\w+)\s*\(", text, flags=re.M)


def test_exports_every_declared_symbol():
    assert LIBRARY_PATH.exists()
    lib = library()
    names = declared_functions()
    assert {"overlap", "nuclear_attraction", "two_center", "three_center"} <= set(names)
    for n in names:
        assert hasattr(lib, n), n


def test_unsupported_class_returns_one(capfd):
    lib = library()
    d = np.zeros(8)
    i = np.zeros(2, np.int32)
    D = ctypes.POINTER(ctypes.c_double)
    I = ctypes.POINTER(ctypes.c_int)
    dp, ip = d.ctypes.data_as(D), i.ctypes.data_as(I)
    # reference three_center.cpp:624-627: unsupported class prints and returns 1
    assert lib.three_center(2, 1, 0, dp, dp, dp, dp, ip, ip, 1, dp, dp, ip, 1, dp, 1, 1, ctypes.c_bool(True),
                            ctypes.c_double(0.0)) == 1
    assert lib.three_center(0, 0, 10, dp, dp, dp, dp, ip, ip, 1, dp, dp, ip, 1, dp, 1, 1, ctypes.c_bool(True),
                            ctypes.c_double(0.0)) == 1
    assert lib.three_center(0, 7, 0, dp, dp, dp, dp, ip, ip, 1, dp, dp, ip, 1, dp, 1, 1, ctypes.c_bool(True),
                            ctypes.c_double(0.0)) == 1
    assert lib.two_center(0, 10, dp, dp, ip, 1, dp, dp, ip, 1, dp, 1, ctypes.c_bool(True), ctypes.c_double(0.0)) == 1
    assert lib.overlap(3, 2, dp, dp, dp, ip, ip, 1, dp, 1, ctypes.c_bool(True)) == 1
    assert lib.nuclear_attraction(0, 7, dp, dp, dp, dp, ip, ip, 1, dp, 1, dp, 1, ctypes.c_bool(True),
                                  ctypes.c_double(0.0)) == 1
    out = capfd.readouterr().out
    assert "does not support angular" in out


def test_device_entry_points_report_specialised_classes():
    lib = library()
    assert lib.eip_pair_matrix_nnz(2, 2) > 0 and lib.eip_pair_matrix_nnz(4, 4) == -1
    assert lib.eip_three_center_nchunks(2, 2, 3) >= 1 and lib.eip_three_center_nchunks(0, 0, 5) == -1
