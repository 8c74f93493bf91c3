"""ctypes plumbing shared by the entry points: argument conversion identical to the reference's
(``ndarray.ctypes.data_as(POINTER(...))``, ``c_int``/``c_double``/``c_bool`` scalars)."""

import ctypes

import numpy as np

D = ctypes.POINTER(ctypes.c_double)
I = ctypes.POINTER(ctypes.c_int)


def dptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(D)


def iptr(a: np.ndarray):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(I)
