"""Fortran-shaped numpy arrays for the COMMON mirror: one extra element per dimension
(index 0 unused) and column-major order, so a[i, j] addresses the reference's a(i,j)
and a[:, j] is a contiguous column."""

import numpy as np


def farray(dtype, *dims):
    return np.zeros(tuple(int(d) + 1 for d in dims), dtype=dtype, order="F")


def real(*dims):
    return farray(np.float32, *dims)


def integer(*dims):
    return farray(np.int32, *dims)
