"""The oracle's arithmetic contract and ran1 against the reference's own gfortran goldens.

tests/golden/compat_expected.txt and tests/golden/ran1_reference.txt are the data files
of the reference's tests (tests/fortran/data/compat_expected.txt,
tests/library/data/ran1_reference.txt), produced by gfortran on glibc 2.43.
"""

import ctypes as C
import pathlib

import numpy as np
import pytest

from oracle import pyoracle as po

GOLD = pathlib.Path(__file__).parent / "golden"
F32 = np.float32

# the 24 inputs of tests/fortran/test_compat.py:34-36
V = [1.0, -0.123456789, 0.0, 1.82, -2.44, 0.1, 0.0999999, 1.0e8, 1.0e9, -0.000123,
     -40.6027222, 99999999.0, -0.0005, 12.3456, 3.3333333e-5, 1.0e-30, 7.5e35, -1.5,
     0.5, 2.5, -2.5, 0.49999997, 123456.789, -9.87654e-12]


def _bits(x):
    return int(np.array([x], dtype=F32).view(np.int32)[0])


def _math_rows():
    lines = (GOLD / "compat_expected.txt").read_text().splitlines()
    k = lines.index("#math")
    return [list(map(int, l.split())) for l in lines[k + 1:k + 1 + len(V)]]


@pytest.mark.parametrize("flavour", ["glibc", "cr"])
def test_glibc_independent_ops_bit_exact(flavour):
    L = po.lib(flavour)
    for i, (x, row) in enumerate(zip(V, _math_rows()), 1):
        x = F32(x)
        y = F32(abs(x) + F32(0.25))
        assert _bits(L.ora_sqrtf(y)) == row[5], (x, "sqrtf")
        assert _bits(L.ora_powi(y, i - 12)) == row[7], (x, "powi")
        assert L.ora_nint(x) == row[8], (x, "nint")
        assert L.ora_ifix(x) == row[9], (x, "ifix")


@pytest.mark.parametrize("flavour", ["glibc", "cr"])
def test_transcendentals_within_one_ulp_of_glibc243(flavour):
    """expf/logf/sinf/cosf are pinned only on the glibc that made the golden; here they
    must agree to one ulp (the values the path uses are the same up to rounding)."""
    L = po.lib(flavour)
    for x, row in zip(V, _math_rows()):
        x = F32(x)
        y = F32(abs(x) + F32(0.25))
        got = [L.ora_expf(F32(x / F32(100.0))), L.ora_logf(y), L.ora_sinf(x), L.ora_cosf(x)]
        for g, want in zip(got, row[:4]):
            assert abs(_bits(g) - want) <= 1, (x, g, want)


def test_nint_rounding_traps():
    L = po.lib()
    assert L.ora_nint(F32(0.49999997)) == 0
    assert L.ora_nint(F32(0.5)) == 1 and L.ora_nint(F32(-0.5)) == -1
    assert L.ora_nint(F32(2.5)) == 3 and L.ora_nint(F32(-2.5)) == -3
    assert L.ora_ifix(F32(-2.5)) == -2 and L.ora_ifix(F32(float("nan"))) == -2147483648


def _ran1_blocks():
    lines = (GOLD / "ran1_reference.txt").read_text().splitlines()
    blocks, i = [], 0
    while i < len(lines):
        blocks.append((lines[i], [int(v) for v in lines[i + 1:i + 1001]], lines[i + 1001]))
        i += 1002
    return blocks


def test_ran1_matches_gfortran_golden():
    """tests/library/test_ran1.py:53-80, restated: 6 seeds x 1000 deviates, the saved
    state after each block, and a continued block."""
    L = po.lib()
    blocks = _ran1_blocks()
    assert len(blocks) == 7
    s = None
    for head, vals, state in blocks[:-1]:
        s = po.Ran1.seeded(int(head.split()[1]))
        out = np.empty(1000, dtype=F32)
        L.ora_ran1_stream(C.byref(s), 1000, po.Pf(out))
        assert out.view(np.int32).tolist() == vals, head
        assert [int(v) for v in state.split()] == s.as_array().tolist(), head
    head, vals, state = blocks[-1]
    assert head == "continued"
    s2 = po.Ran1.from_array(s.as_array())   # restore from the saved state
    out = np.empty(1000, dtype=F32)
    L.ora_ran1_stream(C.byref(s2), 1000, po.Pf(out))
    assert out.view(np.int32).tolist() == vals
    assert [int(v) for v in state.split()] == s2.as_array().tolist()


def test_minimage_closed_form():
    L = po.lib()
    rng = np.random.default_rng(5)
    for v in rng.uniform(-1.6, 1.6, (200, 3)).astype(F32):
        x, y, z = (C.c_float(float(t)) for t in v)
        L.ora_minimage(C.byref(x), C.byref(y), C.byref(z))
        for got, t in zip((x.value, y.value, z.value), v):
            want = F32(t - F32(np.sign(t) * np.floor(abs(float(t)) + 0.5)))
            assert F32(got) == want
            assert -0.5 <= got <= 0.5
