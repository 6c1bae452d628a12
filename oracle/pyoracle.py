"""ctypes wrapper of the CPU oracle (oracle/build/liboracle_*.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product package never imports it.

Arrays are numpy; grid-indexed arrays use the reference's 1-based numbering with an
unused element 0 (see epsr_oracle.h).
"""

from __future__ import annotations

import ctypes as C
import os
import pathlib
import subprocess

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
BUILD = HERE / "build"
F32 = np.float32

_LIBS: dict = {}


def build():
    subprocess.run(["sh", str(HERE / "build.sh")], check=True)


def lib(flavour: str = "glibc"):
    """flavour: 'glibc' (host libm, as the reference), 'cr' (correctly rounded), 'omp'."""
    if flavour not in _LIBS:
        path = BUILD / f"liboracle_{flavour}.so"
        if not path.exists():
            build()
        L = C.CDLL(str(path))
        _declare(L)
        _LIBS[flavour] = L
    return _LIBS[flavour]


fp = C.POINTER(C.c_float)
ip = C.POINTER(C.c_int)
i64p = C.POINTER(C.c_int64)
dp = C.POINTER(C.c_double)


class Ran1(C.Structure):
    _fields_ = [("mcall", C.c_int32), ("idum", C.c_int32), ("iy", C.c_int32), ("iv", C.c_int32 * 32)]

    @classmethod
    def seeded(cls, idum):
        s = cls()
        s.mcall, s.idum, s.iy = 0, int(idum), 0
        return s

    def as_array(self):
        return np.array([self.mcall, self.idum, self.iy] + list(self.iv), dtype=np.int32)

    @classmethod
    def from_array(cls, a):
        s = cls()
        s.mcall, s.idum, s.iy = int(a[0]), int(a[1]), int(a[2])
        for i in range(32):
            s.iv[i] = int(a[3 + i])
        return s


class Box(C.Structure):
    _fields_ = [("nmol", C.c_int), ("natom", C.c_int), ("ntype", C.c_int), ("npairs", C.c_int),
                ("nmoltype", C.c_int), ("xa", fp), ("ya", fp), ("za", fp), ("itype", ip),
                ("npairref", ip), ("n1extr", ip), ("n2extr", ip), ("naf", ip), ("nal", ip),
                ("mtype", ip), ("cfrac", fp), ("natomr", C.c_float), ("sizefactor", C.c_float),
                ("dmax", C.c_float), ("cellst", C.c_float)]


class Centres(C.Structure):
    _fields_ = [("xadd", fp), ("yadd", fp), ("zadd", fp)]


class Tables(C.Structure):
    _fields_ = [("ld", C.c_int), ("npt2", C.c_int), ("xstar", fp), ("xavs", fp), ("xavs1", fp),
                ("ecoul", fp), ("ecoul1", fp)]


class Cells(C.Structure):
    _fields_ = [("ncelldiva", C.c_int), ("ncelldivb", C.c_int), ("ncelldivc", C.c_int),
                ("ncell", C.c_int), ("ncelldivsq", C.c_int), ("maxcellno", C.c_int),
                ("ncellneighb", C.c_int), ("nmolsincell", ip), ("molsincell", ip),
                ("molcellno", ip), ("iacellneighb", ip), ("ibcellneighb", ip), ("iccellneighb", ip)]


class Ledger(C.Structure):
    _fields_ = [("nmol", C.c_int), ("nnbrefmax", C.c_int), ("nergpair", C.c_int),
                ("nergpairmax", C.c_int), ("ergpair", fp), ("molergpair", ip), ("ergpairref", ip),
                ("nergpairm", ip)]


class MCState(C.Structure):
    _fields_ = [("wholerot", fp), ("wholetrans", fp), ("nwhole", ip), ("nmovem", ip),
                ("nstepreject", ip), ("simt", C.c_float), ("stepr", C.c_float),
                ("stepm", C.c_float), ("rejrate", C.c_float), ("ireset", C.c_int)]


class Intra(C.Structure):
    _fields_ = [("qtype", ip), ("imageatom", ip), ("mref", ip), ("nbond", ip), ("iabond", ip), ("tbond", fp),
                ("ldb", C.c_int), ("amassr", fp), ("amassa", fp), ("ngrot", ip), ("nrf", ip), ("nrl", ip),
                ("labaxi", ip), ("rotf", ip), ("rotl", ip), ("labrot", ip), ("isdih", ip), ("ndihedralang", ip),
                ("labrotdih", ip), ("dihedralang", fp), ("ldd", C.c_int), ("xav", fp), ("xav1", fp),
                ("dmaxmol", C.c_float), ("tempb", C.c_float), ("simt", C.c_float), ("tempang", C.c_float),
                ("tempdih", C.c_float), ("refine_intra_fac", C.c_float), ("refine_intra", C.c_int)]


class MCFull(C.Structure):
    _fields_ = [("wholerot", fp), ("wholetrans", fp), ("internrot", fp), ("interntrans", fp), ("nwhole", ip),
                ("nmovem", ip), ("ninter", ip), ("nmovei", ip), ("nstepreject", ip), ("simt", C.c_float),
                ("stepr", C.c_float), ("stepm", C.c_float), ("stepri", C.c_float), ("stepmi", C.c_float),
                ("rejrate", C.c_float), ("tempangwt", C.c_float), ("tempdihwt", C.c_float), ("ireset", C.c_int)]


class DMS(C.Structure):
    _fields_ = [("nfile", C.c_int), ("nq", C.c_int), ("ldq", C.c_int), ("npairs", C.c_int), ("q", fp),
                ("datai", fp), ("soqintra", fp), ("soqinter", fp), ("ciwt", fp), ("cwt", fp), ("weightx", dp),
                ("xfile", ip), ("nqmind", ip), ("nqmaxd", ip), ("nqmin", ip), ("nqmax", ip), ("szeros", fp),
                ("porodamp", fp), ("rshmin", fp), ("porodq", C.c_float), ("porodpwr", C.c_float),
                ("nsmoop", C.c_int), ("x", fp), ("npt1", C.c_int), ("rho", C.c_float)]


class Grid(C.Structure):
    _fields_ = [("npt", C.c_int), ("npt1", C.c_int), ("npt2", C.c_int), ("nptmin", C.c_int),
                ("nptmax", C.c_int)]


def _declare(L):
    P = C.POINTER
    sig = {
        "ora_nint": (C.c_int, [C.c_float]),
        "ora_ifix": (C.c_int, [C.c_float]),
        "ora_powi": (C.c_float, [C.c_float, C.c_int]),
        "ora_sqrtf": (C.c_float, [C.c_float]),
        "ora_expf": (C.c_float, [C.c_float]),
        "ora_cosf": (C.c_float, [C.c_float]),
        "ora_sinf": (C.c_float, [C.c_float]),
        "ora_logf": (C.c_float, [C.c_float]),
        "ora_libm_is_cr": (C.c_int, []),
        "ora_ran1": (C.c_float, [P(Ran1)]),
        "ora_ran1_stream": (None, [P(Ran1), C.c_int, fp]),
        "ora_minimage": (None, [fp, fp, fp]),
        "ora_atogen": (C.c_int, [P(Box), P(Centres), C.c_int, C.c_int, fp, ip, fp, C.c_int, P(Ran1),
                                 fp, fp, fp, fp, ip]),
        "ora_atogen_counts": (C.c_int, [P(Box), P(Centres), C.c_int, C.c_int, ip, fp, C.c_int,
                                        P(Ran1), i64p, i64p]),
        "ora_gensq": (None, [C.c_int, C.c_int, C.c_int, C.c_int, fp, fp, fp, C.c_int, ip, fp, fp,
                             C.c_int]),
        "ora_lorch": (C.c_float, [C.c_float, C.c_float]),
        "ora_stog": (None, [C.c_float, fp, fp, C.c_int, fp, fp, C.c_int, ip, ip, C.c_int, C.c_int, fp]),
        "ora_en_atom_inter_scale": (C.c_float, [C.c_float] * 7 + [P(Tables), C.c_int, C.c_float]),
        "ora_en_mole_pair": (C.c_float, [P(Box), P(Centres), P(Tables), C.c_int, C.c_int]),
        "ora_cells_init": (C.c_int, [P(Cells), C.c_int, fp, fp, fp, C.c_int, C.c_float]),
        "ora_cells_free": (None, [P(Cells)]),
        "ora_make_neighbour_list_from_cells": (C.c_int, [P(Cells), C.c_int, ip]),
        "ora_get_nnbrefmax": (C.c_int, [P(Cells)]),
        "ora_ledger_free": (None, [P(Ledger)]),
        "ora_fmolec1": (C.c_int, [P(Box), P(Centres), P(Tables), P(Cells), fp, P(Ledger), fp, fp, dp]),
        "ora_sumsq": (None, [C.c_int, C.c_int, C.c_int, fp, C.c_double, C.c_double, fp, fp]),
        "ora_sumgr": (None, [C.c_int, C.c_int, C.c_int, fp, fp, C.c_double, C.c_double, fp, fp, fp]),
        "ora_invert_cwt": (C.c_int, [dp, dp, dp, dp, C.c_float, fp, C.c_int, C.c_int, P(Ran1)]),
        "ora_press": (None, [P(Box), P(Centres), P(Tables), P(Ledger), C.c_int, C.c_float, C.c_float, fp, fp, dp,
                             fp, fp]),
        "ora_en_change": (C.c_float, [P(Box), P(Centres), P(Tables), P(Ledger), C.c_int, fp, ip]),
        "ora_molfit": (C.c_int, [P(Box), fp, fp, fp, P(Centres), P(Tables), P(Cells), P(Ledger),
                                 P(MCState), P(Ran1), fp, ip]),
        "ora_inita_grid": (None, [C.c_float, C.c_float, C.c_float, C.c_float, C.c_int, fp, P(Grid)]),
        "ora_uniform": (None, [fp, C.c_float, C.c_float, C.c_float, C.c_int]),
        "ora_j1qr": (C.c_double, [C.c_float, C.c_float]),
        "ora_pofr": (C.c_float, [C.c_float]),
        "ora_foqols": (None, [C.c_int, fp, C.c_int, fp, fp, C.c_float, C.c_float, C.c_float, C.c_float,
                              C.c_float, C.c_float, fp, C.c_int]),
        "ora_truncate": (C.c_float, [C.c_float, C.c_float, C.c_float]),
        "ora_truncatec": (C.c_float, [C.c_float, C.c_float]),
        "ora_chsph": (C.c_float, [C.c_float, C.c_float, C.c_float]),
        "ora_generate_rp": (None, [C.c_int, C.c_int, fp, fp, fp, fp, C.c_int, fp, fp, fp, fp, fp,
                                   C.c_float, C.c_float, fp, fp, C.c_float, C.c_float, fp, fp, fp, fp,
                                   C.c_int]),
        "ora_expand": (C.c_float, [C.c_int, C.c_int, fp, fp, C.c_float]),
        "ora_dnext": (C.c_float, [C.c_float, C.c_float, C.c_float, fp]),
        "ora_calc_pwrexpq": (C.c_int, [C.c_int, C.c_int, C.c_float, C.c_float, C.c_float, C.c_float,
                                       C.c_float, fp, fp, C.c_float]),
        "ora_calc_pwrexpr_grad": (C.c_int, [C.c_int, C.c_int, C.c_float, C.c_float, C.c_float,
                                            C.c_float, C.c_float, fp, fp, C.c_float, C.c_float,
                                            C.c_float]),
        "ora_generate_ep": (C.c_float, [C.c_int, C.c_int, fp, fp, fp, C.c_int, ip, C.c_int, fp, C.c_int,
                                        C.c_float, C.c_float, C.c_float, C.c_float, C.c_float,
                                        C.c_float, C.c_float, C.c_float, C.c_float, C.c_float, fp, fp,
                                        fp, fp, fp, C.c_int, fp]),
        "ora_smoop": (None, [C.c_int, C.c_int, C.c_int, fp]),
        "ora_fit_expansion": (None, [fp, fp, fp, fp, fp, fp, fp, fp, fp, C.c_int, C.c_int, C.c_int,
                                     C.c_int, C.c_int, C.c_int, P(Ran1), ip]),
        "ora_ptgnsq_diff": (None, [C.c_int, C.c_int, C.c_int, fp, fp, C.c_int, fp, fp, fp, ip, ip, C.c_float,
                                   fp]),
        "ora_ptgnsq_rfactor": (C.c_float, [C.c_int, C.c_int, fp, fp]),
        "ora_ptgnsq_diff_x": (None, [C.c_int, C.c_int, C.c_int, fp, fp, C.c_int, fp, fp, dp, ip, fp, ip, ip,
                                     C.c_float, fp]),
        "ora_ftback": (None, [C.c_int, C.c_int, fp, fp, C.c_int, C.c_float, C.c_float, fp, fp]),
        "ora_data_minus_sim": (None, [P(DMS), fp, fp, fp, fp]),
        "ora_da_uniform": (C.c_float, [C.c_ulonglong, C.c_int]),
        "ora_molfit_da": (C.c_int, [P(Box), fp, fp, fp, P(Centres), P(Tables), P(Cells), P(Ledger),
                                    P(MCState), P(Ran1), fp, ip, fp, C.c_ulonglong, ip, dp, fp]),
        "ora_en_atom_intra": (C.c_float, [P(Box), P(Tables), P(Intra), C.c_int, C.c_int, fp, fp]),
        "ora_en_dihedral": (C.c_float, [fp, fp, fp, C.c_float]),
        "ora_en_mole_dihedral": (C.c_float, [P(Box), P(Intra), C.c_int]),
        "ora_en_mole_intra": (C.c_float, [P(Box), P(Tables), P(Intra), C.c_int, fp, fp, fp]),
        "ora_molfit_full": (C.c_int, [P(Box), fp, fp, fp, P(Centres), P(Tables), P(Intra), P(Cells), P(Ledger),
                                      P(MCFull), P(Ran1), fp, fp, ip, C.c_int, C.c_int]),
        "ora_descriptor": (None, [C.c_int, fp, fp, fp, ip, C.c_int, fp, fp, fp, ip, ip, C.c_int, C.c_int,
                                  C.c_float, C.c_int, C.c_float, C.c_float, C.c_float, C.c_float, C.c_float,
                                  dp]),
        "ora_mlp_forward": (None, [C.c_int, C.c_int, fp, C.c_int, C.c_int, fp, C.c_int, dp]),
        "ora_mlp_grad": (C.c_double, [C.c_int, C.c_int, fp, C.c_int, C.c_int, fp, C.c_int, fp, dp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args


def P(a, ctype=C.c_float):
    """pointer to a contiguous numpy array"""
    assert a.flags["C_CONTIGUOUS"], "array must be contiguous"
    return a.ctypes.data_as(C.POINTER(ctype))


def Pi(a):
    assert a.dtype == np.int32
    return P(a, C.c_int)


def Pf(a):
    assert a.dtype == np.float32
    return P(a, C.c_float)


class OracleBox:
    """Holds contiguous copies of a synthetic System's arrays and the ctypes views."""

    def __init__(self, s, xa=None, ya=None, za=None, xadd=None, yadd=None, zadd=None):
        self.s = s
        c = np.ascontiguousarray
        self.xa = c(s.xa if xa is None else xa, dtype=F32).copy()
        self.ya = c(s.ya if ya is None else ya, dtype=F32).copy()
        self.za = c(s.za if za is None else za, dtype=F32).copy()
        self.xadd = c(s.xadd if xadd is None else xadd, dtype=F32).copy()
        self.yadd = c(s.yadd if yadd is None else yadd, dtype=F32).copy()
        self.zadd = c(s.zadd if zadd is None else zadd, dtype=F32).copy()
        self.itype = c(s.itype, dtype=np.int32)
        self.npairref = c(s.npairref.reshape(-1), dtype=np.int32)
        self.n1extr = c(s.n1extr, dtype=np.int32)
        self.n2extr = c(s.n2extr, dtype=np.int32)
        self.naf = c(s.naf, dtype=np.int32)
        self.nal = c(s.nal, dtype=np.int32)
        self.mtype = c(s.mtype, dtype=np.int32)
        self.cfrac = c(s.cfrac, dtype=F32)
        self.box = Box(s.nmol, s.natom, s.ntype, s.npairs, s.nmoltype, Pf(self.xa), Pf(self.ya),
                       Pf(self.za), Pi(self.itype), Pi(self.npairref), Pi(self.n1extr),
                       Pi(self.n2extr), Pi(self.naf), Pi(self.nal), Pi(self.mtype), Pf(self.cfrac),
                       F32(s.natomr), F32(1.0), F32(s.dmax), F32(s.cellst))
        self.centres = Centres(Pf(self.xadd), Pf(self.yadd), Pf(self.zadd))


def ran1_seq(idum, n, flavour="glibc"):
    s = Ran1.seeded(idum)
    out = np.empty(n, dtype=F32)
    lib(flavour).ora_ran1_stream(C.byref(s), n, Pf(out))
    return out, s
