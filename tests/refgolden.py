"""Reader of tests/golden/ref_tinybox.npz (written by tests/golden/make_ref_vectors.py from the
reference's own routines) and builders of the oracle's state from it.

The fixture keeps the reference's 1-based layout (element 0 unused, molecule and atom
numbers 1-based); the builders convert to the oracle's 0-based conventions.
"""

from __future__ import annotations

import ctypes as C
import functools
import os

import numpy as np

from oracle import pyoracle as po

F32 = np.float32
PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ref_tinybox.npz")


@functools.lru_cache(maxsize=None)
def load():
    raw = np.load(PATH, allow_pickle=False)
    d = {k: raw[k] for k in raw.files}
    for k, v in list(d.items()):
        if v.dtype.kind == "U" and v.shape == () and str(v).startswith("="):
            d[k] = d[str(v)[1:]]
    return d


def group(prefix):
    d = load()
    n = len(prefix) + 1
    return {k[n:]: v for k, v in d.items() if k.startswith(prefix + ".")}


def count(name):
    d = load()
    return len({k.split(".")[1] for k in d if k.startswith(name + ".")})


class State:
    """Oracle-side copies of the recorded box: topology (0-based), coordinates, tables."""

    def __init__(self, coords, tables):
        t = group("topo")
        self.t = t
        self.nmol, self.natom = int(t["nmol"]), int(t["natom"])
        self.ntype, self.npairs, self.nmoltype = int(t["ntype"]), int(t["npairs"]), int(t["nmoltype"])
        c = np.ascontiguousarray
        self.xa, self.ya, self.za = (c(coords[k], F32).copy() for k in ("xa", "ya", "za"))
        self.xadd, self.yadd, self.zadd = (c(coords[k], F32).copy() for k in ("xadd", "yadd", "zadd"))
        self.itype = c(t["itype"] - 1, np.int32)
        self.npairref = c((t["npairref"] - 1).reshape(-1), np.int32)
        self.n1extr = c(t["n1extr"] - 1, np.int32)
        self.n2extr = c(t["n2extr"] - 1, np.int32)
        self.naf = c(t["naf"] - 1, np.int32)
        self.nal = c(t["nal"] - 1, np.int32)
        self.mtype = c(t["mtype"] - 1, np.int32)
        self.cfrac = c(t["cfrac"], F32)
        self.box = po.Box(self.nmol, self.natom, self.ntype, self.npairs, self.nmoltype, po.Pf(self.xa),
                          po.Pf(self.ya), po.Pf(self.za), po.Pi(self.itype), po.Pi(self.npairref),
                          po.Pi(self.n1extr), po.Pi(self.n2extr), po.Pi(self.naf), po.Pi(self.nal),
                          po.Pi(self.mtype), po.Pf(self.cfrac), F32(t["natomr"]), F32(t["sizefactor"]),
                          F32(t["dmax"]), F32(t["cellst"]))
        self.centres = po.Centres(po.Pf(self.xadd), po.Pf(self.yadd), po.Pf(self.zadd))
        self.tab = {k: c(tables[k], F32).reshape(-1).copy() for k in ("xavs", "xavs1", "ecoul", "ecoul1")}
        self.xstar = c(tables["xstar"], F32)
        self.ld = int(tables["xavs"].shape[1])
        self.npt2 = int(tables["npt2"])
        self.tables = po.Tables(self.ld, self.npt2, po.Pf(self.xstar), po.Pf(self.tab["xavs"]),
                                po.Pf(self.tab["xavs1"]), po.Pf(self.tab["ecoul"]), po.Pf(self.tab["ecoul1"]))


def ledger_from_record(cap, nmol):
    """An oracle ledger holding a recorded ledger (0-based partners and entries)."""
    M = int(cap["nnbrefmax"])
    ne = int(cap["nergpair"])
    keep = {}
    keep["nergpairm"] = np.ascontiguousarray(cap["nergpairm"], np.int32)
    mol = np.zeros((nmol, M), np.int32)
    ref = np.zeros((nmol, M), np.int32)
    for im in range(nmol):
        n = int(cap["nergpairm"][im])
        mol[im, :n] = cap["molergpair"][1:n + 1, im] - 1
        ref[im, :n] = cap["ergpairref"][1:n + 1, im] - 1
    keep["molergpair"] = np.ascontiguousarray(mol.reshape(-1))
    keep["ergpairref"] = np.ascontiguousarray(ref.reshape(-1))
    keep["ergpair"] = np.ascontiguousarray(np.concatenate([cap["ergpair"].astype(F32), np.zeros(1, F32)]))
    led = po.Ledger()
    led.nmol, led.nnbrefmax, led.nergpair, led.nergpairmax = nmol, M, ne, ne + 1
    led.ergpair = po.Pf(keep["ergpair"])
    led.molergpair = po.Pi(keep["molergpair"])
    led.ergpairref = po.Pi(keep["ergpairref"])
    led.nergpairm = po.Pi(keep["nergpairm"])
    led._keep = keep
    return led


def ledger_arrays(led, nmol):
    """(nergpairm, partners per row, entry values per row) of an oracle ledger, 1-based
    partners so rows compare with the record directly."""
    M = led.nnbrefmax
    nm = np.ctypeslib.as_array(led.nergpairm, shape=(nmol,)).copy()
    mol = np.ctypeslib.as_array(led.molergpair, shape=(nmol * M,)).reshape(nmol, M)
    ref = np.ctypeslib.as_array(led.ergpairref, shape=(nmol * M,)).reshape(nmol, M)
    erg = np.ctypeslib.as_array(led.ergpair, shape=(max(led.nergpair, 1),))
    rows = [(mol[im, :nm[im]] + 1).copy() for im in range(nmol)]
    refs = [(ref[im, :nm[im]] + 1).copy() for im in range(nmol)]
    vals = [erg[ref[im, :nm[im]]].copy() for im in range(nmol)]
    return nm, rows, refs, vals, erg[:led.nergpair].copy()


def record_rows(cap, nmol):
    nm = cap["nergpairm"].astype(np.int32)
    rows = [cap["molergpair"][1:nm[im] + 1, im].astype(np.int32) for im in range(nmol)]
    refs = [cap["ergpairref"][1:nm[im] + 1, im].astype(np.int32) for im in range(nmol)]
    erg = cap["ergpair"].astype(F32)
    vals = [erg[r - 1] for r in refs]
    return nm, rows, refs, vals, erg


def self_energies(cap, nmol):
    """Each molecule's intramolecular energy as the reference stored it (slot 1 of its row
    when that slot is the molecule itself; 0 when nothing was stored)."""
    out = np.zeros(nmol, F32)
    for im in range(nmol):
        if int(cap["nergpairm"][im]) > 0 and int(cap["molergpair"][1, im]) == im + 1:
            out[im] = cap["ergpair"][int(cap["ergpairref"][1, im]) - 1]
    return out


def ran1(a):
    return po.Ran1.from_array(np.asarray(a, np.int32))


def f32_bits(a):
    return np.ascontiguousarray(a, F32).view(np.int32)


def as_ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def intra_from_record(st, tables, tempang=None, tempdih=None):
    """The oracle's intramolecular description (0-based) of the recorded box; tables supply
    the reference potential xav/xav1 on the same grid as the tables of `st`."""
    t = st.t
    c = np.ascontiguousarray
    keep = {}
    na = st.natom
    keep["qtype"] = c(t["qtype"], np.int32)
    keep["imageatom"] = c(t["imageatom"] - 1, np.int32)
    keep["mref"] = c(t["mref"] - 1, np.int32)
    keep["nbond"] = c(t["nbond"], np.int32)
    ldb = int(t["iabond"].shape[1]) - 1
    keep["iabond"] = c(t["iabond"][:, 1:] - 1, np.int32).reshape(-1)
    keep["tbond"] = c(t["tbond"][:, 1:], F32).reshape(-1)
    keep["amassr"] = c(t["amassr"], F32)
    keep["amassa"] = c(t["amassa"], F32)
    keep["ngrot"] = c(t["ngrot"], np.int32)
    keep["nrf"] = c(t["nrf"] - 1, np.int32)
    keep["nrl"] = c(t["nrl"] - 1, np.int32)
    ng = int(t["labaxi"].shape[0]) - 1
    keep["labaxi"] = c(t["labaxi"][1:, 1:3] - 1, np.int32).reshape(-1)
    keep["rotf"] = c(t["rotf"][1:] - 1, np.int32)
    keep["rotl"] = c(t["rotl"][1:] - 1, np.int32)
    keep["labrot"] = c(t["labrot"][1:] - 1, np.int32)
    keep["isdih"] = c([1 if str(g) == "DIH" else 0 for g in t["rotgrp"][1:]], np.int32)
    keep["ndihedralang"] = c(t["ndihedralang"][1:ng + 1], np.int32)
    ldd = max(int(t["labrotdih"].shape[1]) - 1, 1)
    lrd = np.zeros((ng, ldd, 2), np.int32)
    dang = np.zeros((ng, ldd), F32)
    if t["labrotdih"].shape[1] > 1:
        lrd[:, :, 0] = t["labrotdih"][1, 1:, 1:].T - 1
        lrd[:, :, 1] = t["labrotdih"][2, 1:, 1:].T - 1
        dang[:, :] = t["dihedralang"][1:, 1:].T
    keep["labrotdih"] = c(lrd.reshape(-1))
    keep["dihedralang"] = c(dang.reshape(-1))
    keep["xav"] = c(tables["xav"], F32).reshape(-1).copy()
    keep["xav1"] = c(tables["xav1"], F32).reshape(-1).copy()
    assert keep["qtype"].shape[0] == na
    it = po.Intra(po.Pi(keep["qtype"]), po.Pi(keep["imageatom"]), po.Pi(keep["mref"]), po.Pi(keep["nbond"]),
                  po.Pi(keep["iabond"]), po.Pf(keep["tbond"]), ldb, po.Pf(keep["amassr"]), po.Pf(keep["amassa"]),
                  po.Pi(keep["ngrot"]), po.Pi(keep["nrf"]), po.Pi(keep["nrl"]), po.Pi(keep["labaxi"]),
                  po.Pi(keep["rotf"]), po.Pi(keep["rotl"]), po.Pi(keep["labrot"]), po.Pi(keep["isdih"]),
                  po.Pi(keep["ndihedralang"]), po.Pi(keep["labrotdih"]), po.Pf(keep["dihedralang"]), ldd,
                  po.Pf(keep["xav"]), po.Pf(keep["xav1"]), F32(t["dmaxmol"]), F32(t["tempb"]), F32(t["simt"]),
                  F32(t["tempangreq"] if tempang is None else tempang),
                  F32(F32(t["tempdihreq"]) * F32(t["tempb"]) if tempdih is None else tempdih),
                  F32(t["refine_intra_fac"]), int(bool(t["refine_intra"])))
    it._keep = keep
    return it
