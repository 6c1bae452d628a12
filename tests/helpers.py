"""Shared test fixtures: synthetic systems plus the tables the oracle builds for them."""

from __future__ import annotations

import ctypes as C
import functools
import os

import numpy as np

from epsr_python_translation_amd import formats, synth, tables
from oracle import pyoracle as po

F32 = np.float32


EPSR_TEST = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "epsr_test")


@functools.lru_cache(maxsize=None)
def epsr_test_files():
    """The reference's EPSR test box: (.ato, .pcof, .EPSR.inp) parsed."""
    return (formats.read_ato(os.path.join(EPSR_TEST, "200DMSO400Water.ato.gz")),
            formats.read_pcof(os.path.join(EPSR_TEST, "200DMSO400Water.pcof.gz")),
            formats.read_keywords(os.path.join(EPSR_TEST, "200DMSO400Water.EPSR.inp.gz")))


@functools.lru_cache(maxsize=None)
def system(name: str):
    if name == "epsr_test":
        return formats.system_from_files(*epsr_test_files(), name="epsr_test")
    if name == "dense_1400":
        # 1400 waters in a 30 A box: every molecule in every ledger row (rows of 1400 slots)
        return synth.tiny(seed=5, nwater=1400, L=30.0)
    if name == "c0":
        return synth.config0()
    if name == "c1":
        return synth.config1()
    if name == "c3":
        return synth.config3()
    if name == "c4":
        return synth.config4()
    if name.startswith("fine"):
        # 0.004 A bins to 8 A: 21 pairs x 2001 bins, a histogram larger than one workgroup's LDS
        parts = name.split("_")
        return synth.tiny(seed=int(parts[1]), nwater=int(parts[2]), ndmso=int(parts[3]), L=16.0, cellst=0.004)
    if name.startswith("tiny"):
        parts = name.split("_")
        seed = int(parts[1]) if len(parts) > 1 else 11
        nw = int(parts[2]) if len(parts) > 2 else 10
        nd = int(parts[3]) if len(parts) > 3 else 0
        return synth.tiny(seed=seed, nwater=nw, ndmso=nd)
    raise KeyError(name)


@functools.lru_cache(maxsize=None)
def grid(name: str):
    return tables.make_grid(system(name))


def foqols(name: str, flavour="glibc", fwhm=0.0, fwhmq=0.0, lorchqwidth=0.0, lorchqwidthq=0.0):
    s, g = system(name), grid(name)
    ldf = g.npt1 + 1
    fo = np.zeros((g.nq + 1) * ldf, dtype=F32)
    po.lib(flavour).ora_foqols(g.nq, po.Pf(g.q), g.npt1, po.Pf(g.x), po.Pf(g.ycoh), F32(s.rho), F32(fwhm),
                               F32(fwhmq), F32(lorchqwidth), F32(lorchqwidthq), F32(s.rmaxgr), po.Pf(fo), ldf)
    return fo, ldf


def truncation(name: str, flavour="glibc"):
    s, g = system(name), grid(name)
    L = po.lib(flavour)
    wfac = np.array([0.0] + [L.ora_truncate(F32(g.x[j]), F32(s.rminpt), F32(s.rmaxpt)) for j in range(1, g.npt + 1)],
                    dtype=F32)
    wfacc = np.array([0.0] + [L.ora_truncatec(F32(g.x[j]), F32(s.rmaxpt)) for j in range(1, g.npt + 1)], dtype=F32)
    wfac[g.npt] = 0.0
    wfacc[g.npt] = 0.0
    return wfac, wfacc


@functools.lru_cache(maxsize=None)
def oracle_tables(name: str, flavour="glibc"):
    """reference potential (generate_rp) + EP (generate_ep) from the oracle, as the
    tables both sides of an MC parity test read."""
    s, g = system(name), grid(name)
    L = po.lib(flavour)
    ld = g.ld
    npairs = s.npairs
    wfac, wfacc = truncation(name, flavour)
    epot = np.zeros(npairs * ld, F32)
    ecoul = np.zeros(npairs * ld, F32)
    epot1 = np.zeros(npairs * ld, F32)
    ecoul1 = np.zeros(npairs * ld, F32)
    qr = np.zeros(npairs, F32)
    L.ora_generate_rp(g.npt, g.npt2, po.Pf(g.x), po.Pf(g.xstar), po.Pf(wfac), po.Pf(wfacc), npairs, po.Pf(s.ecor),
                      po.Pf(s.dcor), po.Pf(s.qcor), po.Pf(qr), po.Pf(qr), F32(s.power), F32(1.0), None, None,
                      F32(0.3), F32(0.0), po.Pf(epot), po.Pf(ecoul), po.Pf(epot1), po.Pf(ecoul1), ld)
    xavs = np.zeros(npairs * ld, F32)
    xavs1 = np.zeros(npairs * ld, F32)
    xavtot = np.zeros(npairs * ld, F32)
    xavtot1 = np.zeros(npairs * ld, F32)
    pcoefp = np.ascontiguousarray(s.pcoefp.astype(F32).copy())
    irefpcof = np.arange(npairs, dtype=np.int32)
    energabs = C.c_float()
    L.ora_generate_ep(g.npt, g.npt2, po.Pf(g.x), po.Pf(g.xstar), po.Pf(wfac), npairs, po.Pi(irefpcof), s.nterms,
                      po.Pf(pcoefp.reshape(-1)), s.nterms + 1, F32(s.pdmin), F32(s.pdstep), F32(s.rmaxpt),
                      F32(s.psigma2), F32(s.rho), F32(0.0), F32(s.dmax), F32(10.0), F32(0.0), F32(0.0),
                      po.Pf(ecoul), po.Pf(xavs), po.Pf(xavs1), po.Pf(xavtot), po.Pf(xavtot1), ld,
                      C.byref(energabs))
    return dict(ld=ld, npt2=g.npt2, xstar=g.xstar.copy(), xavs=xavs, xavs1=xavs1, ecoul=ecoul, ecoul1=ecoul1,
                epot=epot, epot1=epot1, wfac=wfac, wfacc=wfacc, pcoefp=pcoefp)


def ctables(t):
    return po.Tables(t["ld"], t["npt2"], po.Pf(t["xstar"]), po.Pf(t["xavs"]), po.Pf(t["xavs1"]),
                     po.Pf(t["ecoul"]), po.Pf(t["ecoul1"]))


def oracle_atogen(name, ngrsamples=0, idum=-12345, flavour="glibc", rminex=None, coords=None):
    s, g = system(name), grid(name)
    ob = po.OracleBox(s, **(coords or {}))
    ldw = g.npt1 + 1
    wav = np.zeros(2 * s.npairs * ldw, F32)
    potex = np.zeros(s.npairs, F32)
    potexmax = np.zeros(s.npairs, F32)
    cs = np.zeros(s.ntype, F32)
    irefpcof = np.arange(s.npairs, dtype=np.int32)
    rminex = np.zeros(s.npairs, F32) if rminex is None else np.ascontiguousarray(rminex, dtype=F32)
    rng = po.Ran1.seeded(idum)
    ngr = po.lib(flavour).ora_atogen(C.byref(ob.box), C.byref(ob.centres), g.npt1, ldw, po.Pf(g.ycoh),
                                     po.Pi(irefpcof), po.Pf(rminex), ngrsamples, C.byref(rng), po.Pf(wav),
                                     po.Pf(potex), po.Pf(potexmax), po.Pf(cs), None)
    return dict(wav=wav, potex=potex, ngr=ngr, ran1=rng.as_array(), ldw=ldw, csampled=cs)


def oracle_cells(name, xadd=None, yadd=None, zadd=None):
    s, g = system(name), grid(name)
    cl = po.Cells()
    xa = np.ascontiguousarray(s.xadd if xadd is None else xadd, dtype=F32)
    ya = np.ascontiguousarray(s.yadd if yadd is None else yadd, dtype=F32)
    za = np.ascontiguousarray(s.zadd if zadd is None else zadd, dtype=F32)
    from epsr_python_translation_amd import tables as T
    po.lib().ora_cells_init(C.byref(cl), s.nmolcell, po.Pf(xa), po.Pf(ya), po.Pf(za), s.nmol, T.sepmax(s, g))
    return cl


def ledger_dict(led, nmol):
    M = led.nnbrefmax
    cnt = np.ctypeslib.as_array(led.nergpairm, shape=(nmol,)).copy()
    mol = np.ctypeslib.as_array(led.molergpair, shape=(nmol * M,)).reshape(nmol, M).copy()
    ref = np.ctypeslib.as_array(led.ergpairref, shape=(nmol * M,)).reshape(nmol, M).copy()
    erg = np.ctypeslib.as_array(led.ergpair, shape=(max(1, led.nergpair),))[:led.nergpair].copy()
    return dict(nergpairm=cnt, molergpair=mol, ergpairref=ref, ergpair=erg, nnbrefmax=M)


def oracle_fmolec1(name, flavour="glibc", ob=None, cells=None):
    s = system(name)
    t = oracle_tables(name)
    ob = ob or po.OracleBox(s)
    cl = cells or oracle_cells(name, ob.xadd, ob.yadd, ob.zadd)
    led = po.Ledger()
    ft, fb = C.c_float(), C.c_float()
    f64 = C.c_double()
    tt = ctables(t)
    ei = eintra(name, flavour)
    po.lib(flavour).ora_fmolec1(C.byref(ob.box), C.byref(ob.centres), C.byref(tt), C.byref(cl), po.Pf(ei),
                                C.byref(led), C.byref(ft), C.byref(fb), C.byref(f64))
    return dict(ledger=led, cells=cl, ob=ob, tables=t, ctab=tt, ftest=ft.value, ftestb=fb.value, ftest64=f64.value)


def initial_steps(s):
    """Initial whole-molecule rotation (fraction of pi) and translation (box fraction) steps."""
    return (np.full(s.nmoltype, 0.05, F32), np.full(s.nmoltype, F32(0.1) / F32(s.dmax), F32))


def box_volume(s):
    """clustercom.volume of a cubic box: alatx * blaty * clatz in float32 (Angstrom^3)."""
    return F32(s.dmax) * F32(s.dmax) * F32(s.dmax)


def oracle_press(name, flavour="glibc", ref=None, tables=None, slots=True):
    """ora_press over the oracle's own ledger of `name` (or of ref = oracle_fmolec1(...))."""
    s, g = system(name), grid(name)
    ref = ref or oracle_fmolec1(name, flavour)
    tt = ctables(tables) if tables is not None else ref["ctab"]
    led = ref["ledger"]
    M = max(1, led.nnbrefmax)
    se = np.zeros(s.nmol * M, F32)
    sw = np.zeros(s.nmol * M, F32)
    e, p = C.c_float(), C.c_float()
    w = C.c_double()
    po.lib(flavour).ora_press(C.byref(ref["ob"].box), C.byref(ref["ob"].centres), C.byref(tt), C.byref(led), g.npt1,
                              F32(s.simt), box_volume(s), C.byref(e), C.byref(p), C.byref(w),
                              po.Pf(se) if slots else None, po.Pf(sw) if slots else None)
    return dict(energy=F32(e.value), pressc=F32(p.value), presst=w.value, slote=se.reshape(s.nmol, M),
                slotw=sw.reshape(s.nmol, M), ref=ref)


def oracle_intra(s, ld, tempang, tempdih):
    """po.Intra of a box with an intramolecular topology (s.intra, intra.IntraTopology);
    xav/xav1 are zero (no intramolecular Coulomb weighting) on tables of leading dimension ld."""
    it = s.intra
    c = np.ascontiguousarray
    keep = dict(qtype=c(it.qtype, np.int32), imageatom=c(it.imageatom, np.int32), mref=c(np.repeat(
        np.arange(s.nmol, dtype=np.int32), s.nal - s.naf + 1)), nbond=c(it.nbond, np.int32),
        iabond=c(it.iabond, np.int32).reshape(-1), tbond=c(it.tbond, F32).reshape(-1), amassr=c(it.amassr, F32),
        amassa=c(it.amassa, F32), ngrot=c(it.ngrot, np.int32), nrf=c(it.nrf, np.int32), nrl=c(it.nrl, np.int32),
        labaxi=c(it.labaxi, np.int32).reshape(-1), rotf=c(it.rotf, np.int32), rotl=c(it.rotl, np.int32),
        labrot=c(it.labrot, np.int32), isdih=c(it.isdih, np.int32), ndihedralang=c(it.ndihedralang, np.int32),
        labrotdih=c(it.labrotdih, np.int32).reshape(-1), dihedralang=c(it.dihedralang, F32).reshape(-1),
        xav=np.zeros(s.npairs * ld, F32), xav1=np.zeros(s.npairs * ld, F32))
    for k, v in keep.items():
        if v.size == 0:
            keep[k] = np.zeros(2, v.dtype)
    ldb = int(it.iabond.shape[1])
    ldd = max(1, int(it.labrotdih.shape[1]))
    o = po.Intra(po.Pi(keep["qtype"]), po.Pi(keep["imageatom"]), po.Pi(keep["mref"]), po.Pi(keep["nbond"]),
                 po.Pi(keep["iabond"]), po.Pf(keep["tbond"]), ldb, po.Pf(keep["amassr"]), po.Pf(keep["amassa"]),
                 po.Pi(keep["ngrot"]), po.Pi(keep["nrf"]), po.Pi(keep["nrl"]), po.Pi(keep["labaxi"]),
                 po.Pi(keep["rotf"]), po.Pi(keep["rotl"]), po.Pi(keep["labrot"]), po.Pi(keep["isdih"]),
                 po.Pi(keep["ndihedralang"]), po.Pi(keep["labrotdih"]), po.Pf(keep["dihedralang"]), ldd,
                 po.Pf(keep["xav"]), po.Pf(keep["xav1"]), F32(it.dmaxmol), F32(it.tempb), F32(s.simt),
                 F32(tempang), F32(tempdih), F32(it.refine_intra_fac), int(it.refine_intra_fac > 0.0))
    o._keep = keep
    return o


def fmolec1_weights(s):
    """(tempang, tempdih) as fmolec1 sets them: tempangreq and tempdihreq * tempb"""
    r = s.mc_run
    return F32(r["tempangreq"]), F32(F32(r["tempdihreq"]) * F32(s.intra.tempb))


@functools.lru_cache(maxsize=None)
def eintra(name: str, flavour="glibc"):
    """The intramolecular energy of every molecule: the synthetic boxes' fixed values, or the
    oracle's en_mole_intra at the box's coordinates for a box read from files."""
    s = system(name)
    if s.eintra is not None:
        return np.ascontiguousarray(s.eintra, dtype=F32)
    t = oracle_tables(name)
    ob = po.OracleBox(s)
    it = oracle_intra(s, t["ld"], *fmolec1_weights(s))
    tt = ctables(t)
    L = po.lib(flavour)
    out = np.array([L.ora_en_mole_intra(C.byref(ob.box), C.byref(tt), C.byref(it), im, None, None, None)
                    for im in range(s.nmol)], F32)
    return out
