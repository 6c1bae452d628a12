"""Intramolecular topology of a box: bonds and angle constraints, image atoms, Q-type flags,
masses, rotational groups and dihedrals, in the 0-based layout of epsr_intra_desc_t
(include/epsr_mi355x.h), built from a parsed .ato file the way the reference's readers
leave it in COMMON:

  nbond/iabond/tbond   minitc's bond records (src/epsr26/library/ato_read_and_save.py:347-356),
                       atom numbers made global (bondnumbers), lengths / dmaxmol
                       (normalise_coordinates, :916-970)
  amassr               sqrtf(m1 m2 / (m1 + m2)) * dmaxmol per type pair (:859-863, :969-970)
  rotational groups    labaxi, rotf/rotl/labrot, DIH dihedrals (:365-452), angles in radians
                       (:971-978)
"""

from __future__ import annotations

import ctypes as C
import dataclasses

import numpy as np

F32 = np.float32
PI4ATAN = F32(3.14159265358979323846)


@dataclasses.dataclass
class IntraTopology:
    qtype: np.ndarray          # int32 [natom]
    imageatom: np.ndarray      # int32 [natom], -1 for none
    nbond: np.ndarray          # int32 [natom]
    iabond: np.ndarray         # int32 [natom, ldb]
    tbond: np.ndarray          # float32 [natom, ldb], molecule box units
    amassr: np.ndarray         # float32 [npairs]
    amassa: np.ndarray         # float32 [ntype]
    nrf: np.ndarray            # int32 [nmol], first group
    nrl: np.ndarray            # int32 [nmol], last group (< nrf: none)
    labaxi: np.ndarray         # int32 [ngroup, 2]
    rotf: np.ndarray           # int32 [ngroup]
    rotl: np.ndarray           # int32 [ngroup]
    labrot: np.ndarray         # int32 [nlabrot]
    isdih: np.ndarray          # int32 [ngroup]
    ndihedralang: np.ndarray   # int32 [ngroup]
    labrotdih: np.ndarray      # int32 [ngroup, ldd, 2]
    dihedralang: np.ndarray    # float32 [ngroup, ldd], radians
    dmaxmol: np.float32
    tempb: np.float32
    refine_intra_fac: np.float32 = F32(0.0)

    @property
    def ngroup(self):
        return int(self.labaxi.shape[0])

    @property
    def ngrot(self):
        return np.maximum(self.nrl - self.nrf + 1, 0).astype(np.int32)

    def desc(self):
        """The ctypes descriptor (epsr_intra_desc_t) and the arrays it points into."""
        from . import _native
        c = np.ascontiguousarray
        keep = dict(qtype=c(self.qtype, np.int32), imageatom=c(self.imageatom, np.int32), nbond=c(self.nbond, np.int32),
                    iabond=c(self.iabond, np.int32).reshape(-1), tbond=c(self.tbond, F32).reshape(-1),
                    amassr=c(self.amassr, F32), amassa=c(self.amassa, F32), nrf=c(self.nrf, np.int32),
                    nrl=c(self.nrl, np.int32), labaxi=c(self.labaxi, np.int32).reshape(-1),
                    rotf=c(self.rotf, np.int32), rotl=c(self.rotl, np.int32), labrot=c(self.labrot, np.int32),
                    isdih=c(self.isdih, np.int32), ndihedralang=c(self.ndihedralang, np.int32),
                    labrotdih=c(self.labrotdih, np.int32).reshape(-1), dihedralang=c(self.dihedralang, F32).reshape(-1))
        for k in ("labaxi", "rotf", "rotl", "labrot", "isdih", "ndihedralang", "labrotdih", "dihedralang"):
            if keep[k].size == 0:
                keep[k] = np.zeros(2, keep[k].dtype)
        ip = lambda a: a.ctypes.data_as(C.POINTER(C.c_int))
        fp = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))
        ldb = int(self.iabond.shape[1]) if self.iabond.ndim == 2 else 0
        ldd = max(1, int(self.labrotdih.shape[1]) if self.labrotdih.ndim == 3 else 1)
        d = _native.IntraDesc(ldb, ip(keep["qtype"]), ip(keep["imageatom"]), ip(keep["nbond"]), ip(keep["iabond"]),
                              fp(keep["tbond"]), fp(keep["amassr"]), fp(keep["amassa"]), self.ngroup,
                              int(self.labrot.size), ldd, ip(keep["nrf"]), ip(keep["nrl"]), ip(keep["labaxi"]),
                              ip(keep["rotf"]), ip(keep["rotl"]), ip(keep["labrot"]), ip(keep["isdih"]),
                              ip(keep["ndihedralang"]), ip(keep["labrotdih"]), fp(keep["dihedralang"]),
                              F32(self.dmaxmol), F32(self.tempb), F32(self.refine_intra_fac),
                              int(self.refine_intra_fac > 0.0))
        return d, keep


def from_ato(a, type_index, npairref, amass, dmaxmol, refine_intra_fac=F32(0.0)):
    """The topology of parsed .ato `a` (formats.AtoFile); type_index: the type of every atom
    (0-based), npairref [ntype, ntype], amass [ntype] float32, dmaxmol = dmax * sizefactor."""
    dmaxmol = F32(dmaxmol)
    natom = a.natom
    ldb = max([len(b) for m in a.molecules for b in m.bonds] + [1])
    qtype = np.zeros(natom, np.int32)
    imageatom = np.full(natom, -1, np.int32)
    nbond = np.zeros(natom, np.int32)
    iabond = np.zeros((natom, ldb), np.int32)
    tbond = np.zeros((natom, ldb), F32)
    nrf, nrl = [], []
    labaxi, rotf, rotl, labrot, isdih, ndih, dihs = [], [], [], [], [], [], []
    atom = 0
    for m in a.molecules:
        for k in range(m.nato):
            ia = atom + k
            qtype[ia] = 1 if m.labels[k][:1] in ("Q", "q") else 0
            im_ = int(m.imageatom[k]) if k < len(m.imageatom) else 0
            imageatom[ia] = atom + im_ - 1 if im_ > 0 else -1
            bl = m.bonds[k] if k < len(m.bonds) else []
            nbond[ia] = len(bl)
            for j, (partner, length) in enumerate(bl):
                iabond[ia, j] = atom + int(partner) - 1
                tbond[ia, j] = F32(length) / dmaxmol
        nrf.append(len(labaxi))
        for g in m.rotgroups:
            labaxi.append((atom + int(g.axis[0]) - 1, atom + int(g.axis[1]) - 1))
            rotf.append(len(labrot))
            labrot.extend(atom + int(x) - 1 for x in g.atoms)
            rotl.append(len(labrot) - 1)
            dih = g.kind.strip().upper() == "DIH"
            isdih.append(1 if dih else 0)
            ndih.append(len(g.dihedrals) if dih else 0)
            dihs.append([(atom + int(d0) - 1, atom + int(d1) - 1, F32(ang) * (PI4ATAN / F32(180.0)))
                         for (d0, d1, ang) in (g.dihedrals if dih else [])])
        nrl.append(len(labaxi) - 1)
        atom += m.nato
    ng = len(labaxi)
    ldd = max([len(x) for x in dihs] + [1])
    labrotdih = np.zeros((ng, ldd, 2), np.int32)
    dang = np.zeros((ng, ldd), F32)
    for g, lst in enumerate(dihs):
        for k, (d0, d1, ang) in enumerate(lst):
            labrotdih[g, k] = (d0, d1)
            dang[g, k] = ang
    amass = np.asarray(amass, F32)
    npr = np.asarray(npairref)
    npairs = int(npr.max()) + 1
    amassr = np.zeros(npairs, F32)
    nt = amass.shape[0]
    for it in range(nt):
        for jt in range(it, nt):
            am1, am2 = amass[it], amass[jt]
            amassr[npr[it, jt]] = F32(np.sqrt(F32(F32(am1 * am2) / F32(am1 + am2)), dtype=F32)) * dmaxmol
    return IntraTopology(qtype=qtype, imageatom=imageatom, nbond=nbond, iabond=iabond, tbond=tbond, amassr=amassr,
                         amassa=amass.copy(), nrf=np.array(nrf, np.int32), nrl=np.array(nrl, np.int32),
                         labaxi=np.array(labaxi, np.int32).reshape(-1, 2), rotf=np.array(rotf, np.int32),
                         rotl=np.array(rotl, np.int32), labrot=np.array(labrot, np.int32),
                         isdih=np.array(isdih, np.int32), ndihedralang=np.array(ndih, np.int32), labrotdih=labrotdih,
                         dihedralang=dang, dmaxmol=dmaxmol, tempb=F32(a.tempb), refine_intra_fac=F32(refine_intra_fac))
