"""EPSR's on-disk formats: `.ato` configurations, `.pcof` potential files and the
keyword files (`.EPSR.inp`, and the head of `.pcof`).

Readers follow the reference's readers value by value; writers reproduce its writers
byte for byte (gfortran's formatted and list-directed output), so that a file the
reference wrote reads and writes back unchanged and a box can be restarted from what
this package writes:

* `read_ato` / `write_ato`: minitc and moloutc
  (src/epsr26/library/ato_read_and_save.py:64-915, 1103-1350), including the saved ran1
  state (library/auxiliary_routines.py:272-320);
* `read_pcof` / `write_pcof`: read_pcof and the .pcof half of save_files_others
  (library/read_and_save_files.py:885-1300, 645-760);
* `read_keywords` / `write_keywords`: the keyword, value, text lines and `data N`
  blocks of .EPSR.inp (read_and_save_files.py:558-640, format `(a11,1x,a,15x,a)`);
* `system_from_files`: the box the GPU path runs (`synth.System`), with the unit
  conversion of normalise_coordinates (ato_read_and_save.py:916-1000) and the pair
  parameters of minitc (:845-858).

Numbers are read the way gfortran reads a REAL*4 (the decimal value correctly rounded to
float32, ties to even) and written the way gfortran writes one (correctly rounded
decimal digits; Ew.d drops the leading zero when the field is one short; list-directed
REAL*4 is G with 9 significant digits in a 16-character field, four blanks in place of
the exponent in F form).  All host-side: no file is read on the GPU.
"""

from __future__ import annotations

import dataclasses
import gzip
import math
from decimal import Decimal
from fractions import Fraction

import numpy as np

F32 = np.float32

# ------------------------------------------------------------------ numbers in


def f32(tok) -> np.float32:
    """A Fortran real constant read into REAL*4: the decimal value rounded once to the
    nearest float32 (ties to even)."""
    t = str(tok).strip().upper().replace("D", "E").replace("Q", "E")
    d = float(t)
    f = F32(d)
    if d == 0.0 or not np.isfinite(f):
        return f
    # float(t) rounds the decimal to double; rounding that again to float32 can differ
    # from one rounding only when the double lies next to a float32 midpoint
    lo = np.nextafter(f, F32(-np.inf))
    hi = np.nextafter(f, F32(np.inf))
    ud = math.ulp(d)
    near = any(abs(d - (float(f) + float(n)) / 2.0) <= 2.0 * ud for n in (lo, hi) if np.isfinite(n))
    if not near:
        return f
    x = Fraction(Decimal(t))
    best = min((lo, f, hi), key=lambda c: (abs(Fraction(float(c)) - x),
                                           int(np.array(c, dtype=F32).view(np.int32)) & 1))
    return F32(best)


def fint(tok) -> int:
    return int(str(tok).strip())


def parse_words(text: str, nspcword: int, nchartext: int = 3020):
    """EPSR's word splitter (library/parse.py:5-55): blank-separated words, stopping once
    more than `nspcword` consecutive blanks follow a word, which is how the free text
    after each value is skipped.  Returns [(first, last)] 0-based inclusive spans."""
    lt = len(text)

    def ch(i):  # 1-based, blank beyond the text
        return text[i - 1] if i <= lt else " "

    spans = []
    nspace = 0
    i = 1
    while ch(i) == " " and i < nchartext:
        i += 1
    nfirst = nlast = i
    while i < nchartext and nspace <= nspcword:
        i += 1
        if ch(i) == " ":
            if nlast >= nfirst:
                spans.append((nfirst - 1, nlast - 1))
                nspace = 1
            else:
                nspace += 1
            nfirst = i + 1
        else:
            nlast = i
    if ch(i) != " " and nlast >= nfirst:
        spans.append((nfirst - 1, nlast - 1))
    return spans


def words(text: str, nspcword: int):
    return [text[a:b + 1] for a, b in parse_words(text, nspcword)]


def _tokens(line: str):
    """List-directed tokens of one record (blanks and commas separate; '/' ends the
    statement); repeat counts r*value expand."""
    out = []
    for part in line.replace(",", " ").split():
        if part.startswith("/"):
            out.append("/")
            break
        if "*" in part and part.split("*", 1)[0].isdigit():
            r, v = part.split("*", 1)
            out.extend([v] * int(r))
            continue
        out.append(part)
    return out


class _Lines:
    """The records of a text file, read the way Fortran READ statements consume them."""

    def __init__(self, text: str):
        self.lines = text.splitlines()
        self.k = 0

    def line(self) -> str:
        if self.k >= len(self.lines):
            raise EOFError("end of file")
        s = self.lines[self.k]
        self.k += 1
        return s

    def items(self, n: int):
        """One list-directed READ of n items: starts at the next record and takes as
        many records as it needs; the rest of its last record is skipped."""
        got = []
        while len(got) < n:
            toks = _tokens(self.line())
            if "/" in toks:
                got.extend(toks[:toks.index("/")])
                break
            got.extend(toks)
        return got[:n]

    def counted(self, per: int):
        """One list-directed READ of a count n followed by n groups of `per` items."""
        toks = _tokens(self.line())
        need = 1 + per * int(toks[0])
        while len(toks) < need:
            toks += _tokens(self.line())
        return toks[:need]

    def at_end(self) -> bool:
        return self.k >= len(self.lines)


def _read_text(path) -> str:
    p = str(path)
    if p.endswith(".gz"):
        with gzip.open(p, "rt", newline="") as f:
            return f.read()
    with open(p, "r", newline="") as f:
        return f.read()


# ------------------------------------------------------------------ numbers out


def _sign(x) -> bool:
    return math.copysign(1.0, float(x)) < 0


def _nonfinite(v: float, w: int) -> str:
    """gfortran's NaN and Infinity fields (Inf when the field is too short)."""
    if math.isnan(v):
        return "NaN".rjust(w)
    s = ("-" if v < 0 else "") + ("Infinity" if w >= 8 + (v < 0) else "Inf")
    return s.rjust(w) if len(s) <= w else "*" * w


def fe(x, w: int, d: int) -> str:
    """gfortran Ew.d: 0.ddddE+xx; the leading zero goes when the field is one short."""
    v = float(F32(x)) if not isinstance(x, float) else x
    if not math.isfinite(v):
        return _nonfinite(v, w)
    neg = _sign(v)
    a = abs(v)
    if a == 0.0:
        digits, ex = "0" * d, 0
    else:
        s = f"{a:.{d - 1}e}"
        mant, e = s.split("e")
        digits = mant.replace(".", "")
        ex = int(e) + 1
    es = f"E{'-' if ex < 0 else '+'}{abs(ex):02d}" if abs(ex) <= 99 else f"{'-' if ex < 0 else '+'}{abs(ex):03d}"
    body = "0." + digits + es
    txt = ("-" if neg else "") + body
    if len(txt) > w:
        txt = ("-" if neg else "") + body[1:]
    return txt.rjust(w) if len(txt) <= w else "*" * w


def ff(x, w: int, d: int) -> str:
    """gfortran Fw.d (a negative value that rounds to zero keeps its minus sign)."""
    v = float(F32(x)) if not isinstance(x, float) else x
    if not math.isfinite(v):
        return _nonfinite(v, w)
    neg = _sign(v)
    body = f"{abs(v):.{d}f}"
    txt = ("-" if neg else "") + body
    if len(txt) > w and body.startswith("0."):
        txt = ("-" if neg else "") + body[1:]
    return txt.rjust(w) if len(txt) <= w else "*" * w


def fi(n: int, w: int = 0) -> str:
    """Iw (w = 0: I0)."""
    s = str(int(n))
    if w == 0:
        return s
    return s.rjust(w) if len(s) <= w else "*" * w


def fa(s: str, w: int) -> str:
    """Aw of a character variable of length w (blank padded or cut)."""
    return (s + " " * w)[:w]


def lw_int(n: int) -> str:
    return " " + str(int(n)).rjust(11)


def lw_real(x) -> str:
    """gfortran's list-directed REAL*4 item: a separator blank and a 16-wide G field
    with 9 significant digits (F form for 0.1 <= |x| < 1e9, four trailing blanks)."""
    v = float(F32(x))
    if not math.isfinite(v):
        return " " + _nonfinite(v, 16)
    neg = _sign(v)
    a = abs(v)
    if a == 0.0:
        body = "0.00000000"
        return " " + (("-" if neg else "") + body).rjust(12) + "    "
    s = f"{a:.8e}"
    mant, e = s.split("e")
    e10 = int(e)
    if -1 <= e10 <= 8:
        body = f"{a:.{8 - e10}f}"
        return " " + (("-" if neg else "") + body).rjust(12) + "    "
    es = f"E{'-' if e10 < 0 else '+'}{abs(e10):02d}"
    return " " + (("-" if neg else "") + mant + es).rjust(16)


def lw_line(items) -> str:
    return "".join(lw_int(v) if isinstance(v, (int, np.integer)) else lw_real(v) for v in items)


# ------------------------------------------------------------------ .ato


@dataclasses.dataclass
class AtoType:
    """One atom type's parameters (ctype, ato_manipulation_rev4.py:283-340)."""

    label: str                 # character*3
    atsymbol: str = "XXX"
    iexchange: int = -1
    eps: np.float32 = F32(0)
    sig: np.float32 = F32(0)
    mass: np.float32 = F32(0)
    charge: np.float32 = F32(0)
    qradius: np.float32 = F32(0)
    natomsphere: np.float32 = F32(1)
    radminsphere: np.float32 = F32(0)
    radmaxsphere: np.float32 = F32(0)
    pseudofac: np.float32 = F32(1)


@dataclasses.dataclass
class RotGroup:
    """A rotational group of a molecule (minitc :369-450); atom numbers are 1-based
    within the molecule, as in the file."""

    kind: str                  # character*3: ROT, DIH, ...
    axis: tuple
    dihedrals: list            # [(a1, a2, angle degrees)]
    atoms: list


@dataclasses.dataclass
class AtoMolecule:
    nato: int
    centre: np.ndarray         # float32[3], Angstrom
    phi: np.ndarray            # float32[3]
    tethered: bool = False
    tetheredatom: int = 0
    nstepreject: int = 0
    scalefac: np.float32 = F32(1)
    labels: list = dataclasses.field(default_factory=list)       # character*3 per atom
    imageatom: list = dataclasses.field(default_factory=list)
    xyz: np.ndarray = None     # float32[nato, 3], Angstrom from the centre
    bonds: list = dataclasses.field(default_factory=list)         # per atom [(atom 1-based, length)]
    rotgroups: list = dataclasses.field(default_factory=list)


@dataclasses.dataclass
class AtoTemplate:
    """A molecule type's template line (minitc :633-700)."""

    name: str
    interntrans: np.float32 = F32(0)
    internrot: np.float32 = F32(0)
    wholerot: np.float32 = F32(0)
    wholetrans: np.float32 = F32(0)
    rfract: np.float32 = F32(0)


@dataclasses.dataclass
class AtoFile:
    nmol: int
    cubic: bool
    dmax: np.float32           # cubic: the box side
    simtk: np.float32
    lattice: tuple             # (alat, blat, clat, phib, thetac, phic)
    tol: np.float32
    stepmi: np.float32
    stepri: np.float32
    stepr: np.float32
    stepm: np.float32
    tempb: np.float32
    tempangreq: np.float32
    tempdihreq: np.float32
    tempangwt: np.float32
    tempdihwt: np.float32
    molecules: list
    types: list                # AtoType in type-number order (first occurrence of the label)
    ecore: np.float32
    dcore: np.float32
    ran1: np.ndarray           # int32[35] = (mcall, idum, iy, iv[32]) or None
    moltypes: list             # molecule-type labels (first atom's label), first-occurrence order
    templates: list            # AtoTemplate per molecule type

    # ---- derived (ftype / minitc bookkeeping)
    def itype(self):
        idx = {t.label: i for i, t in enumerate(self.types)}
        return np.array([idx[lab] for m in self.molecules for lab in m.labels], dtype=np.int32)

    def mtype(self):
        idx = {lab: i for i, lab in enumerate(self.moltypes)}
        return np.array([idx[m.labels[0]] for m in self.molecules], dtype=np.int32)

    @property
    def natom(self):
        return sum(m.nato for m in self.molecules)

    def qtype(self):
        return np.array([1 if lab[:1] in ("Q", "q") else 0 for m in self.molecules for lab in m.labels],
                        dtype=np.int32)


def _type_order(molecules):
    seen = []
    for m in molecules:
        for lab in m.labels:
            if lab not in seen:
                seen.append(lab)
    return seen


def read_ato(path) -> AtoFile:
    """minitc (ato_read_and_save.py:64-700): the configuration, its atom types, ecore and
    dcore, the ran1 state and the molecule templates.  Lengths stay in Angstrom (the
    conversion to box fractions is `system_from_files`'s)."""
    L = _Lines(_read_text(path))
    head = L.line()
    w = words(head, 5)
    if len(w) != 3:
        cubic = False
        it = _tokens(head)
        nmol, simtk = fint(it[0]), f32(it[1])
        a = [f32(v) for v in L.items(3)]
        b = [f32(v) for v in L.items(3)]
        lattice = (a[0], a[1], a[2], b[0], b[1], b[2])
        dmax = F32(0)
    else:
        cubic = True
        it = _tokens(head)
        nmol, dmax, simtk = fint(it[0]), f32(it[1]), f32(it[2])
        lattice = (dmax, dmax, dmax, F32(90.0), F32(0.0), F32(0.0))
    line2 = L.line()
    w2 = words(line2, 5)
    v = [f32(t) for t in _tokens(line2)[:6]]
    tol, stepmi, stepri, stepr, stepm, tempb = v
    tempangreq, tempdihreq, tempangwt, tempdihwt = F32(1.0), F32(10.0), F32(1.0), F32(1.0)
    if len(w2) == 7:
        tempdihreq = f32(w2[6])
    if len(w2) > 7:
        tempangreq, tempdihreq = f32(w2[6]), f32(w2[7])
    if len(w2) > 9:
        tempangwt, tempdihwt = f32(w2[8]), f32(w2[9])
    mols = []
    for _ in range(nmol):
        text = L.line()
        wm = words(text, 10)
        t = _tokens(text)
        m = AtoMolecule(nato=fint(t[0]), centre=np.array([f32(x) for x in t[1:4]], dtype=F32),
                        phi=np.array([f32(x) for x in t[4:7]], dtype=F32))
        if len(wm) > 7:
            m.tethered = wm[7][:1] == "T"
            if len(wm[7]) > 1:
                try:
                    m.tetheredatom = int(wm[7][1:])
                except ValueError:
                    m.tetheredatom = 0
        m.nstepreject = fint(wm[8]) if len(wm) > 8 else 0
        m.scalefac = f32(wm[9]) if len(wm) > 9 else F32(1.0)
        xyz = np.zeros((m.nato, 3), dtype=F32)
        for ia in range(m.nato):
            wl = words(L.line(), 10)
            m.labels.append(fa(wl[0], 3))
            img = 0
            if len(wl) > 2:
                img = fint(wl[2])
                if img < 0 or img > m.nato:
                    img = 0
            m.imageatom.append(img)
            xyz[ia] = [f32(x) for x in L.items(3)]
            b = L.counted(2)               # the bond count and its (atom, length) pairs: one READ
            nb = fint(b[0])
            m.bonds.append([(fint(b[1 + 2 * j]), f32(b[2 + 2 * j])) for j in range(nb)])
        m.xyz = xyz
        ngrot = fint(L.items(1)[0])
        for _g in range(ngrot):
            kind = fa(L.line()[1:4], 3)
            ax = tuple(fint(x) for x in L.items(2))
            dih = []
            if kind == "DIH":
                nd = fint(L.items(1)[0])
                for _k in range(nd):
                    d = L.items(3)
                    dih.append((fint(d[0]), fint(d[1]), f32(d[2])))
            wr = words(L.line(), 5)
            nrot = fint(wr[0])
            atoms = [fint(x) for x in wr[1:]]
            if len(atoms) < nrot:
                atoms += [fint(x) for x in L.items(nrot - len(atoms))]
            m.rotgroups.append(RotGroup(kind=kind, axis=ax, dihedrals=dih, atoms=atoms[:max(nrot, len(atoms))]))
        mols.append(m)
    order = _type_order(mols)
    types = {lab: AtoType(label=lab) for lab in order}
    pi = F32(4.0) * F32(math.atan(1.0))
    for _ in range(len(order)):
        w1 = words(L.line(), 8)
        a1 = fa(w1[0], 3)
        ats = fa(w1[1], 3) if len(w1) > 1 else "XXX"
        iexch = fint(w1[2]) if len(w1) > 2 else -1
        text = L.line()
        sp = parse_words(text, 11)
        wv = [text[a:b + 1] for a, b in sp]
        eps, sig, mass, q, rq = [f32(x) for x in _tokens(text[sp[0][0]:sp[min(4, len(sp) - 1)][1] + 1])[:5]]
        natomsph = f32(wv[5]) if len(wv) > 5 else F32(1.0)
        if len(wv) > 7:
            radmin, radmax = f32(wv[6]), f32(wv[7])
            if radmax <= 0.0:
                radmax = F32(0.5) * sig
        else:
            radmin, radmax = F32(0.0), F32(0.5) * sig
        pseudo = F32(1.0)
        if len(wv) > 8:
            pseudo = f32(wv[8])
            if pseudo < 1.0:
                pseudo = F32(1.0)
        volsph = F32(4.0) * pi * (radmax * radmax * radmax - radmin * radmin * radmin) / F32(3.0)
        if abs(natomsph) > 0.0 and natomsph < 0.0:
            natomsph = abs(natomsph) * volsph / pseudo
        if a1 in types:
            types[a1] = AtoType(label=a1, atsymbol=ats, iexchange=iexch, eps=eps, sig=sig, mass=mass, charge=q,
                                qradius=rq, natomsphere=natomsph, radminsphere=radmin, radmaxsphere=radmax,
                                pseudofac=pseudo)
    ecore, dcore = [f32(x) for x in L.items(2)]
    ran1 = None
    try:
        r = L.items(35)
        if len(r) == 35:
            ran1 = np.array([fint(x) for x in r], dtype=np.int64).astype(np.int32)
    except (EOFError, ValueError):
        ran1 = None
    moltypes = []
    for m in mols:
        if m.labels[0] not in moltypes:
            moltypes.append(m.labels[0])
    templates = [AtoTemplate(name=f"moltype{i + 1:02d}", wholerot=stepr, wholetrans=stepm)
                 for i in range(len(moltypes))]
    if ran1 is not None:
        for _ in range(len(moltypes)):
            if L.at_end():
                break
            wt = words(L.line(), 11)
            if len(wt) <= 1:
                continue
            try:
                k = int(wt[0]) - 1
            except ValueError:
                break
            if not 0 <= k < len(moltypes):
                break
            vals = [F32(0)] * 5
            for j in range(5):
                if len(wt) > 2 + j:
                    vals[j] = f32(wt[2 + j])
            templates[k] = AtoTemplate(wt[1], *vals)
    return AtoFile(nmol=nmol, cubic=cubic, dmax=dmax, simtk=simtk, lattice=lattice, tol=tol, stepmi=stepmi,
                   stepri=stepri, stepr=stepr, stepm=stepm, tempb=tempb, tempangreq=tempangreq,
                   tempdihreq=tempdihreq, tempangwt=tempangwt, tempdihwt=tempdihwt, molecules=mols,
                   types=[types[lab] for lab in order], ecore=ecore, dcore=dcore, ran1=ran1,
                   moltypes=moltypes, templates=templates)


def _box_volume(a: AtoFile):
    alat, blat, clat, phib, thetac, phic = (F32(x) for x in a.lattice)
    if a.cubic:
        return alat * blat * clat
    rad = math.pi / 180.0
    blaty = blat * F32(math.sin(float(phib) * rad))
    clatz = clat * F32(math.cos(float(thetac) * rad))
    return alat * blaty * clatz


def ato_composition(a: AtoFile):
    """moloutc's closing block (ato_read_and_save.py:1318-1336): the number densities
    (coarse-grained, actual) and per type (cfraccg, cfrac) over the atoms that are not
    Q-type."""
    it = a.itype()
    qt = a.qtype()
    nt = len(a.types)
    cfrac = np.zeros(nt, dtype=F32)
    cfraccg = np.zeros(nt, dtype=F32)
    natomr = 0
    natomrcg = F32(0)
    for i in range(it.shape[0]):
        t = a.types[it[i]]
        w = t.natomsphere * t.pseudofac
        if qt[i] == 0:
            natomr += 1
            natomrcg = natomrcg + w
        cfrac[it[i]] = cfrac[it[i]] + F32(1.0)
        cfraccg[it[i]] = cfraccg[it[i]] + w
    vol = _box_volume(a)
    with np.errstate(divide="ignore", invalid="ignore"):
        rho = F32(natomr) / vol
        rhocg = natomrcg / vol
        if natomr > 0:
            cfrac = (cfrac / F32(natomr)).astype(F32)
            cfraccg = (cfraccg / natomrcg).astype(F32)
    return rhocg, rho, cfraccg, cfrac


def format_ato(a: AtoFile) -> str:
    """moloutc (ato_read_and_save.py:1103-1350), line for line."""
    out = []
    if a.cubic:
        out.append(" " + fi(a.nmol, 7) + " " + fe(a.dmax, 15, 8) + " " + fe(a.simtk, 15, 8))
    else:
        alat, blat, clat, phib, thetac, phic = a.lattice
        out.append(" " + fi(a.nmol, 7) + " " + fe(a.simtk, 15, 8))
        out.append(" " + "".join(" " + fe(x, 15, 8) for x in (alat, blat, clat)))
        out.append(" " + "".join(" " + fe(x, 15, 8) for x in (phib, thetac, phic)))
    out.append(" " + "".join(" " + fe(x, 12, 5) for x in (a.tol, a.stepmi, a.stepri, a.stepr, a.stepm, a.tempb,
                                                          a.tempangreq, a.tempdihreq, a.tempangwt, a.tempdihwt)))
    for im, m in enumerate(a.molecules, start=1):
        ta = m.tetheredatom if 1 <= m.tetheredatom <= m.nato else 0
        tt = ("T" + f"{ta:05d}") if m.tethered else "F"
        out.append(" " + fi(m.nato) + "".join(" " + fe(x, 15, 8) for x in (*m.centre, *m.phi)) + "  " + fa(tt, 6)
                   + " " + fi(m.nstepreject) + " " + ff(m.scalefac, 10, 5) + " " + fi(im, 5))
        for ia in range(m.nato):
            out.append(" " + fa(m.labels[ia], 3) + "  " + fi(ia + 1, 5) + "  " + fi(m.imageatom[ia], 5))
            out.append("".join(" " + ff(x, 14, 8) for x in m.xyz[ia]))
            bl = m.bonds[ia]
            first = " " + fi(len(bl), 4)
            recs = []
            cur = first
            for j, (an, ln) in enumerate(bl):
                if j > 0 and j % 5 == 0:
                    recs.append(cur)
                    cur = ""
                cur += " " + fi(an, 4) + " " + fe(ln, 9, 3)
            recs.append(cur)
            out.extend(recs)
        out.append(_f204([len(m.rotgroups)]))
        for g in m.rotgroups:
            out.append(" " + fa(g.kind, 3))
            out.append(_f204(list(g.axis)))
            if g.kind == "DIH":
                out.append(_f204([len(g.dihedrals)]))
                for d in g.dihedrals:
                    out.append(" " + fi(d[0], 4) + " " + fi(d[1], 4) + " " + ff(d[2], 10, 3))
            out.append(_f204([len(g.atoms)] + list(g.atoms)))
    for t in a.types:
        out.append(" " + fa(t.label, 3) + " " + fa(t.atsymbol, 3) + " " + fi(t.iexchange, 2))
        out.append("".join(" " + ff(x, 10, 6) for x in (t.eps, t.sig, t.mass, t.charge, t.qradius))
                   + " " + fe(t.natomsphere, 13, 6) + " " + ff(t.radminsphere, 10, 4) + " "
                   + ff(t.radmaxsphere, 10, 4) + " " + fe(t.pseudofac, 13, 6))
    out.append("".join(" " + ff(x, 14, 8) for x in (a.ecore, a.dcore)))
    if a.ran1 is not None:
        out.append(lw_line([int(v) for v in a.ran1]))
    written = []
    mt = a.mtype()
    for im in range(a.nmol):
        k = int(mt[im])
        if k in written:
            continue
        written.append(k)
        tp = a.templates[k]
        name = tp.name.split(" ")[0]
        out.append(" " + fi(len(written), 3) + " " + name + "".join(
            " " + fe(x, 13, 6) for x in (tp.interntrans, tp.internrot, tp.wholerot, tp.wholetrans))
            + " " + ff(tp.rfract, 10, 6))
    rhocg, rho, cfraccg, cfrac = ato_composition(a)
    out.append(" Atomic number density (coarse grain and actual) = " + " " + fe(rhocg, 14, 7) + " " + fe(rho, 14, 7))
    for i, t in enumerate(a.types):
        if t.label[:1] in ("Q", "q"):
            continue
        out.append(" " + fa(t.atsymbol, 3) + "  " + fi(0, 1) + "  " + fe(cfraccg[i], 12, 5) + " " * 10
                   + fe(cfrac[i], 12, 5) + "  " + fa(t.label, 3))
    return "\n".join(out) + "\n"


def _f204(items):
    """(1x,14(i4,1x)): the group's trailing blank is positioning only and never written;
    after 14 items the format reverts to the group on a new record."""
    recs = []
    cur = " "
    for j, v in enumerate(items):
        if j > 0 and j % 14 == 0:
            recs.append(cur[:-1])
            cur = ""
        cur += fi(v, 4) + " "
    recs.append(cur[:-1])
    return "\n".join(recs)


def write_ato(path, a: AtoFile):
    with open(path, "w", newline="") as f:
        f.write(format_ato(a))


# ------------------------------------------------------------------ keyword files


@dataclasses.dataclass
class Keyword:
    name: str
    value: str                 # the value words as they stand in the file
    text: str = ""             # the help text after them


@dataclasses.dataclass
class KeywordFile:
    """A `.EPSR.inp`-style file: a title line, keyword lines (`(a11,1x,a,15x,a)`), and
    per-dataset blocks opened by `data N`; `q` ends it."""

    title: str
    keywords: list             # [Keyword]
    blocks: list               # [[Keyword]] one per `data N` block
    block_word: str = "data"

    def get(self, name, default=None, block=None):
        src = self.keywords if block is None else self.blocks[block]
        for k in src:
            if k.name == name:
                return k.value
        return default

    def real(self, name, default=None, word=0, block=None):
        v = self.get(name, None, block)
        if v is None:
            return default
        return f32(v.split()[word])

    def int(self, name, default=None, word=0, block=None):
        v = self.get(name, None, block)
        if v is None:
            return default
        return fint(v.split()[word])

    def set(self, name, value: str, block=None):
        src = self.keywords if block is None else self.blocks[block]
        for k in src:
            if k.name == name:
                k.value = value
                return
        src.append(Keyword(name, value, ""))


def _keyword_line(text: str, nspc: int):
    sp = parse_words(text, nspc)
    if not sp:
        return None
    name = text[sp[0][0]:sp[0][1] + 1]
    if len(sp) == 1:
        return Keyword(name, "", "")
    a, b = sp[1][0], sp[-1][1]
    return Keyword(name, text[a:b + 1], text[b + 1:].strip(" "))


def read_keywords(path, nspc: int = 12) -> KeywordFile:
    """Keyword lines as save_files writes them and read_input_file reads them
    (read_and_save_files.py:103-190, 558-640; words split as its `_parse(12)` splits
    them): the first word is the keyword, the following words (blank runs of at most
    `nspc`) its value, the rest free text."""
    lines = _read_text(path).splitlines()
    title = lines[0] if lines else ""
    kws, blocks = [], []
    cur = kws
    for text in lines[1:]:
        if not text.strip():
            continue
        k = _keyword_line(text, nspc)
        if k.name == "q" and not k.value:
            break
        if k.name == "data" and k.value.strip().isdigit():
            blocks.append([])
            cur = blocks[-1]
            continue
        cur.append(k)
    return KeywordFile(title=title, keywords=kws, blocks=blocks)


def _format_keyword(k: Keyword) -> str:
    return fa(k.name, 11) + " " + k.value + " " * 15 + fa(k.text, 256)


def format_keywords(kf: KeywordFile, title_suffix=True) -> str:
    out = [kf.title]
    out += [_format_keyword(k) for k in kf.keywords]
    for i, b in enumerate(kf.blocks, start=1):
        out += ["", kf.block_word + " " + fi(i, 3), ""]
        out += [_format_keyword(k) for k in b]
    out.append("q")
    return "\n".join(out) + "\n"


def write_keywords(path, kf: KeywordFile):
    with open(path, "w", newline="") as f:
        f.write(format_keywords(kf))


# ------------------------------------------------------------------ .pcof


@dataclasses.dataclass
class PcofPair:
    """One pair potential of a .pcof file (read_pcof :1080-1262)."""

    tlab1: str
    tlab2: str
    coulfac: np.float32
    rminex: np.float32
    pcoefex: np.float32
    coordex: np.float32
    rmaxex: np.float32
    pcoefexmax: np.float32
    gaussians: list            # [(rp, wp, hp)]
    pcoefp: np.ndarray         # float32[ncoeffp]: pcoefp(1..ncoeffp)


@dataclasses.dataclass
class PcofFile:
    head: KeywordFile          # the parameters before `q`
    pairs: list
    tail: list                 # the lines after the pairs (three-body, coordination constraints), verbatim

    @property
    def ncoeffp(self):
        n = self.head.int("ncoeffp", 0)
        return n if n > 0 or not self.pairs else len(self.pairs[0].pcoefp)


def read_pcof(path) -> PcofFile:
    """read_pcof (read_and_save_files.py:885-1300) for a file that exists: the keyword
    head up to `q`, the number of pairs, then per pair its label line, the rminex /
    pcoefex / coordex [/ rmaxex / pcoefexmax] [/ Gaussians] line and ncoeffp
    coefficients (one list-directed READ).  ireset handling belongs to
    `system_from_files`."""
    L = _Lines(_read_text(path))
    kws = []
    while True:
        text = L.line()
        k = _keyword_line(text, 11)
        if k is None:
            continue
        if k.name == "q":
            break
        kws.append(k)
    head = KeywordFile(title=None, keywords=kws, blocks=[])
    ncoeffp = head.int("ncoeffp", 0)
    npairs = fint(L.items(1)[0])
    pairs = []
    for _ in range(npairs):
        wl = words(L.line(), 11)
        coulfac = f32(wl[3]) if len(wl) > 3 else F32(1.0)
        wv = words(L.line(), 11)
        nwords = len(wv)
        rminex = f32(wv[0])
        pcoefex = f32(wv[1])
        if rminex == 0.0:
            pcoefex = F32(0.0)
        iword = 2
        coordex = F32(0.0)
        if iword < nwords:
            coordex = f32(wv[iword])
            iword += 1
        rmaxex = pcoefexmax = F32(0.0)
        npg = (nwords - iword) // 3
        if nwords - iword - 3 * npg == 2:
            rmaxex, pcoefexmax = f32(wv[iword]), f32(wv[iword + 1])
            iword += 2
            if rmaxex == 0.0:
                pcoefexmax = F32(0.0)
        gauss = []
        while iword + 2 < nwords:
            gauss.append((f32(wv[iword]), f32(wv[iword + 1]), f32(wv[iword + 2])))
            iword += 3
        if abs(pcoefex) < F32(1.0e-10):
            pcoefex = F32(0.0)
        # a head without ncoeffp (the reference then takes its default from the run's
        # symbol table) gives the count of the coefficient record itself
        coef = np.array([f32(x) for x in (L.items(ncoeffp) if ncoeffp > 0 else _tokens(L.line()))], dtype=F32)
        pairs.append(PcofPair(tlab1=fa(wl[0], 3), tlab2=fa(wl[1], 3), coulfac=coulfac, rminex=rminex,
                              pcoefex=pcoefex, coordex=coordex, rmaxex=rmaxex, pcoefexmax=pcoefexmax,
                              gaussians=gauss, pcoefp=coef))
    tail = []
    while not L.at_end():
        tail.append(L.line())
    return PcofFile(head=head, pairs=pairs, tail=tail)


def format_pcof(p: PcofFile) -> str:
    """The .pcof half of save_files_others (read_and_save_files.py:645-760)."""
    out = [_format_keyword(k) for k in p.head.keywords]
    out.append("q")
    out.append(lw_int(len(p.pairs)))
    for ip, pr in enumerate(p.pairs, start=1):
        out.append(fa(pr.tlab1, 3) + " " + fa(pr.tlab2, 3) + " " + " " * 5 + fi(ip) + "  " + ff(pr.coulfac, 10, 5))
        if not pr.gaussians:
            if pr.rmaxex <= 0.0:
                out.append(lw_line([pr.rminex, pr.pcoefex, pr.coordex]))
            else:
                out.append(lw_line([pr.rminex, pr.pcoefex, pr.coordex, pr.rmaxex, pr.pcoefexmax]))
        else:
            head = [pr.rminex, pr.pcoefex, pr.coordex] + ([pr.rmaxex, pr.pcoefexmax] if pr.rmaxex > 0.0 else [])
            s = "".join(" " + ff(x, 10, 5) for x in head)
            s += "".join(" " + ff(g[0], 9, 4) + " " + ff(g[1], 9, 4) + " " + fe(g[2], 11, 4) for g in pr.gaussians)
            out.append(s)
        out.append(lw_line([F32(x) for x in pr.pcoefp]))
    out += p.tail
    return "\n".join(out) + "\n"


def write_pcof(path, p: PcofFile):
    with open(path, "w", newline="") as f:
        f.write(format_pcof(p))


# ------------------------------------------------------------------ .wts (neutron weights)


@dataclasses.dataclass
class XrayRecord:
    """One atom type's X-ray form factor: cxt + sum_i axt_i exp(-bxt_i (Q/4 pi)^2), with the
    modification alphaxt, deltaxt (get_wts/_read_xray_coeffs, epsr_standalone_rev3.py:
    4285-4296, 4517-4537); a1 the type label (character*3), a2 the element read beside it."""

    a1: str
    a2: str
    cxt: float
    axt: tuple                 # 5 doubles
    bxt: tuple                 # 5 doubles
    alphaxt: np.float32
    deltaxt: np.float32


@dataclasses.dataclass
class WtsFile:
    """A weights file (get_wts, epsr_standalone_rev3.py:4180-4330).  Neutron: the density
    factor, then per type pair its labels and the intramolecular and intermolecular weights
    (the rest of each value record is not read).  X-ray (a file name containing
    `X-ray.wts`, `x-ray.wts` or `XWTS.wts`): the normalisation line -- rhofac, normx,
    iafformaff, or one old-style value they are deduced from (:4240-4266) -- then per atom
    type its labels and form-factor coefficients."""

    rhofac: np.float32
    pairs: list                # [(a1, a2, w_intra, w_inter)] in file order, labels character*3
    xray: bool = False
    normx: int = 0
    iafformaff: int = 0
    records: list = dataclasses.field(default_factory=list)   # [XrayRecord]


def is_xray_wts(path) -> bool:
    return any(k in str(path) for k in ("X-ray.wts", "x-ray.wts", "XWTS.wts"))


def _read_xray_wts(L: _Lines) -> WtsFile:
    first = L.line()
    w = words(first, 5)
    if len(w) == 3:
        rhofac, normx, iafformaff = f32(w[0]), fint(w[1]), fint(w[2])
    else:
        # the old single value: its sign gives per-atom (1) or per-molecule (2) data, its
        # size the form factors and the normalisation (normx 1 <-> 2 in the new numbering)
        rhofac = f32(w[0])
        nrho = 2 if rhofac < 0 else 1
        iafformaff = int(np.trunc(F32(F32(abs(rhofac)) + F32(1)) / F32(2)))
        normx = int(np.trunc(F32(F32(abs(rhofac)) - F32(2 * (iafformaff - 1)))))
        normx = 2 if normx == 1 else 1
        rhofac = f32(nrho)
    recs = []
    while not L.at_end():
        lab = L.line()
        if not lab.strip():
            continue
        a1 = fa(lab[1:4], 3)                               # (1x,a3,1x,a4): a2 keeps the
        a2 = fa((lab[5:9] + "    ")[:4][1:], 3)            # rightmost 3 of its 4 characters
        try:
            v = L.items(13)
        except EOFError:
            break
        if len(v) < 13:
            break
        num = [float(t.replace("d", "e").replace("D", "e")) for t in v[:11]]
        recs.append(XrayRecord(a1=a1, a2=a2, cxt=num[0], axt=tuple(num[1:11:2]), bxt=tuple(num[2:11:2]),
                               alphaxt=f32(v[11]), deltaxt=f32(v[12])))
    return WtsFile(rhofac=rhofac, pairs=[], xray=True, normx=normx, iafformaff=iafformaff, records=recs)


def read_wts(path) -> WtsFile:
    L = _Lines(_read_text(path))
    if is_xray_wts(path):
        return _read_xray_wts(L)
    rhofac = f32(L.items(1)[0])
    pairs = []
    while not L.at_end():
        lab = L.line()
        if not lab.strip():
            continue
        a1, a2 = fa(lab[1:4], 3), fa(lab[5:8], 3)          # (1x,a3,1x,a3)
        try:
            w = L.items(2)
        except EOFError:
            break
        pairs.append((a1, a2, f32(w[0]), f32(w[1])))
    return WtsFile(rhofac=rhofac, pairs=pairs)


# ------------------------------------------------------------------ the box


def system_from_files(ato, pcof=None, inp=None, name=None):
    """The GPU path's box (`synth.System`) from an .ato file, the potential coefficients of a
    .pcof file and the run parameters of an .EPSR.inp file (paths or parsed objects).

    Coordinates in box fractions as normalise_coordinates leaves them (centres / dmax,
    atom offsets / (dmax * sizefactor), float32 divisions); types numbered by first
    occurrence (ftype); pair parameters as minitc forms them (ecor = sqrt(eps_i eps_j),
    dcor = (sig_i + sig_j)/2, qcor = q_i q_j 1390); EP coefficients per type pair matched
    by label in either order (read_pcof :1263-1290), with ireset's reset applied;
    rmaxgr <= 0 or beyond half the box becomes half the box (epsr_standalone_rev3.py:2938-2941).

    The box carries the file's intramolecular topology (bonds and angle constraints, image
    atoms, rotational groups, dihedrals, masses: intra.from_ato) and MC run parameters
    (stepr/stepm/stepri/stepmi normalised as normalise_coordinates does, ato_read_and_save.py:
    916-940; tempangwt/tempdihwt; per-template step sizes; per-molecule nstepreject), so the
    intramolecular energies (en_mole_intra) and moves run on the GPU; eintra is None.

    The GPU path is cubic: a non-cubic box, Q-type atoms, tethered molecules or an
    intramolecular Coulomb weighting (ref_intra's second word > 0) are rejected.
    """
    from . import synth

    a = ato if isinstance(ato, AtoFile) else read_ato(ato)
    pc = pcof if (pcof is None or isinstance(pcof, PcofFile)) else read_pcof(pcof)
    kf = inp if (inp is None or isinstance(inp, KeywordFile)) else read_keywords(inp)
    if not a.cubic:
        raise ValueError("system_from_files: the GPU path handles cubic boxes only")
    if a.qtype().any():
        raise ValueError("system_from_files: Q-type atoms are not handled by the GPU path")
    if any(m.tethered for m in a.molecules):
        raise ValueError("system_from_files: tethered molecules are not handled by the GPU path")
    refine_intra_fac = F32(0.0)
    if kf is not None and kf.get("ref_intra") is not None:
        w = kf.get("ref_intra").split()
        refine_intra_fac = f32(w[0]) if w else F32(0.0)
        if len(w) > 1 and f32(w[1]) > 0.0:
            raise NotImplementedError("system_from_files: ref_intra's Coulomb weighting (coul_intra) needs the "
                                      "intramolecular reference potential xav, which this package does not build")
    dmax = F32(a.dmax)
    sizefactor = kf.real("sizefactor", F32(1.0)) if kf is not None else F32(1.0)
    dmaxmol = F32(dmax * sizefactor)
    itype = a.itype()
    xa, ya, za, naf, nal = [], [], [], [], []
    xadd = np.empty(a.nmol, F32)
    yadd = np.empty(a.nmol, F32)
    zadd = np.empty(a.nmol, F32)
    atom = 0
    for im, m in enumerate(a.molecules):
        xadd[im], yadd[im], zadd[im] = (F32(c) / dmax for c in m.centre)
        off = (m.xyz / dmaxmol).astype(F32)
        xa.append(off[:, 0])
        ya.append(off[:, 1])
        za.append(off[:, 2])
        naf.append(atom)
        nal.append(atom + m.nato - 1)
        atom += m.nato
    nt = len(a.types)
    eps = np.array([t.eps for t in a.types], F32)
    sig = np.array([t.sig for t in a.types], F32)
    charge = np.array([t.charge for t in a.types], F32)
    npairref = np.zeros((nt, nt), np.int32)
    n1, n2 = [], []
    ip = 0
    for it in range(nt):
        for jt in range(it, nt):
            npairref[it, jt] = npairref[jt, it] = ip
            n1.append(it)
            n2.append(jt)
            ip += 1
    n1 = np.array(n1, np.int32)
    n2 = np.array(n2, np.int32)
    ecor = np.sqrt(eps[n1] * eps[n2]).astype(F32)
    dcor = (F32(0.5) * (sig[n1] + sig[n2])).astype(F32)
    qcor = (charge[n1] * charge[n2] * F32(1390)).astype(F32)
    _, _, cfraccg, cfrac = ato_composition(a)
    kw = {"cfraccg": cfraccg}
    npairs = ip
    if pc is not None:
        ncoeffp = pc.ncoeffp
        ireset = kf.int("ireset", 0) if kf is not None else 0
        pcoefp = np.zeros((npairs, ncoeffp + 1), F32)
        rminex = np.zeros(npairs, F32)
        pcoefex = np.zeros(npairs, F32)
        coulfac = np.ones(npairs, F32)
        rminfac = pc.head.real("rminfac", F32(0.0))
        roverlap = pc.head.real("roverlap", F32(0.0))
        for p in range(npairs):
            la, lb = a.types[n1[p]].label, a.types[n2[p]].label
            found = None
            for pr in pc.pairs:
                if (pr.tlab1 == la and pr.tlab2 == lb) or (pr.tlab1 == lb and pr.tlab2 == la):
                    found = pr
                    break
            if found is None:
                # a pair the file lacks gets new coefficients and the minimum distance
                # rminfac * dcor, at least roverlap (read_and_save_files.py:1297-1319)
                pcoefp[p, 1] = 1.0
                r0 = F32(F32(rminfac) * dcor[p])
                rminex[p] = r0 if r0 >= F32(roverlap) else F32(roverlap)
                continue
            if any(abs(g[1]) > 0.0 for g in found.gaussians):
                raise NotImplementedError("system_from_files: Gaussian or modified-Morse terms of the reference "
                                          "potential (.pcof) are not built by this package")
            if found.rmaxex > 0.0:
                raise NotImplementedError("system_from_files: container forces (rmaxex > 0 in the .pcof) are not "
                                          "handled by this package")
            coulfac[p] = found.coulfac
            pcoefp[p, 1:] = found.pcoefp[:ncoeffp]
            if ireset != 0:
                if ireset < 2:
                    pcoefp[p, 1] = 1.0
                pcoefp[p, 2:] = 0.0
            efac = pcoefp[p, 1]
            if any(g[0] < 0.0 or g[1] < 0.0 for g in found.gaussians):
                efac = F32(0.0)
            pcoefp[p, 1] = efac
            rminex[p] = found.rminex
            pcoefex[p] = F32(0.0) if ireset == 2 else found.pcoefex
        for key, kname in (("rminpt", "rminpt"), ("rmaxpt", "rmaxpt"), ("power", "power"),
                           ("psigma2", "psigma2"), ("pdstep", "pdstep")):
            v = pc.head.real(kname, None)
            if v is not None:
                kw[key] = float(v)
        kw["rminex"] = rminex
        kw["pcoefex"] = pcoefex
        kw["coulfac"] = coulfac
        for key in ("refpotfac", "expecf"):
            v = pc.head.real(key, None)
            if v is not None:
                kw[key] = float(v)
        ex = pc.head.get("expecf")
        if ex is not None and len(ex.split()) > 1:
            kw["expecfpwr"] = float(f32(ex.split()[1]))
        rpt = pc.head.get("reppottype")
        if rpt is not None and rpt.split() and rpt.split()[0].lower().startswith("harm"):
            raise NotImplementedError("system_from_files: the harmonic repulsive core is not built by this package")
    else:
        pcoefp = np.zeros((npairs, 2), F32)
        pcoefp[:, 1] = 1.0
    cellst, rmaxgr, qstep, nq = F32(0.03), F32(0.0), F32(0.05), 600
    if kf is not None:
        cellst = kf.real("cellst", cellst)
        rmaxgr = kf.real("rmaxgr", rmaxgr)
        qstep = kf.real("qstep", qstep)
        nq = kf.int("nq", nq)
        for key, kname, conv in (("ntimes", "ntimes", int), ("nmolcell", "nmolcell", int),
                                 ("ngrsamples", "ngrsamples", int), ("rejrate", "rejrate", float)):
            v = kf.get(kname)
            if v is not None:
                kw[key] = conv(fint(v.split()[0]) if conv is int else float(f32(v.split()[0])))
    if rmaxgr > F32(0.5) * dmax or rmaxgr <= 0.0:
        rmaxgr = F32(0.5) * dmax
    sysname = name or (a.templates[0].name if a.templates else "ato")
    from . import intra as IT
    amass = np.array([t.mass for t in a.types], F32)
    topo = IT.from_ato(a, itype, npairref, amass, dmaxmol, refine_intra_fac)
    stepm = F32(a.stepm) / dmax
    if stepm > 0.25:
        stepm = F32(0.25)
    tpl = a.templates
    mc_run = dict(stepr=F32(a.stepr), stepm=F32(stepm), stepri=F32(a.stepri), stepmi=F32(a.stepmi) / dmaxmol,
                  tempangwt=F32(a.tempangwt), tempdihwt=F32(a.tempdihwt), tempangreq=F32(a.tempangreq),
                  tempdihreq=F32(a.tempdihreq),
                  wholerot=np.array([t.wholerot for t in tpl], F32), wholetrans=np.array([t.wholetrans for t in tpl], F32),
                  internrot=np.array([t.internrot for t in tpl], F32),
                  interntrans=np.array([t.interntrans for t in tpl], F32),
                  nstepreject=np.array([m.nstepreject for m in a.molecules], np.int32),
                  intra=kf.int("intra", 100) if kf is not None else 100,
                  rotfreq=kf.int("rotfreq", 5) if kf is not None else 5,
                  iinit=kf.int("iinit", 0) if kf is not None else 0)
    return synth.System(
        name=sysname, L=float(dmax), type_names=[t.label.strip() for t in a.types],
        xa=np.ascontiguousarray(np.concatenate(xa)), ya=np.ascontiguousarray(np.concatenate(ya)),
        za=np.ascontiguousarray(np.concatenate(za)), itype=itype, xadd=xadd, yadd=yadd, zadd=zadd,
        naf=np.array(naf, np.int32), nal=np.array(nal, np.int32), mtype=a.mtype(),
        mol_names=[t.name for t in a.templates], eintra=None, npairref=npairref,
        n1extr=n1, n2extr=n2, cfrac=cfrac, eps=eps, sig=sig, charge=charge, ecor=ecor, dcor=dcor, qcor=qcor,
        pcoefp=pcoefp, cellst=float(cellst), rmaxgr=float(rmaxgr), qstep=float(qstep), nq=int(nq),
        simtk=float(a.simtk), intra=topo, mc_run=mc_run,
        qradius=np.array([t.qradius for t in a.types], F32), **kw)


def update_ato(a: AtoFile, xadd, yadd, zadd, xa, ya, za, ran1=None, sizefactor=F32(1.0)) -> AtoFile:
    """A copy of `a` with the box's current coordinates (box fractions, as the GPU path
    holds them) turned back into Angstrom the way unnormalise_coordinates does
    (ato_read_and_save.py:1001-1080: value * dmax, float32) and, if given, the ran1 state."""
    b = dataclasses.replace(a, molecules=[dataclasses.replace(m) for m in a.molecules])
    dmax = F32(a.dmax)
    dmaxmol = F32(dmax * F32(sizefactor))
    atom = 0
    for im, m in enumerate(b.molecules):
        m.centre = np.array([F32(xadd[im]) * dmax, F32(yadd[im]) * dmax, F32(zadd[im]) * dmax], F32)
        sl = slice(atom, atom + m.nato)
        m.xyz = np.stack([np.asarray(xa[sl], F32) * dmaxmol, np.asarray(ya[sl], F32) * dmaxmol,
                          np.asarray(za[sl], F32) * dmaxmol], axis=1).astype(F32)
        atom += m.nato
    if ran1 is not None:
        b.ran1 = np.asarray(ran1, np.int32).copy()
    return b
