"""The atom-level COMMON data the hot path reads (src/epsr26/common/molecom.py):
coordinates relative to the molecule centre, atom types, the pair-type table and the
reference-potential parameters.  Filled by common.setup; sized to the box."""

from ._arrays import integer, real
import numpy as np

F32 = np.float32


def reset(n_atom=0, n_type=0, n_pairs=0):
    global xa, ya, za, itype, npairref, n1extr, n2extr, cfrac, charge, qradius
    global ecor, dcor, qcor, natom, natomr, ntype, npairs, sizefactor, power, simt
    global stepr, stepm, stepri, stepmi, tol, tolout, xs, ys, zs
    global tempb, tempangwt, tempdihwt, tempangreq, tempdihreq
    xa, ya, za = real(n_atom), real(n_atom), real(n_atom)
    itype = integer(n_atom)
    npairref = integer(n_type, n_type)
    n1extr, n2extr = integer(n_pairs), integer(n_pairs)
    cfrac, charge, qradius = real(n_type), real(n_type), real(n_type)
    ecor, dcor, qcor = real(n_pairs), real(n_pairs), real(n_pairs)
    natom, ntype, npairs = n_atom, n_type, n_pairs
    natomr = F32(n_atom)
    sizefactor = F32(1.0)
    power = F32(12.0)
    simt = F32(0.0)
    stepr = stepm = F32(1.0)
    stepri = stepmi = F32(0.0)
    tol = tolout = F32(0.0)
    xs = ys = zs = F32(0.0)
    tempb = F32(0.0)
    tempangwt = tempangreq = F32(1.0)
    tempdihwt = tempdihreq = F32(0.0)


reset()
