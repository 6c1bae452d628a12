"""The molecule-level COMMON data (src/epsr26/common/clustercom.py): centres, atom ranges,
molecule types, box size and the MC step sizes and counters."""

from ._arrays import farray, integer, real
import numpy as np

F32 = np.float32


def reset(n_mol=0, n_moltype=0):
    global xadd, yadd, zadd, naf, nal, nato, mtype, nmol, nmoltype, dmax
    global wholerot, wholetrans, nwhole, nmovem, nstepreject, ngrot, internrot, interntrans, ninter, nmovei
    global tethered, tetheredatom
    xadd, yadd, zadd = real(n_mol), real(n_mol), real(n_mol)
    naf, nal, nato, mtype = integer(n_mol), integer(n_mol), integer(n_mol), integer(n_mol)
    ngrot = integer(n_mol)
    nstepreject = integer(n_mol)
    # molecules tethered to a point (the GPU path moves none of them: the drop-in refuses)
    tethered, tetheredatom = farray(np.bool_, n_mol), integer(n_mol)
    nmol, nmoltype = n_mol, n_moltype
    dmax = F32(0.0)
    wholerot, wholetrans = real(n_moltype), real(n_moltype)
    internrot, interntrans = real(n_moltype), real(n_moltype)
    nwhole, nmovem = integer(2, n_moltype), integer(2, n_moltype)
    ninter, nmovei = integer(2, n_moltype), integer(2, n_moltype)


reset()
