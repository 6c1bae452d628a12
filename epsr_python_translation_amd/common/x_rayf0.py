"""X-ray weights (src/epsr26/common/x_rayf0.py): the Q-dependent pair weights weightx(iq,
ip, ix) of X-ray file ix (float64, scaled by the concentrations as get_wts leaves them) and
the self-scattering shift backx(iq, ix)."""

from ._arrays import farray

import numpy as np

MXFILE = 2


def reset(nq=0, npairs=0):
    global weightx, backx
    weightx = farray(np.float64, nq, npairs, MXFILE)
    backx = farray(np.float64, nq, MXFILE)


reset()
