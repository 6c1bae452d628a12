"""The refinement-run COMMON data (src/epsr26/common/epsr16.py) the hot path reads: the
r grid (x, xstar), the histogram and Q grids, the foqols transform, the datasets'
Q ranges and the Metropolis rejection target."""

from ._arrays import integer, real
import numpy as np

F32 = np.float32


def reset(n_pt=0, n_q=0, n_file=0, n_pt1=0):
    global x, xstar, ycoh, q, foqols, nqmin, nqmax, sigma2
    global npt, npt1, npt2, nptmin, nptmax, nq, nfile, cellst, rho, rminpt, rmaxpt, rmaxgr
    global rejrate, qminp, qminm, nmolcell
    x, xstar, ycoh = real(n_pt), real(n_pt), real(n_pt)
    q = real(n_q)
    foqols = real(n_pt1, n_q)          # foqols(j, i)
    nqmin, nqmax = integer(n_file), integer(n_file)
    sigma2 = real(n_file)
    npt, npt1, npt2, nptmin, nptmax, nq, nfile = n_pt, n_pt1, 0, 0, 0, n_q, n_file
    cellst = rho = rminpt = rmaxpt = rmaxgr = F32(0.0)
    rejrate = F32(0.75)
    qminp = qminm = F32(0.0)
    nmolcell = 1


reset()
