"""Potential COMMON data (src/epsr26/common/potential_par.py): the tables looked up by the
energy routines, the close-contact counters of atogen and the empirical-potential
coefficients and fit parameters."""

from ._arrays import integer, real
import numpy as np

F32 = np.float32


def reset(npt=0, npairs=0, nterms=0, nref=0, nfile=0):
    global xavs, xavs1, ecoul, ecoul1, xavtot, xavtot1, epot, epot1, wfac, wfacc
    global potex, potexmax, irefpcof, rminex, rmaxex, efacp, pcoefp, pcoefd, npcoeff, pstepd
    global pdmin, pdstep, pdmax, psigma2, rbroad, paccept, potfac, feedback, gaussian, npits, npitss
    global expsq, pwt, pwtold, absolute_energy, erequnit, ereqstep, pressfac, energabs
    xavs, xavs1, ecoul, ecoul1 = real(npt, npairs), real(npt, npairs), real(npt, npairs), real(npt, npairs)
    xavtot, xavtot1, epot, epot1 = real(npt, npairs), real(npt, npairs), real(npt, npairs), real(npt, npairs)
    wfac, wfacc = real(npt), real(npt)
    potex, potexmax = real(npairs), real(npairs)
    irefpcof = integer(npairs)
    rminex, rmaxex = real(npairs), real(npairs)
    efacp = real(nref)
    pcoefp = real(nterms, nref)       # pcoefp(ic, irefp)
    pcoefd = real(nterms, nfile)      # pcoefd(ic, id)
    npcoeff = integer(nfile)
    pstepd = real(nfile)
    expsq = None
    pwt, pwtold = real(nterms), real(nterms)
    pdmin, pdstep, pdmax, psigma2, rbroad = F32(0.0), F32(0.5), F32(12.0), F32(0.3), F32(0.0)
    paccept, potfac, feedback = F32(0.05), F32(1.0), F32(1.0)
    gaussian = False
    npits = npitss = 1
    absolute_energy, erequnit, ereqstep = F32(10.0), F32(0.0), F32(0.0)
    pressfac, energabs = F32(1.0), F32(0.0)


reset()
