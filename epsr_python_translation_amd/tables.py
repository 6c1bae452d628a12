"""Radial and Q grids of the refinement loop (host side, exact float32 recurrences).

These are the serial setup recurrences of inita/initpt; everything data-parallel built
on them (foqols, truncation functions, potential tables) runs on the GPU.

References (src/epsr26/programs/epsr_standalone_rev3.py): radial grid :2947-2999,
uniform distribution :5391-5419, Q grid :3025-3026, xstar :3748-3749.
"""

from __future__ import annotations

import dataclasses

import numpy as np

F32 = np.float32
MPZ = 3000   # epsr_dimension.mpz (src/epsr26/common/epsr_dimension.py:13-21)


@dataclasses.dataclass
class Grid:
    x: np.ndarray        # [npt+1], x[1] = 0, Angstrom
    xstar: np.ndarray    # [npt+1], x / dmax
    ycoh: np.ndarray     # [npt+1]
    r2: np.ndarray       # [npt1+1], right edges r2 used by the foqols recurrence
    q: np.ndarray        # [nq+1]
    npt: int
    npt1: int
    npt2: int
    nptmin: int
    nptmax: int
    nq: int

    @property
    def ld(self):
        """leading dimension of grid-indexed tables"""
        return self.npt + 1


def inita_grid(cellst, rmaxgr, rminpt, rmaxpt, mpz=MPZ):
    """r grid of inita (:2947-2999); returns x (length mpz+1) and the index limits."""
    x = np.zeros(mpz + 1, dtype=F32)
    celstt = F32(cellst)
    rmaxgr = F32(rmaxgr)
    rminpt = F32(rminpt)
    rmaxpt = F32(rmaxpt)
    x1dold = F32(-0.5) * celstt
    cellhf = F32(0.25) * celstt
    x[1] = 0.0
    j = 2
    npt1 = npt2 = nptmin = nptmax = 2
    x1d = F32(x1dold + celstt)
    while j < mpz and x1d < rmaxgr:
        x[j] = x1d - cellhf
        x1dold = x1d
        cellhf = F32(0.5) * celstt
        if x1d < rminpt:
            nptmin = j + 1
        if x1d < rmaxpt:
            nptmax = j + 1
        if x1d < rmaxgr:
            npt1 = j
        j += 1
        x1d = F32(x1dold + celstt)
    x[j] = x1d - cellhf
    npt = j
    npt2 = j - 2
    npt1 = min(npt1, npt2)
    return x[:npt + 1].copy(), npt, npt1, npt2, nptmin, nptmax


def uniform(xst, rho, rmaxgr, ndim):
    """uniform (:5391-5419): atoms per shell of a uniform fluid of density rho."""
    ycoh = np.zeros(ndim + 1, dtype=F32)
    pi = F32(np.pi)
    rhofac = F32(F32(F32(4.0) * pi) * F32(rho)) / F32(3.0)
    x1 = F32(0.0)
    x13 = F32(0.0)
    xst = F32(xst)
    xstep = xst * F32(0.5)
    x2 = F32(x1 + xstep)
    j = 1
    while x2 < F32(rmaxgr) and j < ndim:
        j += 1
        x2 = F32(x1 + xstep)
        x23 = F32(F32(x2 * x2) * x2)
        ycoh[j] = rhofac * F32(x23 - x13)
        xstep = xst
        x1 = x2
        x13 = x23
    return ycoh


def r2_chain(x, npt1):
    """r2(j) = x(j) + (x(j) - r1), r1 <- r2: the bin edges of the foqols loop (:3045-3052)."""
    r2 = np.zeros(npt1 + 1, dtype=F32)
    r1 = F32(0.0)
    for j in range(1, npt1 + 1):
        rr = x[j]
        rstep = F32(rr - r1)
        r2[j] = F32(rr + rstep)
        r1 = r2[j]
    return r2


def make_grid(system, mpz=MPZ):
    """All grids for a synthetic System (Angstrom units; rho from the box)."""
    x, npt, npt1, npt2, nptmin, nptmax = inita_grid(system.cellst, system.rmaxgr, system.rminpt,
                                                    system.rmaxpt, mpz)
    # xstar = x / dmax (initpt, :3748-3749)
    xstar = (x / F32(system.dmax)).astype(F32)
    ycoh = uniform(system.cellst, system.rho, system.rmaxgr, npt)
    q = np.zeros(system.nq + 1, dtype=F32)
    q[1:] = F32(system.qstep) * np.arange(system.nq, dtype=F32)
    return Grid(x=x, xstar=xstar, ycoh=ycoh, r2=r2_chain(x, npt1), q=q, npt=npt, npt1=npt1,
                npt2=npt2, nptmin=nptmin, nptmax=nptmax, nq=system.nq)


def sepmax(system, grid):
    """Neighbour-list range of the refinement loop, in box fractions:
    xstar(nptmax) + maxmolsize + 3/dmax (src/epsr26/programs/epsr_standalone_rev3.py:1259)."""
    nat = system.nal - system.naf + 1
    r = np.sqrt(system.xa.astype(np.float64) ** 2 + system.ya.astype(np.float64) ** 2 +
                system.za.astype(np.float64) ** 2)
    maxmolsize = F32(2.0 * r.max()) if nat.max() > 1 else F32(0.0)
    return F32(F32(grid.xstar[grid.nptmax] + maxmolsize) + F32(F32(3.0) / F32(system.dmax)))
