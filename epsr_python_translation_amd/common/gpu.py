"""The device mirror of the COMMON state: one context, the box topology, the cell lists
and pair-energy ledger, and the intramolecular energies of the molecules.  The drop-in
routines upload the coordinates and tables they read from COMMON before each call and
write what they change back, so COMMON stays the state a caller sees."""

from __future__ import annotations

import numpy as np
import torch

from .. import device as D
from . import clustercom, epsr16, molecom, potential_par

F32 = np.float32

_ctx = None
_box = None
_ledger = None
_version = -1
version = 0          # bumped by common.setup whenever the topology changes
eintra = None        # [nmol] intramolecular energy of each molecule (0-based)
intra = None         # the box's intramolecular topology (intra.IntraTopology), if loaded


def reset():
    global _box, _ledger, _version, version, eintra, intra
    _box = None
    _ledger = None
    _version = -1
    version += 1
    eintra = None
    intra = None


def ctx() -> D.Context:
    global _ctx
    if _ctx is None:
        _ctx = D.Context(0)
    return _ctx


def box() -> D.DeviceBox:
    """The box built from COMMON's topology (rebuilt after common.setup)."""
    global _box, _ledger, _version
    if _box is None or _version != version:
        n, m = molecom.natom, clustercom.nmol
        nt = molecom.ntype
        _box = D.DeviceBox(ctx(), itype=molecom.itype[1:n + 1] - 1, naf=clustercom.naf[1:m + 1] - 1,
                           nal=clustercom.nal[1:m + 1] - 1, mtype=clustercom.mtype[1:m + 1] - 1,
                           npairref=np.ascontiguousarray(molecom.npairref[1:nt + 1, 1:nt + 1]) - 1,
                           n1extr=molecom.n1extr[1:molecom.npairs + 1] - 1,
                           n2extr=molecom.n2extr[1:molecom.npairs + 1] - 1, cfrac=molecom.cfrac[1:nt + 1],
                           natomr=molecom.natomr, sizefactor=molecom.sizefactor, dmax=clustercom.dmax,
                           cellst=epsr16.cellst, xa=molecom.xa[1:n + 1], ya=molecom.ya[1:n + 1],
                           za=molecom.za[1:n + 1], xadd=clustercom.xadd[1:m + 1], yadd=clustercom.yadd[1:m + 1],
                           zadd=clustercom.zadd[1:m + 1], nmoltype=clustercom.nmoltype)
        if intra is not None:
            _box.set_intra(intra)
        _ledger = None
        _version = version
    return _box


def upload_coords():
    b = box()
    n, m = molecom.natom, clustercom.nmol
    for name, src, k in (("xa", molecom.xa, n), ("ya", molecom.ya, n), ("za", molecom.za, n),
                         ("xadd", clustercom.xadd, m), ("yadd", clustercom.yadd, m), ("zadd", clustercom.zadd, m)):
        getattr(b, name).copy_(torch.from_numpy(np.ascontiguousarray(src[1:k + 1])))
    return b


def download_coords():
    b = box()
    n, m = molecom.natom, clustercom.nmol
    co = b.get_coords()
    molecom.xa[1:n + 1], molecom.ya[1:n + 1], molecom.za[1:n + 1] = co["xa"], co["ya"], co["za"]
    clustercom.xadd[1:m + 1], clustercom.yadd[1:m + 1], clustercom.zadd[1:m + 1] = co["xadd"], co["yadd"], co["zadd"]


def _columns(a, ncol):
    """Columns 1..ncol of a column-major (ld, ncol+1) COMMON table: [ncol*ld] contiguous."""
    return np.ascontiguousarray(np.asfortranarray(a)[:, 1:ncol + 1].T).reshape(-1)


def tables() -> D.DeviceTables:
    """The tables the energy routines look up: potential_par.xavs/xavs1/ecoul/ecoul1 on
    epsr16.xstar, up to epsr16.npt2."""
    c = ctx()
    npairs = molecom.npairs
    ld = potential_par.xavs.shape[0]
    return D.DeviceTables(c, ld, epsr16.npt2, c.tensor(epsr16.xstar[:ld]),
                          c.tensor(_columns(potential_par.xavs, npairs)),
                          c.tensor(_columns(potential_par.xavs1, npairs)),
                          c.tensor(_columns(potential_par.ecoul, npairs)),
                          c.tensor(_columns(potential_par.ecoul1, npairs)))


def ledger() -> D.Ledger:
    global _ledger
    box()
    if _ledger is None:
        _ledger = D.Ledger(ctx(), _box)
    return _ledger


def has_ledger() -> bool:
    return _ledger is not None and _ledger.built
