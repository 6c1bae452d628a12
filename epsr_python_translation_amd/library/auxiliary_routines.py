"""auxiliary_routines.ran1 and smoop (src/epsr26/library/auxiliary_routines.py:324-385,
744-764) over the COMMON mirror."""

from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from ..common import gpu, ran1com
from .. import device as D

F32 = np.float32


def ran1():
    """One deviate of the Park-Miller + Bays-Durham generator, advancing ran1com."""
    st = ran1com.state()
    out = np.zeros(1, F32)
    lib = gpu.ctx().lib
    rc = lib.epsr_ran1_stream(st.ctypes.data_as(C.POINTER(C.c_int32)), 1, out.ctypes.data_as(C.POINTER(C.c_float)))
    if rc != 0:
        raise D.EpsrError(f"epsr_ran1_stream failed with {rc}")
    ran1com.set_state(st)
    return F32(out[0])


def smoop(nqmin, nqmax, nsmoo, diff):
    """3-point smoothing of diff(nqmin..nqmax), nsmoo times, in place (on the GPU)."""
    c = gpu.ctx()
    t = c.tensor(np.ascontiguousarray(diff, dtype=F32))
    D.smoop(c, nqmin, nqmax, nsmoo, t)
    diff[:] = t.cpu().numpy()
