"""The ran1 state (src/epsr26/common/ran1com.py): mcall, idum, iy, iv(1:32)."""

import numpy as np

from ._arrays import integer


def reset():
    global mcall, idum, iy, iv
    mcall, idum, iy = 0, 0, 0
    iv = integer(32)


def state():
    """The 35-int32 state of the C ABI (mcall, idum, iy, iv(1:32))."""
    return np.array([mcall, idum, iy] + [int(v) for v in iv[1:33]], dtype=np.int32)


def set_state(a):
    global mcall, idum, iy
    mcall, idum, iy = int(a[0]), int(a[1]), int(a[2])
    iv[1:33] = a[3:35]


reset()
