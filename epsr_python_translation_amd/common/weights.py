"""Dataset weights (src/epsr26/common/weights.py): cwt(id, ic) for the fit of each
dataset, cwtpot(ic, id) for the potential update, xfile(id) (X-ray data: unsupported)."""

from ._arrays import integer, real


def reset(nfile=0, npairs=0):
    global cwt, cwtpot, xfile
    cwt = real(nfile, npairs)
    cwtpot = real(npairs, nfile)
    xfile = integer(nfile)


reset()
