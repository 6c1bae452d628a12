"""Access to tests/golden/ref_xray.npz (written by tests/golden/make_ref_xray_vectors.py from
the reference's own run): the X-ray weights of get_wts and the data-minus-simulation block,
converted to the oracle's and the library's 0-based layouts."""

from __future__ import annotations

import functools
import os

import numpy as np

F32 = np.float32
PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ref_xray.npz")


@functools.lru_cache(maxsize=None)
def load():
    return dict(np.load(PATH, allow_pickle=False))


def group(prefix):
    z = load()
    n = len(prefix) + 1
    return {k[n:]: v for k, v in z.items() if k.startswith(prefix + ".")}


def dms_inputs(ldq=None):
    """The block's inputs as [dataset][ldq] / [pair][ldq] / [xfile][ldq][pair] arrays."""
    d = group("dms.0")
    nq, nf, np_ = int(d["nq"]), int(d["nfile"]), int(d["npairs"])
    ldq = ldq or nq + 1
    rows = lambda a, n: np.ascontiguousarray(np.pad(a.T, ((0, 0), (0, ldq - (nq + 1)))).astype(F32)).reshape(n, ldq)
    wx = d["weightx"]                                  # [nq+1, npairs, nx]
    nx = wx.shape[2]
    weightx = np.zeros((nx, ldq, np_))
    weightx[:, :nq + 1, :] = np.transpose(wx, (2, 0, 1))
    q = np.zeros(ldq, F32)
    q[:nq + 1] = d["q"]
    return dict(nq=nq, nfile=nf, npairs=np_, ldq=ldq, q=q, datai=rows(d["datai"], nf),
                soqintra=rows(d["soqintra"], np_), soqinter=rows(d["soqinter"], np_),
                qwind=rows(d["qwind"], np_), datat=rows(d["datat"], nf), weightx=np.ascontiguousarray(weightx),
                ciwt=np.ascontiguousarray(d["ciwt"].astype(F32)), cwt=np.ascontiguousarray(d["cwt"].astype(F32)),
                ciwt64=d["ciwt"], cwt64=d["cwt"], xfile=d["xfile"].astype(np.int32),
                nqmind=d["nqmind"].astype(np.int32), nqmaxd=d["nqmaxd"].astype(np.int32),
                nqmin=d["nqmin"].astype(np.int32), nqmax=d["nqmax"].astype(np.int32),
                szeros=d["szeros"].astype(F32), porodamp=d["porodamp"].astype(F32), rshmin=d["rshmin"].astype(F32),
                porodq=F32(d["porodq"]), porodpwr=F32(d["porodpwr"]), nsmoop=int(d["nsmoop"]),
                x=np.ascontiguousarray(d["x"].astype(F32)), npt1=int(d["npt1"]), rho=F32(d["rho"]),
                szerostemp=d["szerostemp"].astype(F32), qminm=F32(d["qminm"]), qminp=F32(d["qminp"]))


def fit_yin(k):
    return group(f"fit.{k}.in")["yin"]
