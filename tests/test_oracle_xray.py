"""The oracle's data-minus-simulation block, ftback and ptgnsq differences against the
reference's own run on an X-ray + neutron case (tests/golden/ref_xray.npz, from
tests/golden/make_ref_xray_vectors.py): bit for bit, glibc flavour (the recording's libm)."""

from __future__ import annotations

import ctypes as C

import numpy as np

from oracle import pyoracle as po
from tests import refxray as RX

F32 = np.float32


def _bits(a):
    return np.ascontiguousarray(a, F32).view(np.int32)


def _dms_struct(d):
    keep = []
    P = lambda a: (keep.append(a), po.Pf(a))[1]
    I = lambda a: (keep.append(a), po.Pi(a))[1]
    wx = d["weightx"]
    keep.append(wx)
    s = po.DMS(d["nfile"], d["nq"], d["ldq"], d["npairs"], P(d["q"]), P(d["datai"]), P(d["soqintra"]),
               P(d["soqinter"]), P(d["ciwt"]), P(d["cwt"]), wx.ctypes.data_as(po.dp), I(d["xfile"]), I(d["nqmind"]),
               I(d["nqmaxd"]), I(d["nqmin"]), I(d["nqmax"]), P(d["szeros"]), P(d["porodamp"]), P(d["rshmin"]),
               d["porodq"], d["porodpwr"], d["nsmoop"], P(d["x"]), d["npt1"], d["rho"])
    return s, keep


def test_weights_hold_float32_values():
    """The reference keeps ciwt/cwt in float64 arrays; their values are float32 products,
    so the float32 copies the oracle and the library take lose nothing."""
    d = RX.dms_inputs()
    assert np.array_equal(d["ciwt"].astype(np.float64), d["ciwt64"])
    assert np.array_equal(d["cwt"].astype(np.float64), d["cwt64"])


def test_data_minus_simulation_matches_the_reference():
    d = RX.dms_inputs()
    assert d["nsmoop"] == 1 and (d["xfile"] > 0).any() and (d["porodamp"] > 0).any() and (d["szeros"] != 0).any()
    s, keep = _dms_struct(d)
    datat = np.zeros((d["nfile"], d["ldq"]), F32)
    szt = np.zeros(d["nfile"], F32)
    dts = np.zeros(d["ldq"], F32)
    grt = np.zeros(d["npt1"] + 1, F32)
    po.lib("glibc").ora_data_minus_sim(C.byref(s), po.Pf(datat), po.Pf(szt), po.Pf(dts), po.Pf(grt))
    nq = d["nq"]
    assert np.array_equal(_bits(datat[:, 1:nq + 1]), _bits(d["datat"][:, 1:nq + 1]))
    assert np.array_equal(_bits(szt), _bits(d["szerostemp"]))


def test_data_minus_simulation_each_branch_matters():
    """Leaving out the background subtraction, the Porod term or the intramolecular
    subtraction changes the result: each is exercised by the recorded case."""
    d = RX.dms_inputs()
    nq = d["nq"]
    base = None
    for change in (None, "nsmoop", "porod", "intra"):
        e = dict(d)
        if change == "nsmoop":
            e["nsmoop"] = 0
        elif change == "porod":
            e["porodamp"] = np.zeros_like(d["porodamp"])
        elif change == "intra":
            e["soqintra"] = np.zeros_like(d["soqintra"])
        s, keep = _dms_struct(e)
        datat = np.zeros((d["nfile"], d["ldq"]), F32)
        szt = np.zeros(d["nfile"], F32)
        po.lib("glibc").ora_data_minus_sim(C.byref(s), po.Pf(datat), po.Pf(szt), po.Pf(np.zeros(d["ldq"], F32)),
                                           po.Pf(np.zeros(d["npt1"] + 1, F32)))
        if base is None:
            base = datat.copy()
        else:
            assert not np.array_equal(_bits(datat[:, 1:nq + 1]), _bits(base[:, 1:nq + 1])), change


def test_ptgnsq_differences_match_the_reference():
    """ptgnsq's first step on the block's output (data = datat, szeros = szerostemp, the
    current intermolecular partials): -qtemp is what the reference hands fit_expansion,
    for the two neutron datasets and the X-ray one."""
    d = RX.dms_inputs()
    nq, nf, ldq = d["nq"], d["nfile"], d["ldq"]
    qt = np.zeros(nf * ldq, F32)
    qmin = d["qminp"] if d["qminp"] > 0.0 else d["qminm"]
    po.lib("glibc").ora_ptgnsq_diff_x(nf, nq, d["npairs"], po.Pf(d["q"]), po.Pf(d["datat"]), ldq, po.Pf(d["qwind"]),
                                      po.Pf(d["cwt"]), d["weightx"].ctypes.data_as(po.dp), po.Pi(d["xfile"]),
                                      po.Pf(d["szerostemp"]), po.Pi(d["nqmin"]), po.Pi(d["nqmax"]), qmin, po.Pf(qt))
    qt = qt.reshape(nf, ldq)
    for k in range(nf):
        n = int(d["nqmax"][k])
        assert np.array_equal(_bits(-qt[k, 1:n + 1]), _bits(RX.fit_yin(k)[1:n + 1])), k
