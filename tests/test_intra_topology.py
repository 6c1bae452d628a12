"""The intramolecular topology and MC run parameters read from an .ato file
(formats.system_from_files -> intra.from_ato) against what the reference's own readers left
in COMMON for the same file (tests/golden/ref_tinybox.npz: `case.ato` and the recorded
topology, written by tests/golden/make_ref_vectors.py)."""

from __future__ import annotations

import numpy as np

from epsr_python_translation_amd import formats as FM
from tests import refgolden as RG

F32 = np.float32


def _system(tmp_path):
    p = tmp_path / "case.ato"
    p.write_text(str(RG.load()["case.ato"]))
    return FM.system_from_files(str(p))


def test_topology_matches_the_reference(tmp_path):
    s = _system(tmp_path)
    t = RG.group("topo")
    it = s.intra
    assert s.eintra is None
    assert np.array_equal(it.qtype, t["qtype"])
    assert np.array_equal(it.imageatom + 1, t["imageatom"])
    assert np.array_equal(it.nbond, t["nbond"])
    for ia in range(s.natom):
        nb = int(t["nbond"][ia])
        assert np.array_equal(it.iabond[ia, :nb] + 1, t["iabond"][ia, 1:nb + 1]), ia
        assert np.array_equal(it.tbond[ia, :nb].view(np.int32), t["tbond"][ia, 1:nb + 1].astype(F32).view(np.int32)), ia
    assert np.array_equal(it.amassr.view(np.int32), t["amassr"].astype(F32).view(np.int32))
    assert np.array_equal(it.amassa.view(np.int32), t["amassa"].astype(F32).view(np.int32))
    assert np.array_equal(it.ngrot, t["ngrot"])
    has = it.ngrot > 0
    assert np.array_equal(it.nrf[has] + 1, t["nrf"][has]) and np.array_equal(it.nrl[has] + 1, t["nrl"][has])
    assert np.array_equal(it.labaxi + 1, t["labaxi"][1:, 1:3])
    assert np.array_equal(it.rotf + 1, t["rotf"][1:]) and np.array_equal(it.rotl + 1, t["rotl"][1:])
    assert np.array_equal(it.labrot + 1, t["labrot"][1:])
    assert F32(it.dmaxmol) == F32(t["dmaxmol"]) and F32(it.tempb) == F32(t["tempb"])


def test_run_parameters_match_the_reference(tmp_path):
    s = _system(tmp_path)
    r = s.mc_run
    cap, cin = RG.group("molfit.0"), RG.group("molfit.0.in")
    for k in ("stepr", "stepm", "stepri", "stepmi", "tempangwt", "tempdihwt"):
        assert F32(r[k]) == F32(cap[k]), k
    for k in ("wholerot", "wholetrans", "internrot", "interntrans"):
        assert np.array_equal(r[k].view(np.int32), cin[k].astype(F32).view(np.int32)), k
    assert np.array_equal(r["nstepreject"], cin["nstepreject"])
