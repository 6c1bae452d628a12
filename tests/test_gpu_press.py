"""press on the GPU against the oracle: the energy (the reference's float32 running sum),
the double virial sum, the pressure and every ledger slot's energy and virial, bit for
bit, on the same ledger and tables."""

import numpy as np
import pytest
import torch

from tests import helpers as H

F32 = np.float32
pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from epsr_python_translation_amd import device as D
    return D.Context(0)


@pytest.mark.parametrize("name", ["tiny_12_6_2", "c0", "epsr_test", "c1"])
def test_press_bit_exact(ctx, name):
    from epsr_python_translation_amd import device as D
    from epsr_python_translation_amd import tables as T
    print(f"press parity {name}", flush=True)
    s, g = H.system(name), H.grid(name)
    ref = H.oracle_press(name)
    t = H.oracle_tables(name)
    box = D.DeviceBox.from_system(ctx, s)
    tab = D.DeviceTables(ctx, t["ld"], t["npt2"], t["xstar"], t["xavs"], t["xavs1"], t["ecoul"], t["ecoul1"])
    led = D.Ledger(ctx, box)
    led.cells_init(s.nmolcell, T.sepmax(s, g))
    led.fmolec1(tab, H.eintra(name))
    e, p, w, se, sw = led.press(tab, g.npt1, s.simt, H.box_volume(s), slots=True)
    torch.cuda.synchronize()
    print(f"  energy {e} kJ/mol, pressure {p} kbar", flush=True)
    M = ref["slote"].shape[1]
    assert np.array_equal(se[:, :M].view(np.int32), ref["slote"].view(np.int32))
    assert np.array_equal(sw[:, :M].view(np.int32), ref["slotw"].view(np.int32))
    assert F32(e).view(np.int32) == ref["energy"].view(np.int32)
    assert np.float64(w).view(np.int64) == np.float64(ref["presst"]).view(np.int64)
    assert F32(p).view(np.int32) == ref["pressc"].view(np.int32)


def test_refinement_press_returns_the_reference_tuple(ctx):
    from epsr_python_translation_amd import pipeline as PL
    from epsr_python_translation_amd import press as P
    s = H.system("epsr_test")
    R = PL.Refinement(ctx, s)
    R.build_tables()
    R.cells_and_ledger()
    energf = np.array([F32(3.0)], F32)
    pressf = np.array([F32(-6.0)], F32)
    energy, pressc, energabs, e3, p3 = R.press(energf=energf, pressf=pressf, wsum=np.array([F32(2.0)]))
    assert np.isfinite(energy) and np.isfinite(pressc) and e3 == 0.0 and p3 == 0.0
    assert energabs == P.abs_energy_ep(s.pcoefp)
    assert energf[0] == F32(3.0) / (F32(s.nmol) * F32(2.0))
    assert pressf[0] == ((F32(6.0) / F32(3.0)) * F32(100.0)) / (F32(6.0222) * H.box_volume(s) * F32(2.0))


def test_press_after_an_mc_sweep_bit_exact(ctx):
    """press over the ledger and coordinates an MC sweep left behind, on the reference's
    test box: the GPU's sweep and the oracle's sweep agree bit for bit, and so must press."""
    import ctypes as C
    from epsr_python_translation_amd import device as D
    from epsr_python_translation_amd import tables as T
    from oracle import pyoracle as po
    print("press after one sweep, epsr_test", flush=True)
    name = "epsr_test"
    s, g = H.system(name), H.grid(name)
    ref = H.oracle_fmolec1(name, "cr")
    ob = ref["ob"]
    wr, wt = H.initial_steps(s)
    nwh = np.zeros(2 * s.nmoltype, np.int32)
    nmv = np.zeros(2 * s.nmoltype, np.int32)
    nsr = np.zeros(s.nmol, np.int32)
    mcs = po.MCState(po.Pf(wr), po.Pf(wt), po.Pi(nwh), po.Pi(nmv), po.Pi(nsr), s.simt, F32(1), F32(1), F32(0.75), 0)
    rng = po.Ran1.seeded(-31)
    fadds, nacc = C.c_float(0.0), C.c_int()
    po.lib("cr").ora_molfit(C.byref(ob.box), po.Pf(ob.xa), po.Pf(ob.ya), po.Pf(ob.za), C.byref(ob.centres),
                            C.byref(ref["ctab"]), C.byref(ref["cells"]), C.byref(ref["ledger"]), C.byref(mcs),
                            C.byref(rng), C.byref(fadds), C.byref(nacc))
    want = H.oracle_press(name, "cr", ref=ref)
    t = H.oracle_tables(name)
    box = D.DeviceBox.from_system(ctx, s)
    tab = D.DeviceTables(ctx, t["ld"], t["npt2"], t["xstar"], t["xavs"], t["xavs1"], t["ecoul"], t["ecoul1"])
    led = D.Ledger(ctx, box)
    led.cells_init(s.nmolcell, T.sepmax(s, g))
    led.fmolec1(tab, H.eintra(name))
    wr2, wt2 = H.initial_steps(s)
    mc = D.MCState(ctx, s.nmol, s.nmoltype, s.simt, wholerot=wr2, wholetrans=wt2)
    st = D.ran1_state_array(-31)
    D.molfit(ctx, box, tab, led, mc, st)
    e, p, w, se, sw = led.press(tab, g.npt1, s.simt, H.box_volume(s), slots=True)
    torch.cuda.synchronize()
    M = want["slote"].shape[1]
    assert np.array_equal(se[:, :M].view(np.int32), want["slote"].view(np.int32))
    assert F32(e).view(np.int32) == want["energy"].view(np.int32)
    assert np.float64(w).view(np.int64) == np.float64(want["presst"]).view(np.int64)
    assert F32(p).view(np.int32) == want["pressc"].view(np.int32)
