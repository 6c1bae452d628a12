"""Ledger (fmolec1), en_change, en_mole_pair and the MC cycle (molfit) on the GPU against
the oracle.

The GPU evaluates cosf/expf correctly rounded; the oracle's 'cr' build does the same, so
ledgers, coordinates, counters, step sizes and the ran1 state must agree bit for bit.
Against the glibc build (the reference's own libm calls) the chains agree until the first
transcendental that glibc rounds differently; that comparison is reported."""

import ctypes as C

import numpy as np
import pytest
import torch

from oracle import pyoracle as po
from tests import helpers as H

F32 = np.float32
pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from epsr_python_translation_amd import device as D
    return D.Context(0)


def _gpu_setup(ctx, name, coords=None):
    from epsr_python_translation_amd import device as D
    from epsr_python_translation_amd import tables as T
    s, g = H.system(name), H.grid(name)
    t = H.oracle_tables(name)
    box = D.DeviceBox.from_system(ctx, s)
    if coords:
        for k, v in coords.items():
            getattr(box, k).copy_(torch.as_tensor(v))
    tab = D.DeviceTables(ctx, t["ld"], t["npt2"], t["xstar"], t["xavs"], t["xavs1"], t["ecoul"], t["ecoul1"])
    led = D.Ledger(ctx, box)
    led.cells_init(s.nmolcell, T.sepmax(s, g))
    f64, fb64 = led.fmolec1(tab, H.eintra(name))
    return box, tab, led, f64, fb64


@pytest.mark.parametrize("name", ["tiny_12_6_2", "c0", "c1", "epsr_test"])
def test_fmolec1_ledger_bit_exact(ctx, name):
    print(f"fmolec1 parity {name}", flush=True)
    s = H.system(name)
    ref = H.oracle_fmolec1(name, "cr")
    rd = H.ledger_dict(ref["ledger"], s.nmol)
    box, tab, led, f64, fb64 = _gpu_setup(ctx, name)
    gd = led.export()
    assert led.nnbrefmax == rd["nnbrefmax"]
    np.testing.assert_array_equal(gd["nergpairm"], rd["nergpairm"])
    for m in range(s.nmol):
        n = rd["nergpairm"][m]
        assert np.array_equal(gd["molergpair"][m, :n], rd["molergpair"][m, :n]), m
        assert np.array_equal(gd["ergpairref"][m, :n], rd["ergpairref"][m, :n]), m
    assert np.array_equal(gd["ergpair"].view(np.int32), rd["ergpair"].view(np.int32))
    # the chain order equals the oracle's cell order
    cl = ref["cells"]
    cnt = np.ctypeslib.as_array(cl.nmolsincell, shape=(cl.ncell,))
    mic = np.ctypeslib.as_array(cl.molsincell, shape=(cl.ncell * cl.maxcellno,)).reshape(cl.ncell, cl.maxcellno)
    order = np.concatenate([mic[c, :cnt[c]] for c in range(cl.ncell)])
    np.testing.assert_array_equal(gd["order"], order)
    # totals: FP64 on the GPU (block partials, any order) against the oracle's FP64 sum
    # of the same entries: both within n u64 sum|e| of the exact sum (u64 = 2^-53)
    pairs = rd["ergpair"][int(np.count_nonzero(H.eintra(name))):].astype(np.float64)
    bound64 = 2 * len(pairs) * 2.0 ** -53 * np.abs(pairs).sum() / s.nmol
    assert abs(f64 - ref["ftest64"]) <= bound64 + 1e-300
    # the reference's float32 running sum obeys the recursive-summation bound
    # |s32 - s| <= (n-1) u sum|e| (u = 2^-24) against the same entries
    bound = (len(pairs) * 2.0 ** -24) * np.abs(pairs).sum() / s.nmol
    rel32 = abs(f64 - ref["ftest"]) / max(1e-30, abs(f64))
    print(f"fmolec1 {name}: ftest fp64 {f64:.9g}, reference fp32 {ref['ftest']:.9g} (rel {rel32:.2e}, "
          f"bound {bound / max(1e-30, abs(f64)):.2e})")
    assert abs(f64 - ref["ftest"]) <= bound + 1e-6 * abs(f64)
    assert abs(fb64 - float(np.sum(H.eintra(name), dtype=np.float64)) / s.nmol) < 1e-12


def test_en_change_and_en_mole_pair_bit_exact(ctx):
    from epsr_python_translation_amd import device as D
    name = "c0"
    s = H.system(name)
    ref = H.oracle_fmolec1(name, "cr")
    box, tab, led, _, _ = _gpu_setup(ctx, name)
    L = po.lib("cr")
    ob = ref["ob"]
    ergnew_ref = np.zeros(max(1, ref["ledger"].nnbrefmax), F32)
    ergnew_gpu = ctx.zeros(max(1, led.nnbrefmax))
    nn = C.c_int()
    rng = np.random.default_rng(7)
    for im in rng.choice(s.nmol, 12, replace=False):
        im = int(im)
        dx = F32(rng.uniform(-0.01, 0.01))
        ob.xadd[im] = F32(ob.xadd[im] + dx)
        box.xadd[im] = float(ob.xadd[im])
        fr = L.ora_en_change(C.byref(ob.box), C.byref(ob.centres), C.byref(ref["ctab"]), C.byref(ref["ledger"]), im,
                             po.Pf(ergnew_ref), C.byref(nn))
        fg = led.en_change(tab, im, ergnew_gpu)
        torch.cuda.synchronize()
        assert F32(fr).view(np.int32) == F32(fg).view(np.int32), (im, fr, fg)
        eg = ergnew_gpu.cpu().numpy()[:nn.value]
        assert np.array_equal(eg.view(np.int32), ergnew_ref[:nn.value].view(np.int32))
    ii = rng.integers(0, s.nmol, 64)
    jj = rng.integers(0, s.nmol, 64)
    ok = ii != jj
    ii, jj = ii[ok], jj[ok]
    got = D.en_mole_pair(ctx, box, tab, ii, jj).cpu().numpy()
    want = np.array([L.ora_en_mole_pair(C.byref(ob.box), C.byref(ob.centres), C.byref(ref["ctab"]), int(a), int(b))
                     for a, b in zip(ii, jj)], F32)
    assert np.array_equal(got.view(np.int32), want.view(np.int32))


def _run_oracle_molfit(name, flavour, nsweep, idum, steps=None):
    s = H.system(name)
    ref = H.oracle_fmolec1(name, flavour)
    ob = ref["ob"]
    wr, wt = H.initial_steps(s) if steps is None else (steps[0].copy(), steps[1].copy())
    nwh = np.zeros(2 * s.nmoltype, np.int32)
    nmv = np.zeros(2 * s.nmoltype, np.int32)
    nsr = np.zeros(s.nmol, np.int32)
    mc = po.MCState(po.Pf(wr), po.Pf(wt), po.Pi(nwh), po.Pi(nmv), po.Pi(nsr), s.simt, F32(1), F32(1), F32(0.75), 0)
    rng = po.Ran1.seeded(idum)
    fadds = C.c_float(0.0)
    nacc = C.c_int()
    for _ in range(nsweep):
        po.lib(flavour).ora_molfit(C.byref(ob.box), po.Pf(ob.xa), po.Pf(ob.ya), po.Pf(ob.za), C.byref(ob.centres),
                                   C.byref(ref["ctab"]), C.byref(ref["cells"]), C.byref(ref["ledger"]), C.byref(mc),
                                   C.byref(rng), C.byref(fadds), C.byref(nacc))
    return dict(ob=ob, ledger=H.ledger_dict(ref["ledger"], s.nmol), wr=wr, wt=wt, nwhole=nwh, nmovem=nmv, nsr=nsr,
                ran1=rng.as_array(), fadds=F32(fadds.value))


@pytest.mark.parametrize("name,nsweep", [("tiny_13_12_3", 3), ("c0", 2), ("c1", 1), ("epsr_test", 2), ("dense_1400", 1)])
def test_molfit_bit_exact_with_correctly_rounded_libm(ctx, name, nsweep):
    from epsr_python_translation_amd import device as D
    print(f"molfit parity {name} sweeps={nsweep}", flush=True)
    s = H.system(name)
    idum = -2718
    ref = _run_oracle_molfit(name, "cr", nsweep, idum)
    box, tab, led, _, _ = _gpu_setup(ctx, name)
    wr, wt = H.initial_steps(s)
    mc = D.MCState(ctx, s.nmol, s.nmoltype, s.simt, wholerot=wr, wholetrans=wt)
    st = D.ran1_state_array(idum)
    for _ in range(nsweep):
        D.molfit(ctx, box, tab, led, mc, st)
        print(f"  sweep: accepted {mc.stats.accepted}/{mc.stats.moves}, exact resolutions {mc.stats.exact_resolutions}", flush=True)
    torch.cuda.synchronize()
    co = box.get_coords()
    ob = ref["ob"]
    for k in ("xa", "ya", "za", "xadd", "yadd", "zadd"):
        assert np.array_equal(co[k].view(np.int32), getattr(ob, k).view(np.int32)), k
    np.testing.assert_array_equal(st, ref["ran1"])
    assert np.array_equal(mc.wholerot.view(np.int32), ref["wr"].view(np.int32))
    assert np.array_equal(mc.wholetrans.view(np.int32), ref["wt"].view(np.int32))
    np.testing.assert_array_equal(mc.nwhole, ref["nwhole"])
    np.testing.assert_array_equal(mc.nmovem, ref["nmovem"])
    np.testing.assert_array_equal(mc.nstepreject.cpu().numpy(), ref["nsr"])
    assert F32(mc.fadds).view(np.int32) == ref["fadds"].view(np.int32)
    gd = led.export()
    assert np.array_equal(gd["ergpair"].view(np.int32), ref["ledger"]["ergpair"].view(np.int32))


def test_molfit_against_glibc_libm_chain(ctx):
    """Same chain with the reference's glibc cosf/expf: identical until the first value
    glibc rounds differently from the correctly rounded result; report the agreement."""
    from epsr_python_translation_amd import device as D
    name = "c0"
    s = H.system(name)
    ref = _run_oracle_molfit(name, "glibc", 1, -99)
    box, tab, led, _, _ = _gpu_setup(ctx, name)
    wr, wt = H.initial_steps(s)
    mc = D.MCState(ctx, s.nmol, s.nmoltype, s.simt, wholerot=wr, wholetrans=wt)
    st = D.ran1_state_array(-99)
    D.molfit(ctx, box, tab, led, mc, st)
    co = box.get_coords()
    same = np.mean(co["xadd"] == ref["ob"].xadd)
    print(f"molfit vs glibc chain: {same * 100:.1f}% of centres identical after one sweep")
    np.testing.assert_array_equal(st, ref["ran1"])      # RNG consumption does not depend on libm
    assert same > 0.5


@pytest.mark.parametrize("change", ["grow", "scale"])
def test_molfit_after_tables_changed_in_place(ctx, change):
    """MC cycles on tables whose arrays were changed in place after they were measured and
    packed (the empirical potential grown past its last non-zero bin, or scaled within its
    range) give the bits of cycles on a copy of the changed tables, packed afresh."""
    from epsr_python_translation_amd import device as D
    name = "c0"
    s = H.system(name)
    t = H.oracle_tables(name)
    ld, npt2 = t["ld"], t["npt2"]
    xavs = np.asarray(t["xavs"], F32).reshape(-1, ld)
    last = int(np.nonzero(np.any(xavs[:, :npt2 + 1] != 0, axis=0))[0].max())
    grown = xavs.copy()
    if change == "grow":
        hi = min(npt2, last + 300)
        assert hi > last + 1
        grown[:, last + 1:hi] = F32(2e-3)
        print(f"tables changed in place: xavs grown from bin {last} to {hi - 1} of {npt2}", flush=True)
    else:
        grown = (grown * F32(1.25)).astype(F32)
        print(f"tables changed in place: xavs scaled by 1.25 up to bin {last}", flush=True)
    outs = []
    for inplace in (True, False):
        box, tab, led, _, _ = _gpu_setup(ctx, name)
        wr, wt = H.initial_steps(s)
        mc = D.MCState(ctx, s.nmol, s.nmoltype, s.simt, wholerot=wr, wholetrans=wt)
        st = D.ran1_state_array(-31)
        D.molfit(ctx, box, tab, led, mc, st)
        torch.cuda.synchronize()
        if inplace:
            tab.xavs.copy_(torch.as_tensor(grown.ravel()))
        else:
            tab = D.DeviceTables(ctx, ld, npt2, t["xstar"], grown.ravel(), t["xavs1"], t["ecoul"], t["ecoul1"])
        torch.cuda.synchronize()
        for _ in range(2):
            D.molfit(ctx, box, tab, led, mc, st)
        torch.cuda.synchronize()
        co = box.get_coords()
        outs.append(dict(co=co, st=st.copy(), erg=led.export()["ergpair"].copy(), wr=mc.wholerot.copy(),
                         wt=mc.wholetrans.copy(), fadds=F32(mc.fadds)))
    a, b = outs
    for k in ("xa", "ya", "za", "xadd", "yadd", "zadd"):
        assert np.array_equal(a["co"][k].view(np.int32), b["co"][k].view(np.int32)), k
    np.testing.assert_array_equal(a["st"], b["st"])
    assert np.array_equal(a["erg"].view(np.int32), b["erg"].view(np.int32))
    assert np.array_equal(a["wr"].view(np.int32), b["wr"].view(np.int32))
    assert np.array_equal(a["wt"].view(np.int32), b["wt"].view(np.int32))
    assert a["fadds"].view(np.int32) == b["fadds"].view(np.int32)


@pytest.mark.parametrize("gap", [1, 599, 5000])
def test_molfit_after_other_ran1_draws(ctx, gap):
    """Cycles separated by `gap` draws of another ran1 consumer (atogen draws one value per
    molecule) read their streams where the caller's state lies: within the window the
    previous cycle uploaded (gap up to nmol + 256), else in a window drawn afresh; the
    chain equals the oracle's with the same draws between cycles."""
    from epsr_python_translation_amd import _native
    from epsr_python_translation_amd import device as D
    name = "c0"
    s = H.system(name)
    idum = -4242
    print(f"molfit with {gap} other draws between cycles", flush=True)
    ref = H.oracle_fmolec1(name, "cr")
    ob = ref["ob"]
    wr, wt = H.initial_steps(s)
    nwh = np.zeros(2 * s.nmoltype, np.int32)
    nmv = np.zeros(2 * s.nmoltype, np.int32)
    nsr = np.zeros(s.nmol, np.int32)
    omc = po.MCState(po.Pf(wr), po.Pf(wt), po.Pi(nwh), po.Pi(nmv), po.Pi(nsr), s.simt, F32(1), F32(1), F32(0.75), 0)
    rng = po.Ran1.seeded(idum)
    fadds = C.c_float(0.0)
    nacc = C.c_int()
    junk = np.empty(gap, np.float32)
    L = po.lib("cr")
    box, tab, led, _, _ = _gpu_setup(ctx, name)
    gwr, gwt = H.initial_steps(s)
    mc = D.MCState(ctx, s.nmol, s.nmoltype, s.simt, wholerot=gwr, wholetrans=gwt)
    st = D.ran1_state_array(idum)
    lib = _native.lib()
    for cycle in range(3):
        if cycle:
            L.ora_ran1_stream(C.byref(rng), gap, po.Pf(junk))
            assert lib.epsr_ran1_stream(st.ctypes.data_as(C.POINTER(C.c_int32)), gap,
                                        junk.ctypes.data_as(C.POINTER(C.c_float))) == 0
        L.ora_molfit(C.byref(ob.box), po.Pf(ob.xa), po.Pf(ob.ya), po.Pf(ob.za), C.byref(ob.centres),
                     C.byref(ref["ctab"]), C.byref(ref["cells"]), C.byref(ref["ledger"]), C.byref(omc), C.byref(rng),
                     C.byref(fadds), C.byref(nacc))
        D.molfit(ctx, box, tab, led, mc, st)
    torch.cuda.synchronize()
    co = box.get_coords()
    for k in ("xa", "ya", "za", "xadd", "yadd", "zadd"):
        assert np.array_equal(co[k].view(np.int32), getattr(ob, k).view(np.int32)), k
    np.testing.assert_array_equal(st, rng.as_array())
    assert np.array_equal(led.export()["ergpair"].view(np.int32),
                          H.ledger_dict(ref["ledger"], s.nmol)["ergpair"].view(np.int32))
    assert F32(mc.fadds).view(np.int32) == F32(fadds.value).view(np.int32)


def test_movem_step_doubling_bit_exact(ctx):
    """A translation step far below the float32 spacing of the centres: movem doubles it
    until the centre moves (molfit.py:747-758), the doubled step persists for the type, and
    the GPU's host-resolved redo of the cycle equals the oracle's chain bit for bit."""
    from epsr_python_translation_amd import device as D
    name = "tiny_13_12_3"
    s = H.system(name)
    wr = np.full(s.nmoltype, 0.05, F32)
    wt = np.full(s.nmoltype, 1.0e-12, F32)
    print("molfit with a 1e-12 translation step (movem doubling)", flush=True)
    ref = _run_oracle_molfit(name, "cr", 1, -77, steps=(wr, wt))
    box, tab, led, _, _ = _gpu_setup(ctx, name)
    mc = D.MCState(ctx, s.nmol, s.nmoltype, s.simt, wholerot=wr.copy(), wholetrans=wt.copy())
    st = D.ran1_state_array(-77)
    D.molfit(ctx, box, tab, led, mc, st)
    torch.cuda.synchronize()
    assert mc.stats.retries > 0, "the step was doubled"
    co = box.get_coords()
    for k in ("xa", "ya", "za", "xadd", "yadd", "zadd"):
        assert np.array_equal(co[k].view(np.int32), getattr(ref["ob"], k).view(np.int32)), k
    np.testing.assert_array_equal(st, ref["ran1"])
    assert np.array_equal(mc.wholetrans.view(np.int32), ref["wt"].view(np.int32))
    assert np.all(mc.wholetrans > 1.0e-12)
    gd = led.export()
    assert np.array_equal(gd["ergpair"].view(np.int32), ref["ledger"]["ergpair"].view(np.int32))


def test_resolver_pass_cap_reports_undecided_moves(ctx):
    """The one-workgroup resolution stops at its pass cap (set small from the host) and the
    cycle fails loudly instead of committing undecided moves.  A cap leaves one grid-wide
    pass, so the finisher has work; with a cap large enough the cycle equals the uncapped one."""
    from epsr_python_translation_amd import _native
    from epsr_python_translation_amd import device as D
    name = "c1"
    s = H.system(name)
    wr, wt = H.initial_steps(s)
    box0, tab0, led0, _, _ = _gpu_setup(ctx, name)
    mc0 = D.MCState(ctx, s.nmol, s.nmoltype, s.simt, wholerot=wr, wholetrans=wt)
    st0 = D.ran1_state_array(-5)
    D.molfit(ctx, box0, tab0, led0, mc0, st0)
    box, tab, led, _, _ = _gpu_setup(ctx, name)
    box.set_launch(_native.LAUNCH_RESOLVE_PASSES, s.nmol + 1)
    mc = D.MCState(ctx, s.nmol, s.nmoltype, s.simt, wholerot=wr, wholetrans=wt)
    st = D.ran1_state_array(-5)
    D.molfit(ctx, box, tab, led, mc, st)
    finisher = mc.stats.resolver_passes - 1
    print(f"resolver: {finisher} finishing passes after one grid pass", flush=True)
    assert finisher >= 1
    co0, co = box0.get_coords(), box.get_coords()
    for k in co:
        assert np.array_equal(co[k].view(np.int32), co0[k].view(np.int32)), k
    np.testing.assert_array_equal(st, st0)
    # no finishing pass: the moves the grid-wide pass left are undecided (how many passes a
    # run needs varies with the timing of concurrent decisions, so the cap is not finisher - 1)
    box2, tab2, led2, _, _ = _gpu_setup(ctx, name)
    box2.set_launch(_native.LAUNCH_RESOLVE_PASSES, 0)
    mc2 = D.MCState(ctx, s.nmol, s.nmoltype, s.simt, wholerot=wr, wholetrans=wt)
    st2 = D.ran1_state_array(-5)
    with pytest.raises(D.EpsrError, match="undecided"):
        D.molfit(ctx, box2, tab2, led2, mc2, st2)
