"""MC cycles back to back (epsr_molfit_cycles): the step sizes adapt on the device between
cycles and the host waits once, so every output must be what the same number of single
epsr_molfit calls leaves -- coordinates, ledger, step sizes, the last cycle's counters and
stats, nstepreject, ran1 state and fadds, bit for bit.  Cases: the 3200-atom box (also
against the correctly rounded oracle's chain), the 60,000-atom water box, a tiny box whose
1e-12 translation step makes movem double it in several cycles (a cycle that must double
commits nothing on the device and is run again alone) over 70 cycles (more than one device
batch of 64), ireset = 1 (no adaptation), and batches with other ran1 draws between them."""

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


def _setup(ctx, name, steps, ireset):
    from epsr_python_translation_amd import device as D
    from epsr_python_translation_amd import tables as T
    s, g = H.system(name), H.grid(name)
    t = H.oracle_tables(name)
    box = D.DeviceBox.from_system(ctx, s)
    tab = D.DeviceTables(ctx, t["ld"], t["npt2"], t["xstar"], t["xavs"], t["xavs1"], t["ecoul"], t["ecoul1"])
    led = D.Ledger(ctx, box)
    led.cells_init(s.nmolcell, T.sepmax(s, g))
    led.fmolec1(tab, H.eintra(name))
    wr, wt = steps
    mc = D.MCState(ctx, s.nmol, s.nmoltype, s.simt, ireset=ireset, wholerot=wr.copy(), wholetrans=wt.copy())
    return s, box, tab, led, mc


def _state(box, led, mc, st):
    torch.cuda.synchronize()
    co = box.get_coords()
    out = {k: co[k].view(np.int32).copy() for k in ("xa", "ya", "za", "xadd", "yadd", "zadd")}
    out.update(ran1=st.copy(), wr=mc.wholerot.view(np.int32).copy(), wt=mc.wholetrans.view(np.int32).copy(),
               nwhole=mc.nwhole.copy(), nmovem=mc.nmovem.copy(), nsr=mc.nstepreject.cpu().numpy().copy(),
               fadds=F32(mc.fadds).view(np.int32), erg=led.export()["ergpair"].view(np.int32).copy(),
               accepted=mc.stats.accepted)
    # (the resolver's pass and exact-sum counts are left out: moves decided within a grid-wide
    # pass depend on which partners other workgroups have decided by then, so those counts
    # vary from run to run while every decision is the chain's)
    return out


def _same(a, b, what):
    for k in a:
        assert np.array_equal(np.asarray(a[k]), np.asarray(b[k])), f"{what}: {k}"


def _steps(s, kind):
    if kind == "halting":
        return np.full(s.nmoltype, 0.05, F32), np.full(s.nmoltype, 1.0e-12, F32)
    return H.initial_steps(s)


@pytest.mark.parametrize("name,ncyc,kind,ireset", [("c0", 6, "default", 0), ("c1", 3, "default", 0),
                                                   ("tiny_13_12_3", 70, "halting", 0),
                                                   ("epsr_test", 4, "default", 1)])
def test_cycles_equal_single_calls(ctx, name, ncyc, kind, ireset):
    from epsr_python_translation_amd import device as D
    print(f"molfit_cycles {name}: {ncyc} cycles back to back ({kind} steps, ireset {ireset})", flush=True)
    idum = -1357
    steps = _steps(H.system(name), kind)
    s, box, tab, led, mc = _setup(ctx, name, steps, ireset)
    st = D.ran1_state_array(idum)
    halts = acc = 0
    for _ in range(ncyc):
        D.molfit(ctx, box, tab, led, mc, st)
        halts += mc.stats.retries > 0
        acc += mc.stats.accepted
    one = _state(box, led, mc, st)
    s, box, tab, led, mc = _setup(ctx, name, steps, ireset)
    st = D.ran1_state_array(idum)
    D.molfit(ctx, box, tab, led, mc, st, ncycles=ncyc)
    many = _state(box, led, mc, st)
    print(f"  cycles with movem doubling: {halts}; last cycle accepted {many['accepted']}", flush=True)
    _same(one, many, f"{name} {ncyc} cycles")
    assert mc.stats.cycles == ncyc and mc.stats.accepted_total == acc
    if kind == "halting":
        assert halts >= 1
    if name == "c0":
        # and the chain is the oracle's
        ref = _oracle_chain(name, ncyc, idum, steps)
        for k in ("xa", "ya", "za", "xadd", "yadd", "zadd"):
            assert np.array_equal(many[k], getattr(ref["ob"], k).view(np.int32)), k
        np.testing.assert_array_equal(many["ran1"], ref["ran1"])
        assert np.array_equal(many["wr"], ref["wr"].view(np.int32))
        assert np.array_equal(many["wt"], ref["wt"].view(np.int32))
        assert many["fadds"] == ref["fadds"].view(np.int32)


def _oracle_chain(name, ncyc, idum, steps):
    s = H.system(name)
    ref = H.oracle_fmolec1(name, "cr")
    ob = ref["ob"]
    wr, wt = steps[0].copy(), steps[1].copy()
    nwh = np.zeros(2 * s.nmoltype, np.int32)
    nmv = np.zeros(2 * s.nmoltype, np.int32)
    nsr = np.zeros(s.nmol, np.int32)
    mc = po.MCState(po.Pf(wr), po.Pf(wt), po.Pi(nwh), po.Pi(nmv), po.Pi(nsr), s.simt, F32(1), F32(1), F32(0.75), 0)
    rng = po.Ran1.seeded(idum)
    fadds = C.c_float(0.0)
    nacc = C.c_int()
    for _ in range(ncyc):
        po.lib("cr").ora_molfit(C.byref(ob.box), po.Pf(ob.xa), po.Pf(ob.ya), po.Pf(ob.za), C.byref(ob.centres),
                                C.byref(ref["ctab"]), C.byref(ref["cells"]), C.byref(ref["ledger"]), C.byref(mc),
                                C.byref(rng), C.byref(fadds), C.byref(nacc))
    return dict(ob=ob, wr=wr, wt=wt, ran1=rng.as_array(), fadds=F32(fadds.value))


def test_batches_with_other_draws_between(ctx):
    """Batches and single cycles with another ran1 consumer between them (atogen's origin
    draws): a batch leaves the window of the call after it drawn and uploaded, and that call
    reads its stream where the caller's state lies."""
    from epsr_python_translation_amd import _native
    from epsr_python_translation_amd import device as D
    name = "c0"
    lib = _native.lib()
    plan = [("batch", 3), ("draw", 599), ("batch", 2), ("draw", 5000), ("single", 1), ("batch", 4), ("single", 1)]
    outs = []
    for batched in (False, True):
        s, box, tab, led, mc = _setup(ctx, name, H.initial_steps(H.system(name)), 0)
        st = D.ran1_state_array(-8642)
        for what, n in plan:
            if what == "draw":
                junk = np.empty(n, F32)
                assert lib.epsr_ran1_stream(st.ctypes.data_as(C.POINTER(C.c_int32)), n,
                                            junk.ctypes.data_as(C.POINTER(C.c_float))) == 0
            elif batched and what == "batch":
                D.molfit(ctx, box, tab, led, mc, st, ncycles=n)
            else:
                for _ in range(n):
                    D.molfit(ctx, box, tab, led, mc, st)
        outs.append(_state(box, led, mc, st))
    _same(outs[0], outs[1], "batches with draws between")


def test_cycles_kernel_timer_counts_every_launch(ctx):
    """Every cycle of a batch is timed: the summed timer counts one mc_pair launch per cycle."""
    from epsr_python_translation_amd import device as D
    name = "c0"
    tctx = D.Context(0, timing=True)
    s, box, tab, led, mc = _setup(tctx, name, H.initial_steps(H.system(name)), 0)
    st = D.ran1_state_array(-11)
    tctx.kernel_ms_sum("mc_pair")
    D.molfit(tctx, box, tab, led, mc, st, ncycles=5)
    ms, n = tctx.kernel_ms_sum("mc_pair")
    print(f"  mc_pair: {n} launches, {ms:.3f} ms", flush=True)
    assert n == 5 and ms > 0.0
    assert tctx.kernel_ms_sum("mc_pair") == (-1.0, 0)
