"""The GPU's intramolecular energies and its complete MC cycle against the reference's own
recorded calls (tests/golden/ref_tinybox.npz, from tests/golden/make_ref_vectors.py: the
reference's epsr program on 4 DMSO + 20 water taken from its EPSR-test box, with bonds, angle
constraints and rotational groups): en_mole_intra of every molecule, fmolec1's ledger, and
the six recorded molfit cycles -- the shake (movei judged on the intramolecular energy, then
internal rotations), four whole-molecule cycles and the cycle with internal rotations -- run
back to back from the recorded start, bit for bit: coordinates, ran1 state, the four step
sizes and their counters, nstepreject, the ledger, fadds/faddbs and nisum.  (The reference's
transcendentals come from glibc and the GPU's are correctly rounded; on these inputs they
agree, as the correctly rounded oracle's run of the same cycles shows.)"""


import numpy as np
import pytest
import torch

from epsr_python_translation_amd import formats as FM
from tests import refgolden as RG

F32 = np.float32
pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from epsr_python_translation_amd import device as D
    return D.Context(0)


def _bits(a):
    return np.ascontiguousarray(a, F32).view(np.int32)


def _setup(ctx, tmp_path):
    from epsr_python_translation_amd import device as D
    p = tmp_path / "case.ato"
    p.write_text(str(RG.load()["case.ato"]))
    s = FM.system_from_files(str(p))
    t0 = RG.group("fmolec1.0")
    box = D.DeviceBox(ctx, itype=s.itype, naf=s.naf, nal=s.nal, mtype=s.mtype, npairref=s.npairref, n1extr=s.n1extr,
                      n2extr=s.n2extr, cfrac=s.cfrac, natomr=s.natomr, sizefactor=1.0, dmax=s.dmax, cellst=s.cellst,
                      xa=t0["xa"], ya=t0["ya"], za=t0["za"], xadd=t0["xadd"], yadd=t0["yadd"], zadd=t0["zadd"],
                      nmoltype=s.nmoltype)
    box.set_intra(s.intra)
    tab = RG.group("fmolec1.0.tables")
    ld = int(tab["xavs"].shape[1])
    dt = {k: ctx.tensor(np.ascontiguousarray(tab[k], F32).reshape(-1)) for k in
          ("xavs", "xavs1", "ecoul", "ecoul1", "xav", "xav1")}
    tables = D.DeviceTables(ctx, ld, int(tab["npt2"]), ctx.tensor(np.ascontiguousarray(tab["xstar"], F32)),
                            dt["xavs"], dt["xavs1"], dt["ecoul"], dt["ecoul1"])
    return s, box, tables, dt


def _set_coords(box, rec):
    for k in ("xa", "ya", "za", "xadd", "yadd", "zadd"):
        getattr(box, k).copy_(torch.from_numpy(np.ascontiguousarray(rec[k], F32)))


def _fmolec1(ctx, s, box, tables, dt, led, k):
    from epsr_python_translation_amd import device as D
    topo = RG.group("topo")
    cells = RG.group("cells.0")
    led.cells_init(int(cells["nmolcellin"]), F32(cells["rmax"]))
    ei = D.en_mole_intra(ctx, box, tables, F32(topo["simt"]), F32(topo["tempangreq"]),
                         F32(F32(topo["tempdihreq"]) * F32(topo["tempb"])), xav=dt["xav"], xav1=dt["xav1"])
    return ei, led.fmolec1(tables, ei)


@pytest.mark.parametrize("k", [0, 1])
def test_en_mole_intra_and_ledger_match_the_reference(ctx, tmp_path, k):
    from epsr_python_translation_amd import device as D
    print(f"en_mole_intra + fmolec1 against the reference's call {k}", flush=True)
    s, box, tables, dt = _setup(ctx, tmp_path)
    cap = RG.group(f"fmolec1.{k}")
    _set_coords(box, cap)
    led = D.Ledger(ctx, box)
    ei, _ = _fmolec1(ctx, s, box, tables, dt, led, k)
    torch.cuda.synchronize()
    assert np.array_equal(_bits(ei.cpu().numpy()), _bits(RG.self_energies(cap, s.nmol)))
    ex = led.export()
    want = RG.record_rows(cap, s.nmol)
    assert np.array_equal(ex["nergpairm"], want[0])
    for im in range(s.nmol):
        n = int(want[0][im])
        assert np.array_equal(ex["molergpair"][im, :n] + 1, want[1][im]), im
        assert np.array_equal(ex["ergpairref"][im, :n] + 1, want[2][im]), im
    assert np.array_equal(_bits(ex["ergpair"][:int(cap["nergpair"])]), _bits(want[4]))


def test_molfit_cycles_replay_the_reference(ctx, tmp_path):
    from epsr_python_translation_amd import device as D
    s, box, tables, dt = _setup(ctx, tmp_path)
    led = D.Ledger(ctx, box)
    ncyc = RG.count("molfit")
    cap0 = RG.group("molfit.0")
    cin0 = RG.group("molfit.0.in")
    mc = D.MCFullState(ctx, s.nmol, s.nmoltype, F32(cap0["simt"]), stepr=F32(cap0["stepr"]),
                       stepm=F32(cap0["stepm"]), stepri=F32(cap0["stepri"]), stepmi=F32(cap0["stepmi"]),
                       rejrate=F32(cap0["rejrate"]), tempangwt=F32(cap0["tempangwt"]),
                       tempdihwt=F32(cap0["tempdihwt"]), ireset=int(cap0["ireset"]), intra=int(cap0["intra"]),
                       rotfreq=int(cap0["rotfreq"]), nisum=int(cap0["nisum"]), wholerot=cin0["wholerot"],
                       wholetrans=cin0["wholetrans"], internrot=cin0["internrot"], interntrans=cin0["interntrans"],
                       nstepreject=cin0["nstepreject"])
    mc.fadds, mc.faddbs = F32(cap0["fadds"]), F32(cap0["faddbs"])
    st = np.ascontiguousarray(cin0["ran1"], np.int32).copy()
    nledger = 0
    kinds = []
    for k in range(ncyc):
        cap, cin, cout = RG.group(f"molfit.{k}"), RG.group(f"molfit.{k}.in"), RG.group(f"molfit.{k}.out")
        # the program around molfit: fmolec1 at the start and after the shake (with the
        # centres folded back into the box first, epsr_standalone_rev3.py:1329-1345)
        if k == 0 or int(RG.group(f"molfit.{k - 1}.out")["nisum"]) == 1:
            f = RG.group(f"fmolec1.{nledger}")
            _set_coords(box, f)
            _fmolec1(ctx, s, box, tables, dt, led, nledger)
            nledger += 1
            # the program restarts the running energies from fmolec1's (float32) totals
            mc.fadds, mc.faddbs = F32(cap["fadds"]), F32(cap["faddbs"])
        co = box.get_coords()
        for key in ("xa", "ya", "za", "xadd", "yadd", "zadd"):
            assert np.array_equal(_bits(co[key]), _bits(cin[key])), f"cycle {k} start {key}"
        assert np.array_equal(st, cin["ran1"]), f"cycle {k} start ran1"
        assert mc.nisum == int(cap["nisum"]) and F32(mc.fadds) == F32(cap["fadds"])
        n, i, r = int(cap["nisum"]), int(cap["intra"]), int(cap["rotfreq"])
        kind = "shake" if n >= i else ("rotation" if n % r == 0 else "whole-molecule")
        kinds.append(kind)
        print(f"cycle {k}: {kind} (nisum {n})", flush=True)
        D.molfit_full(ctx, box, tables, led, mc, st, xav=dt["xav"], xav1=dt["xav1"])
        torch.cuda.synchronize()
        co = box.get_coords()
        for key in ("xa", "ya", "za", "xadd", "yadd", "zadd"):
            assert np.array_equal(_bits(co[key]), _bits(cout[key])), f"cycle {k} {key}"
        assert np.array_equal(st, cout["ran1"]), f"cycle {k} ran1"
        for key in ("wholerot", "wholetrans", "internrot", "interntrans"):
            assert np.array_equal(_bits(getattr(mc, key)), _bits(cout[key])), f"cycle {k} {key}"
        for key in ("nwhole", "nmovem", "ninter", "nmovei"):
            assert np.array_equal(getattr(mc, key).reshape(s.nmoltype, 2).T, cout[key][1:3]), f"cycle {k} {key}"
        assert np.array_equal(mc.nstepreject.cpu().numpy(), cout["nstepreject"]), f"cycle {k} nstepreject"
        ex = led.export()
        ne = int(cout["nergpair"])
        assert np.array_equal(_bits(ex["ergpair"][:ne]), _bits(cout["ergpair"])), f"cycle {k} ledger"
        assert F32(mc.fadds) == F32(cout["fadds"]) and F32(mc.faddbs) == F32(cout["faddbs"]), f"cycle {k} fadds"
        assert mc.nisum == int(cout["nisum"])
    assert set(kinds) == {"shake", "rotation", "whole-molecule"}
