"""atogen on the GPU against the oracle: bit-exact histograms, potex, ngrsamples and
the ran1 state left behind."""

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


def _gpu_atogen(ctx, name, ngrsamples, idum, rminex=None):
    from epsr_python_translation_amd import device as D
    s, g = H.system(name), H.grid(name)
    box = D.DeviceBox.from_system(ctx, s)
    ldw = g.npt1 + 1
    st = D.ran1_state_array(idum)
    rminex = np.zeros(s.npairs, F32) if rminex is None else rminex
    counts = ctx.zeros(2 * s.npairs * ldw, torch.int32)
    wav, potex, _, ngr = D.atogen(ctx, box, g.npt1, ldw, ctx.tensor(g.ycoh),
                                  ctx.tensor(np.arange(s.npairs, dtype=np.int32)), ctx.tensor(rminex),
                                  ngrsamples, st, counts=counts)
    torch.cuda.synchronize()
    return dict(wav=wav.cpu().numpy(), potex=potex.cpu().numpy(), ngr=ngr, ran1=st, counts=counts.cpu().numpy())


@pytest.mark.parametrize("name,ngr,idum", [
    ("tiny_11_10_0", 0, -1),          # every molecule an origin (addwav 2)
    ("tiny_12_6_2", 0, -5),
    ("tiny_14_40_0", 7, -99),         # sampled origins (addwav 1), under-represented types forced
    ("c0", 0, -12345),                # 200 DMSO + 400 water, all 600 molecules
    ("c0", 150, -31),                 # 3200 atoms, 150 sampled origins
    ("epsr_test", 0, -7),             # the reference's EPSR test box read from its .ato
    ("epsr_test", 200, -3),
    ("fine_7_30_6", 0, -19),          # 21 pairs x 2001 bins: the pair columns split into LDS-sized groups
    ("fine_7_30_6", 12, -23),
])
@pytest.mark.parametrize("close", [True, False], ids=["close-contacts", "no-close-contacts"])
def test_atogen_bit_exact(ctx, name, ngr, idum, close):
    """With every rminex 2.2 the kernel counts close contacts; with none it takes the path
    without them (and, at most 4 atom types, the packed histogram offsets)."""
    print(f"atogen parity {name} ngr={ngr} close={close}", flush=True)
    s = H.system(name)
    rminex = np.full(s.npairs, 2.2 if close else 0.0, F32)
    ref = H.oracle_atogen(name, ngrsamples=ngr, idum=idum, rminex=rminex)
    got = _gpu_atogen(ctx, name, ngr, idum, rminex=rminex)
    assert got["ngr"] == ref["ngr"]
    np.testing.assert_array_equal(got["ran1"], ref["ran1"])
    g = H.grid(name)
    sl = np.s_[:, :, 1:g.npt1 + 1]
    w_ref = ref["wav"].reshape(2, s.npairs, -1)[sl]
    w_got = got["wav"].reshape(2, s.npairs, -1)[sl]
    assert np.array_equal(w_ref.view(np.int32), w_got.view(np.int32))
    assert np.array_equal(ref["potex"].view(np.int32), got["potex"].view(np.int32))


def test_atogen_large_box_sampled_bit_exact(ctx):
    """60,000-atom water box with the reference default of 1000 origin molecules."""
    print("atogen parity c1 ngr=1000", flush=True)
    ref = H.oracle_atogen("c1", ngrsamples=0, idum=-4242)
    got = _gpu_atogen(ctx, "c1", 0, -4242)
    assert got["ngr"] == ref["ngr"] == 1000
    np.testing.assert_array_equal(got["ran1"], ref["ran1"])
    s, g = H.system("c1"), H.grid("c1")
    sl = np.s_[:, :, 1:g.npt1 + 1]
    assert np.array_equal(ref["wav"].reshape(2, s.npairs, -1)[sl].view(np.int32),
                          got["wav"].reshape(2, s.npairs, -1)[sl].view(np.int32))


def test_atogen_large_box_full_mode_invariants(ctx):
    """Every pair of the 60,000-atom box (1.8e9 distances): exact pair-count invariants.
    Over all bins, counts of each pair type equal n_i n_j (or n_i(n_i-1)/2 less the
    intramolecular pairs) minus the pairs beyond the last bin, which are counted
    exactly by the histogram of the same kernel with a wider range."""
    print("atogen invariants c1 full", flush=True)
    s, g = H.system("c1"), H.grid("c1")
    got = _gpu_atogen(ctx, "c1", s.nmol, -3)
    c = got["counts"].reshape(2, s.npairs, -1).astype(np.int64)
    inter = c[1].sum()
    # all molecule pairs x 9 atom pairs, minus those beyond npt1
    total_pairs = s.nmol * (s.nmol - 1) // 2 * 9
    assert 0.45 * total_pairs < inter <= total_pairs
    # the 2:1 H/O ratio fixes the partial sums of a homogeneous box only statistically,
    # but O-H and H-O orderings fold into one pair type, so OH counts ~ 4 x OO counts
    oo, oh, hh = c[1].sum(axis=1)
    assert abs(oh / oo - 4.0) < 0.05 and abs(hh / oo - 4.0) < 0.05
    # the normalised g(r) tends to 1 at large r (uniform fluid), as the reference's
    # normalisation by ycoh implies
    wav = got["wav"].reshape(2, s.npairs, -1)
    tail = wav[1, :, g.npt1 - 200:g.npt1 + 1] / g.ycoh[None, g.npt1 - 200:g.npt1 + 1]
    assert np.all(np.abs(tail.mean(axis=1) - 1.0) < 0.02)


def test_atogen_large_box_full_mode_bit_exact(ctx):
    """Every pair of the 60,000-atom box (1.8e9 distances, every molecule an origin), with
    close contacts counted: the histograms, potex and the ran1 state against the all-cores
    oracle (the sequential oracle's counts, origin molecules spread over threads)."""
    print("atogen parity c1 full (oracle on all cores)", flush=True)
    s, g = H.system("c1"), H.grid("c1")
    rminex = np.full(s.npairs, 2.2, F32)
    ref = H.oracle_atogen("c1", ngrsamples=s.nmol, idum=-77, rminex=rminex, flavour="omp")
    got = _gpu_atogen(ctx, "c1", s.nmol, -77, rminex=rminex)
    assert got["ngr"] == ref["ngr"] == s.nmol
    np.testing.assert_array_equal(got["ran1"], ref["ran1"])
    sl = np.s_[:, :, 1:g.npt1 + 1]
    assert np.array_equal(ref["wav"].reshape(2, s.npairs, -1)[sl].view(np.int32),
                          got["wav"].reshape(2, s.npairs, -1)[sl].view(np.int32))
    assert np.array_equal(ref["potex"].view(np.int32), got["potex"].view(np.int32))
    assert got["potex"].sum() > 0


@pytest.mark.parametrize("name", ["c1", "c0", "epsr_test", "fine_7_30_6"])
def test_atogen_binning_exhaustive(ctx, name):
    """Every float32 squared distance from 0 past the last counted bin (about 1e9 values for
    the 60,000-atom box): the pair kernel's bin (hardware square root, exact thresholds near
    bin edges) and close-contact test equal the correctly rounded NINT(sqrtf(dsq) / stepcl)
    and sqrtf(dsq) * dmax < 2.2."""
    from epsr_python_translation_amd import device as D
    s, g = H.system(name), H.grid(name)
    stepcl = F32(F32(s.cellst) / F32(s.dmax))
    top = F32(F32(g.npt1 + 3) * stepcl)
    hi = int(np.array(F32(top * top * F32(1.01)), F32).view(np.uint32))
    print(f"atogen binning sweep {name}: stepcl={stepcl} npt1={g.npt1} floats={hi}", flush=True)
    bad, seen, edge = D.atogen_bin_check(ctx, stepcl, g.npt1, F32(s.dmax), F32(2.2), 0, hi)
    print(f"  {seen} values, {edge} near a bin edge, {bad} disagreements", flush=True)
    assert seen == hi and edge >= 100 * g.npt1 and bad == 0
