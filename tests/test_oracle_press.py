"""The oracle's press (oracle/ora_press.c), pinned by what the mathematics fixes:

* no interactions (zero tables): energy 0, virial 0, pressure = the ideal gas's
  N kT / V converted to kbar (100 / 6.0222 per kJ mol^-1 A^-3);
* each pair's energy equals the ledger's entry for that pair (en_mole_pair, a separate
  routine with its own association of the same terms) to rounding;
* the virial is the derivative of the energy under a uniform scaling of the molecule
  centres (atoms move with their centres): presst = dE/ds at s = 1, checked by central
  differences on the reference's own test box; a dropped prfac, a sign or a swapped
  separation fails it;
* an independent float64 evaluation (numpy, all pairs) on a small box."""

import ctypes as C

import numpy as np

from oracle import pyoracle as po
from tests import helpers as H

F32 = np.float32


def test_ideal_gas_without_interactions():
    name = "tiny_12_6_2"
    s = H.system(name)
    t = dict(H.oracle_tables(name))
    for k in ("xavs", "xavs1", "ecoul", "ecoul1"):
        t[k] = np.zeros_like(t[k])
    r = H.oracle_press(name, tables=t)
    assert r["energy"] == 0.0 and r["presst"] == 0.0
    want = F32(float(F32(s.nmol) * F32(s.simt)) * 100.0 / float(F32(6.0222) * H.box_volume(s)))
    assert r["pressc"] == want
    # N kT / V: kJ mol^-1 A^-3 -> J m^-3 (1e3 / N_A / 1e-30) -> kbar (1e-8)
    kbar = s.nmol * 0.008314 * s.simtk / float(H.box_volume(s)) * 1e3 / 6.0222e23 / 1e-30 * 1e-8
    assert abs(float(r["pressc"]) - kbar) < 1e-4 * kbar


def test_pair_energies_equal_the_ledger_entries():
    name = "epsr_test"
    r = H.oracle_press(name)
    led = H.ledger_dict(r["ref"]["ledger"], H.system(name).nmol)
    worst = 0.0
    for m in range(led["nergpairm"].shape[0]):
        for j in range(led["nergpairm"][m]):
            n = led["molergpair"][m, j]
            if n <= m:
                assert r["slote"][m, j] == 0.0 and r["slotw"][m, j] == 0.0
                continue
            e_led = float(led["ergpair"][led["ergpairref"][m, j]])
            e = float(r["slote"][m, j])
            worst = max(worst, abs(e - e_led))
    # up to 100 atom-pair terms of up to ~10^2 kJ/mol each, float32 (u = 6e-8): the two
    # associations differ by rounding only (seen: 3.8e-5); a dropped term is >= 1e-2
    assert worst < 1e-4, worst


def test_virial_is_the_derivative_under_centre_scaling():
    """With tables whose interpolation is U(r) = r (slope 1 in every bin, so the two-bin
    slope average press uses is exact), each pair's virial equals the derivative of the
    pair's energy (en_mole_pair, another routine) when the partner's centre separation is
    scaled: W = sum (Rc . r)/|r| = dE/ds."""
    name = "epsr_test"
    s, g = H.system(name), H.grid(name)
    t = dict(H.oracle_tables(name))
    ld, npairs = t["ld"], s.npairs
    lin = np.zeros(npairs * ld, F32)
    one = np.zeros(npairs * ld, F32)
    for ic in range(npairs):
        lin[ic * ld + 1:ic * ld + g.npt1 + 1] = t["xstar"][1:g.npt1 + 1]
        one[ic * ld + 1:ic * ld + g.npt1 + 1] = 1.0
    t.update(xavs=lin, xavs1=one, ecoul=np.zeros_like(lin), ecoul1=np.zeros_like(lin))
    ref = H.oracle_fmolec1(name)
    r = H.oracle_press(name, ref=ref, tables=t)
    led = H.ledger_dict(ref["ledger"], s.nmol)
    ob, tt, L = ref["ob"], H.ctables(t), po.lib()
    h = 1e-3
    checked = 0
    rng = np.random.default_rng(3)
    for m in rng.choice(s.nmol - 1, 60, replace=False):
        for j in range(led["nergpairm"][m]):
            n = int(led["molergpair"][m, j])
            if n <= m:
                continue
            cm = np.array([ob.xadd[m], ob.yadd[m], ob.zadd[m]], np.float64)
            cn = np.array([ob.xadd[n], ob.yadd[n], ob.zadd[n]], np.float64)
            if np.linalg.norm(cn - cm) >= 0.25:
                continue    # no image change, every atom pair well inside the last bin npt1
            e = []
            for sc in (1.0 + h, 1.0 - h):
                p = (cm + sc * (cn - cm)).astype(F32)
                ob.xadd[n], ob.yadd[n], ob.zadd[n] = p
                e.append(L.ora_en_mole_pair(C.byref(ob.box), C.byref(ob.centres), C.byref(tt), int(m), n))
            ob.xadd[n], ob.yadd[n], ob.zadd[n] = cn.astype(F32)
            fd = (e[0] - e[1]) / (2 * h)
            w = float(r["slotw"][m, j])
            na = (s.nal[m] - s.naf[m] + 1) * (s.nal[n] - s.naf[n] + 1)
            assert abs(fd - w) <= 2e-3 * na * 0.5 + 1e-4 * abs(w), (m, n, fd, w)
            checked += 1
    assert checked > 500

    # the sign convention of the pressure: U increasing with r (an attraction) gives a
    # positive virial sum and a pressure below the ideal gas's
    ideal = float(F32(s.nmol) * F32(s.simt)) * 100.0 / float(F32(6.0222) * H.box_volume(s))
    assert r["presst"] > 0.0 and r["pressc"] < ideal


def test_against_an_independent_float64_evaluation():
    name = "tiny_12_6_2"
    s, g = H.system(name), H.grid(name)
    t = H.oracle_tables(name)
    r = H.oracle_press(name)
    ld = t["ld"]
    stepcl = np.float64(F32(s.cellst) / F32(s.dmax))
    X = np.stack([s.xa, s.ya, s.za], 1).astype(np.float64)
    Cc = np.stack([s.xadd, s.yadd, s.zadd], 1).astype(np.float64)
    e_tot = 0.0
    w_tot = 0.0
    mag = 0.0
    for m in range(s.nmol):
        for n in range(m + 1, s.nmol):
            rc = Cc[n] - Cc[m]
            rc -= np.where(rc >= 0, np.floor(rc + 0.5), np.ceil(rc - 0.5))
            for ia in range(s.naf[m], s.nal[m] + 1):
                for ja in range(s.naf[n], s.nal[n] + 1):
                    d = (X[ja] + Cc[n]) - (X[ia] + Cc[m])
                    d -= np.where(d >= 0, np.floor(d + 0.5), np.ceil(d - 0.5))
                    r_ = float(np.sqrt(d @ d))
                    iz = int(r_ / stepcl) + 2
                    if iz > g.npt1:
                        continue
                    o = int(s.npairref[s.itype[ia], s.itype[ja]]) * ld
                    dt = r_ - float(t["xstar"][iz])
                    e = (t["xavs"][o + iz] + t["xavs1"][o + iz] * dt) + (t["ecoul"][o + iz] + t["ecoul1"][o + iz] * dt)
                    slope = 0.5 * (t["xavs1"][o + iz] + t["xavs1"][o + iz - 1]) + \
                        0.5 * (t["ecoul1"][o + iz] + t["ecoul1"][o + iz - 1])
                    e_tot += float(e)
                    mag += abs(float(e))
                    w_tot += float(slope) * float(rc @ d) / r_
    assert abs(float(r["energy"]) * s.nmol - e_tot) <= 1e-5 * mag
    assert abs(r["presst"] - w_tot) <= 1e-5 * abs(w_tot) + 1e-3
