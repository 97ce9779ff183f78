"""Ensemble statistics on the device (bdem_get_particle_stats) and the silo
discharge driver, on the reference's own silo deck.

The statistics are reductions of the state, so they are checked against
numpy on the state the library returns: counts exactly, the fastest
particle exactly, sums to double rounding.
"""
import math
import os

import numpy as np
import pytest

import bdem
from bdem import scenarios as S
from bdem.deck import Deck, run_silo
from helpers import make_device, surface_profile_np
from oracle_py import Oracle

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DECKS = os.path.join(ROOT, "tests", "golden", "decks")


def _check(dev, min_y, box, radius_of):
    st = dev.get_state()
    live = st["pid"] >= 0
    p = st["pos"].astype(np.float64)
    v = st["vel"].astype(np.float32) * np.float32(0.010)
    sp = np.sqrt((v.astype(np.float32) ** 2).sum(axis=1, dtype=np.float32)).astype(np.float64)
    above = live & (st["pos"][:, 1] > np.float32(min_y))
    got = dev.particle_stats(min_y, box)
    assert got["count"] == int(above.sum())
    if above.any():
        assert got["mean_speed"] == pytest.approx(sp[above].mean(), rel=1e-6)
        k = np.argmax(np.where(above, sp, -1.0))
        assert got["max_speed"] == pytest.approx(sp[k], rel=1e-6)
        assert sp[st["pid"] == got["max_pid"]][0] == pytest.approx(sp[k], rel=1e-6)
    lo, hi = np.array(box[0]), np.array(box[1])
    f = st["pos"]
    inside = live & np.all((f >= lo.astype(np.float32)) & (f < hi.astype(np.float32)), axis=1)
    assert got["box_count"] == int(inside.sum())
    r = radius_of(st)
    vol = (4.0 / 3.0) * np.pi * (r[inside].astype(np.float64) ** 3)
    assert got["box_solid_fraction"] == pytest.approx(vol.sum() / np.prod(hi - lo), rel=1e-9)
    return got


def test_stats_on_a_dense_random_pack():
    print("[case] stats random pack", flush=True)
    deck = S.random_deck(20000, seed=5, L=30.0, polydisperse=True)
    dev = make_device(deck)
    dev.step(20)
    rad = deck["radius"]
    order_r = lambda st: rad[st["pid"]]
    g = _check(dev, 10.0, ((5.0, 5.0, 5.0), (20.0, 25.0, 20.0)), order_r)
    assert g["count"] > 0 and 0.0 < g["box_solid_fraction"] < 0.75
    _check(dev, -np.inf, ((0.0, 0.0, 0.0), (30.0, 30.0, 30.0)), order_r)
    dev.close()


def test_silo_deck_discharge():
    print("[case] silo SPH_30 discharge", flush=True)
    d = Deck(os.path.join(DECKS, "2.Silo"), "SPH_30")
    dev = d.device()
    d.load_state(dev)
    rad = np.array([p.radius for p in d.pobj], np.float32)
    by_type = lambda st: rad[st["ptype"]]
    n_above = _check(dev, d.info.hatch_height, ((0.0, 30.0, 0.0), (14.0, 60.0, 5.0)), by_type)["count"]
    assert n_above == d.sim.Num_Particles          # the start file holds everything above the hatch
    curve = run_silo(dev, d, 0.4)
    pct = [c[1] for c in curve]
    print("discharged % at", [f"{t:.3f}s {p:.1f}" for t, p in curve])
    assert all(b >= a for a, b in zip(pct, pct[1:]))   # particles only leave (the floor is below the hatch)
    assert pct[-1] > 0.0
    _check(dev, d.info.hatch_height, ((0.0, 0.0, 0.0), (14.0, 30.0, 5.0)), by_type)
    dev.close()


def _same(a, b):
    return np.array_equal(np.isnan(a), np.isnan(b)) and np.array_equal(a[~np.isnan(a)], b[~np.isnan(b)])


def test_surface_profile_is_numpys():
    """Per-bin highest top y + r equals numpy's on the returned state exactly,
    along x and z, with ranges that cut the bed and bins left empty."""
    print("[case] surface profile, collapse deck 3000 steps", flush=True)
    deck = S.collapse_deck()
    dev = make_device(deck)
    dev.step(3000)
    st = dev.get_state()
    live = st["pid"] >= 0
    for axis, lo, hi, bins in ((0, 0.0, 20.0, 20), (0, -2.5, 9.3, 37), (2, 3.0, 17.0, 4096), (0, 4.0, 5.0, 1)):
        got = dev.surface_profile(lo, hi, bins, axis)
        want = surface_profile_np(st["pos"], 0.5, lo, hi, bins, axis, live)
        assert _same(got, want), (axis, lo, hi, bins)
    with pytest.raises(bdem.lib.BdemError):
        dev.surface_profile(0.0, 1.0, 8, axis=1)
    with pytest.raises(bdem.lib.BdemError):
        dev.surface_profile(0.0, 1.0, 4097)
    dev.close()


@pytest.mark.parametrize("deg", [20.0, 35.0])
def test_inclined_layer_gives_its_angle(deg):
    """One sphere per bin with its centre on y = 8 + tan(deg) x: the fitted
    free surface has that slope."""
    print(f"[case] inclined layer {deg} deg", flush=True)
    deck = S.collapse_deck()
    n = len(deck["pos"])
    lo, hi, bins = 1.0, 19.0, 18
    x = lo + (np.arange(bins) + 0.5) * (hi - lo) / bins
    pos = deck["pos"].copy()
    pos[:, 1] = 1.0                       # everything else low: a floor layer
    pos[:bins, 0] = x
    pos[:bins, 1] = 8.0 + math.tan(math.radians(deg)) * (x - 10.0)
    pos[:bins, 2] = 10.0
    dev = make_device(dict(deck, pos=pos))
    top = dev.surface_profile(lo, hi, bins)
    slope, _, ang = bdem.lib.surface_angle(top, lo, hi)
    assert ang == pytest.approx(deg, abs=1e-3)
    assert n > bins
    dev.close()


def test_column_collapse_statistics_match_fp64_oracle():
    """A long run compared by ensemble statistics: 1,596 spheres slump from a
    column into a ramp over 50,000 steps.  The device (float) and the FP64
    oracle part ways after a few thousand steps, as a chaotic flow must, but
    the ramp's angle of repose, the solid fraction of the bed and the mean
    height agree."""
    print("[case] column collapse 50000 steps vs FP64 oracle", flush=True)
    deck = S.collapse_deck()
    dev = make_device(deck)
    dev.step(50000)
    ora = Oracle(deck, bits=64)
    ora.step(50000)
    lo, hi, bins = 2.0, 14.0, 12
    res = {}
    for name, st in (("device", dev.get_state()), ("fp64", ora.state())):
        top = surface_profile_np(st["pos"], 0.5, lo, hi, bins)
        _, _, ang = bdem.lib.surface_angle(top, lo, hi)
        p = np.asarray(st["pos"], np.float64)
        inside = (p[:, 0] >= 2) & (p[:, 0] < 12) & (p[:, 1] >= 0) & (p[:, 1] < 3) & (p[:, 2] >= 0) & (p[:, 2] < 20)
        phi = inside.sum() * (4.0 / 3.0) * math.pi * 0.125 / (10.0 * 3.0 * 20.0)
        res[name] = (ang, phi, p[:, 1].mean(), np.abs(np.asarray(st["vel"], np.float64)).max())
        print(f"{name}: angle {ang:.2f} deg, bed solid fraction {phi:.3f}, mean height {p[:, 1].mean():.3f} cm, "
              f"max |v| {res[name][3]:.2f} cm/s", flush=True)
    d, o = res["device"], res["fp64"]
    got = dev.particle_stats(-np.inf, ((2, 0, 0), (12, 3, 20)))
    assert got["box_solid_fraction"] == pytest.approx(d[1], rel=1e-6)
    assert 10.0 < d[0] < 30.0
    assert abs(d[0] - o[0]) < 2.5
    assert abs(d[1] - o[1]) < 0.03
    assert abs(d[2] / o[2] - 1.0) < 0.03
    dev.close()
    ora.close()
