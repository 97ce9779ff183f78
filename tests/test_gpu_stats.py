"""Ensemble statistics on the device (bdem_get_particle_stats) and the silo
discharge driver, on the reference's own silo deck.

The statistics are reductions of the state, so they are checked against
numpy on the state the library returns: counts exactly, the fastest
particle exactly, sums to double rounding.
"""
import os

import numpy as np
import pytest

from bdem import scenarios as S
from bdem.deck import Deck, run_silo
from helpers import make_device

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
