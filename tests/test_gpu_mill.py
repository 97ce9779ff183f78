"""The ball mill on the device: lifters turned inside every step, the
dissipation tallies, and the reference's known answer.

The deck is the reference's own Mill_SPH30 (242 steel balls, a 90 cm drum
with 8 lifters, 22 RPM).  The library turns the lifters on the device before
each step (bdem_set_lifter_rotation); the oracle turns them on the host as
Rotate_VolObject_Inc does (Input/Processing_Mills.h:7-60).  Tallies are added
one contact at a time with float atomics on the device and in particle order
in the oracle, so they agree to rounding, not bit for bit.
"""
import os
import time

import numpy as np
import pytest

from bdem.deck import Deck
from helpers import rel_err, by_pid
from oracle_py import Oracle

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DECKS = os.path.join(ROOT, "tests", "golden", "decks")
PUBLISHED = dict(particle=319.724, lifter=156.378, drum=149.878, total=625.98)
TOL = 1e-5


def _mill(drum_moving=True):
    d = Deck(os.path.join(DECKS, "3.BallMill"), "Mill_SPH30")
    dev = d.device()
    d.load_state(dev)
    dev.reset_energy()
    dev.drum_rotation(drum_moving)
    a = d.info.rotate_about
    dev.set_lifter_rotation(d.info.rad_per_step, (a.x, a.y, a.z))
    return d, dev


def _v(f):
    return (f.x, f.y, f.z)


def test_device_lifter_turn_is_the_host_turn():
    print("[case] lifter turn 2000 steps", flush=True)
    d, dev = _mill()
    sc = d.scenario()
    ora = Oracle(sc)
    dev.step(2000)
    for _ in range(2000):
        ora.turn_lifters(sc["rad_per_step"], sc["lifter_about"])
    got, want = dev.get_dobj(), ora.get_dobj()
    assert len(got) == len(want) == 8
    for g, w in zip(got, want):
        assert _v(g.COM) == _v(w.COM)
        assert [_v(g.vertex[k]) for k in range(8)] == [_v(w.vertex[k]) for k in range(8)]
        for j in range(g.num_faces):
            assert _v(g.faces[j].normal) == _v(w.faces[j].normal)
            assert _v(g.faces[j].centroid) == _v(w.faces[j].centroid)
            assert g.faces[j].area == w.faces[j].area
    # 2000 steps of RadPerStep: the lifters moved by 2000 |RadPerStep| about the axis
    a0 = np.arctan2(d.dobj[0].COM.y - 45.0, d.dobj[0].COM.x - 45.0)
    a1 = np.arctan2(got[0].COM.y - 45.0, got[0].COM.x - 45.0)
    assert abs((a0 - a1) - 2000 * d.info.rad_per_step) < 1e-4
    dev.close()
    ora.close()


def test_mill_steps_and_tallies_match_oracle():
    print("[case] mill 400 steps with tallies", flush=True)
    d, dev = _mill()
    sc = d.scenario()
    ora = Oracle(sc)
    ora.set_drum_rotation(True)
    dev.step(400)
    ora.run(400, sc["rad_per_step"], sc["lifter_about"])
    sg, so = by_pid(dev.get_state()), by_pid(ora.state())
    for key in ("pos", "vel", "avel"):
        err = rel_err(sg[key], so[key], 1e-3)
        print(f"mill {key}: max rel err {err:.3e}")
        assert err <= TOL, key
    eg = dev.energy().astype(np.float64)
    eo, _ = ora.energy()
    print(f"tallies (J): device {eg}, oracle {eo}")
    assert np.all(eo > 0)
    assert np.allclose(eg, eo, rtol=1e-4, atol=0)
    dev.close()
    ora.close()


def _revolution(drum_moving):
    d, dev = _mill(drum_moving)
    steps = d.info.steps_per_rev
    t0 = time.perf_counter()
    dev.step(steps)
    dev.sync()
    dt = time.perf_counter() - t0
    e = dev.energy().astype(np.float64)
    p = e / (steps * d.sim.InitalDelta_t)
    print(f"drum {'moving' if drum_moving else 'static'}: {steps} steps in {dt:.2f} s "
          f"({steps / dt:.0f} steps/s): particle {p[0]:.2f} lifter {p[1]:.2f} drum {p[2]:.2f} "
          f"total {p.sum():.2f} W", flush=True)
    assert dev.status() & 1 == 0
    dev.close()
    return p


def test_mill_power_one_revolution():
    """The known answer (Projects/3.BallMill/ReadMe.txt:5-6); tolerances
    cover the run-to-run spread of this chaotic system.  As shipped, the
    drum contact moves at MillRotVel = -|omega R| * 0.01 (main.cpp:717-721):
    the particle and lifter powers match the published ones and the drum
    power is about 36 % low.  With the drum contact static (lifters still
    turning) all three powers and the total match the published figures,
    so those come from a build whose drum contact did not move; both
    semantics are asserted."""
    print("[case] mill one revolution", flush=True)
    p = _revolution(True)
    assert abs(p[0] / PUBLISHED["particle"] - 1) < 0.04
    assert abs(p[1] / PUBLISHED["lifter"] - 1) < 0.10
    assert 80.0 < p[2] < 105.0          # the shipped drum contact: about 92 W
    q = _revolution(False)
    assert abs(q.sum() / PUBLISHED["total"] - 1) < 0.04
    assert abs(q[0] / PUBLISHED["particle"] - 1) < 0.04
    assert abs(q[2] / PUBLISHED["drum"] - 1) < 0.08
