"""Dissipation tallies on the CPU, pinned by energy balance.

The tallies (Tally_EnergyDis_PP / DisLift / DisSurf, GPU_Memory_Transfer.cuh:
97-99) add each contact's dissipated work.  Energy conservation fixes what
they must sum to: the kinetic energy a system loses through its contacts.
The reference defines the increments for its own law
(Collision_Response.cuh:25-89, 401-475); the history laws use this
library's definition (damping work, the spring energy a clipped normal force
does not return, and sliding friction work).  A wrong factor (the half per
side of a pair, dt) or a missing term fails these checks:

  - head-on pair impact, no friction, no gravity: the pair tally equals the
    kinetic energy lost, for the reference law, LSD and Hertz-Mindlin;
  - a sphere sliding on the floor until it rolls (LSD, Hertz-Mindlin): the
    loss is 2/7 of the initial kinetic energy (a solid sphere rolls at 5/7
    of its sliding speed) and the surface tally equals it;
  - a cube landing face down on the floor without friction (the reference
    law's face-face wall contact, no moment): the surface tally equals the
    kinetic energy lost;
  - polyhedra contacts are tallied: pairs, lifters and walls of a cube mill.

All in FP64 with dt = 2e-6 s (125+ steps per contact), where the discrete
work matches the energy change to better than 1 %.
"""
import numpy as np
import pytest

from bdem import scenarios as S
from bdem import types as T
from oracle_py import Oracle

DT = 2e-6


def _head_on(law, e=0.6):
    d = S.random_deck(2, seed=1, L=20.0, law=law)
    d["pos"] = np.array([[10, 10, 10], [11.002, 10, 10]], np.float32)
    d["vel"] = np.array([[20, 0, 0], [-20, 0, 0]], np.float32)
    d["avel"] = np.zeros((2, 3), np.float32)
    d["sim"].Force_Field = T.f3(0, 0, 0)
    d["sim"].EnergyCalc = True
    d["sim"].InitalDelta_t = DT
    if law == "ref":
        d["pobj"][0] = S.sphere_pobj(0.5, 2500.0, e, 1e5, 0.0)
    return d


@pytest.mark.parametrize("law", ["ref", "lsd", "hm"])
def test_pair_tally_is_the_kinetic_energy_lost(law):
    d = _head_on(law)
    o = Oracle(d, bits=64)
    m = float(d["pobj"][0].mass)
    k0 = 0.5 * m * np.sum(o.state()["vel"] ** 2)
    o.step(3000)                                   # 6 ms: the impact is over
    k1 = 0.5 * m * np.sum(o.state()["vel"] ** 2)
    _, exact = o.energy()
    assert k1 < 0.5 * k0                           # a real, damped impact
    assert exact[0] == pytest.approx(k0 - k1, rel=1e-2)
    assert exact[1] == 0.0 and exact[2] == 0.0
    o.close()


@pytest.mark.parametrize("law", ["lsd", "hm"])
def test_sliding_sphere_surface_tally_is_two_sevenths(law):
    d = S.random_deck(1, seed=1, L=20.0, law=law)
    d["sim"].EnergyCalc = True
    d["sim"].InitalDelta_t = DT
    d["sim"].Roll_Res = 0.0
    r, m = 0.5, float(d["pobj"][0].mass)
    d["avel"] = np.zeros((1, 3), np.float32)
    d["vel"] = np.zeros((1, 3), np.float32)
    d["pos"] = np.array([[10, r - 0.0005, 10]], np.float32)
    o = Oracle(d, bits=64)
    o.step(20000)                                   # come to rest on the floor
    y = o.state()["pos"][0][1]
    o.close()
    d["pos"] = np.array([[10, y, 10]], np.float32)
    d["vel"] = np.array([[50, 0, 0]], np.float32)
    o = Oracle(d, bits=64)
    k0 = 0.5 * m * 50.0 ** 2
    o.step(40000)                                   # slides, then rolls
    st = o.state()
    v, w = st["vel"][0], st["avel"][0]
    assert abs(w[2] * r + v[0]) < 1e-2 * v[0]       # rolling, v = -w r (the LSD spring still rings)
    k1 = 0.5 * m * np.sum(v ** 2) + 0.5 * (0.4 * m * r * r) * np.sum(w ** 2)
    _, exact = o.energy()
    assert k0 - k1 == pytest.approx(2.0 / 7.0 * k0, rel=1e-3)
    assert exact[2] == pytest.approx(k0 - k1, rel=1e-3)
    o.close()


def test_cube_landing_face_down_surface_tally():
    q = np.array([1, 0, 0, 0], np.float32)
    d = S.poly_pair_deck((5, 0.6, 5), (9, 5, 9), q, q, va=(0, -80, 0), mu=0.0, gravity=False)
    for p in d["pobj"]:
        p.surface_Fric_static = 0.0
        p.surface_Fric_kinetic = 0.0
    d["sim"].EnergyCalc = True
    d["sim"].InitalDelta_t = DT
    o = Oracle(d, bits=64)
    m = float(d["pobj"][0].mass)
    k0 = 0.5 * m * 80.0 ** 2
    o.step(1500)                                    # 3 ms: bounced off
    st = o.state()
    a = int(np.nonzero(st["pid"] == 0)[0][0])
    assert st["vel"][a][1] > 10.0                   # moving up again
    assert np.abs(st["avel"][a]).max() == 0.0       # face-face: no moment
    k1 = 0.5 * m * np.sum(st["vel"][a] ** 2)
    _, exact = o.energy()
    assert exact[2] == pytest.approx(k0 - k1, rel=1e-2)
    o.close()


def test_polyhedra_contacts_are_tallied():
    d = S.poly_mill_deck(300)
    d["sim"].EnergyCalc = True
    o = Oracle(d)
    o.set_drum_rotation(True)
    o.run(40, d["rad_per_step"], d["lifter_about"])
    e, exact = o.energy()
    assert np.all(exact > 0.0)                      # pairs, lifters and the drum wall
    assert np.allclose(e, exact * 1e-4, rtol=1e-3)  # float sums of the same increments, x 1E-4 (J)
    o.close()
