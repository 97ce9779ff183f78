"""Parity at the headline size: the 1,048,576-sphere silo (Hertz-Mindlin
with tangential history, radii U[0.2, 0.4] cm, cylinder over a 30-degree
cone, outlet closed).

The device settles the fill with the outlet closed, as the benchmark does
(24,000 steps at dt = 2e-5 s), until the pack is in contact (about five
contacts per particle); that state, with empty spring history, then starts
both a fresh context and the oracle at the configured dt.
After 3 more steps every position, velocity and spin must agree to 1e-5
(observed bit-exact) and the contact lists of all particles must be equal.
"""
import numpy as np
import pytest

from bdem import scenarios as S
from helpers import make_pair, make_device, rel_err, by_pid, oracle_positions_out

pytestmark = pytest.mark.gpu

N = 1048576
TOL = 1e-5


def test_full_size_silo_matches_oracle():
    deck = S.silo_deck(N)
    dev = make_device(deck)
    sim = deck["sim"]
    dt = sim.InitalDelta_t
    sim.InitalDelta_t = 2e-5
    dev.update_sim_info(sim)
    for done in range(2000, 24001, 2000):
        dev.step(2000)
        dev.sync()
        print(f"[case] full-size silo: settle {done}/24000", flush=True)
    sim.InitalDelta_t = dt
    st = by_pid(dev.get_state())
    dev.close()
    assert np.array_equal(st["pid"], np.arange(N, dtype=np.int32))
    start = dict(deck, sim=sim, pos=st["pos"], vel=st["vel"], avel=st["avel"])
    print("[case] full-size silo: 3 steps, device and oracle", flush=True)
    dev, ora = make_pair(start)
    lg, ng = dev.debug_contacts()
    lo, no = ora.contacts()
    assert np.array_equal(ng, no)
    k = float(no.mean())
    print(f"mean contacts per particle {k:.3f}", flush=True)
    assert k > 4.0
    m = np.minimum(no, 32)
    cols = np.arange(32)[None, :] < m[:, None]
    assert np.array_equal(np.where(cols, lg, 0), np.where(cols, lo, 0))
    dev.step(3)
    ora.step(3)
    sg, so = by_pid(dev.get_state()), by_pid(ora.state())
    for key in ("pos", "vel", "avel"):
        err = rel_err(sg[key], so[key], 1e-3)
        exact = float(np.mean(sg[key].astype(np.float64) == so[key]))
        print(f"silo 1M 3 steps {key}: max rel err {err:.3e}, bit-exact fraction {exact:.6f}", flush=True)
        assert err <= TOL, key
    dev.close()
    ora.close()


def test_full_size_hull_drum_matches_oracle():
    """The 200,000-hull lifter drum of the polyhedral mill configuration
    (each hull its own shape row; hulls start in contact, about 13 listed
    candidates each): the device and the oracle from the same state, lifters
    turning.  Contact lists bit-exact; after 2 steps positions, velocities,
    spins and orientations within 1e-5 (sin/cos of the quaternion update
    differ between device and host libm)."""
    print("[case] full-size hull drum: building 200,000 hulls", flush=True)
    deck = S.hull_mill_deck(200000)
    dev, ora = make_pair(deck)
    for d in (dev,):
        d.drum_rotation(True)
        d.set_lifter_rotation(deck["rad_per_step"], deck["lifter_about"])
    ora.set_drum_rotation(True)
    lg, ng = dev.debug_contacts()
    lo, no = ora.contacts()
    assert np.array_equal(ng, no)
    k = float(no.mean())
    print(f"listed candidates per hull {k:.3f}", flush=True)
    assert k > 10.0
    m = np.minimum(no, 32)
    cols = np.arange(32)[None, :] < m[:, None]
    assert np.array_equal(np.where(cols, lg, 0), np.where(cols, lo, 0))
    dev.step(2)
    ora.run(2, deck["rad_per_step"], deck["lifter_about"])
    sg, so = by_pid(dev.get_state()), by_pid(ora.state())
    assert np.array_equal(sg["pid"], so["pid"])
    so["pos"] = oracle_positions_out(deck, so)
    for key in ("pos", "vel", "avel", "quat"):
        scale = max(float(np.max(np.abs(so[key]))), 1e-30)
        err = rel_err(sg[key], so[key], 1e-3 * scale)
        exact = float(np.mean(sg[key].astype(np.float64) == so[key]))
        print(f"hull drum 200k 2 steps {key}: max rel err {err:.3e}, bit-exact fraction {exact:.6f}", flush=True)
        assert err <= TOL, key
    dev.close()
    ora.close()
