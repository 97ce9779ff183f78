"""Dissipation tallies on the device against the oracle: the history laws
(this library's definition) and polyhedra (the reference's, through its
pair, wall and lifter responses).

Each contact's increment is added with a float atomic on the device and in
particle order by the oracle, so the totals agree to rounding (relative
1e-4), not bit for bit.  The states must not change when the tallies are on:
they stay within the usual 1e-5 of the oracle (spheres: bit-exact).
"""
import numpy as np
import pytest

import bdem
from bdem import scenarios as S
from helpers import make_pair, rel_err, by_pid, oracle_positions_out

pytestmark = pytest.mark.gpu


def _states_close(deck, dev, ora, tol):
    sg, so = by_pid(dev.get_state()), by_pid(ora.state())
    assert np.array_equal(sg["pid"], so["pid"])
    so["pos"] = oracle_positions_out(deck, so)
    for key in ("pos", "vel", "avel"):
        scale = max(float(np.max(np.abs(so[key]))), 1e-30)
        err = rel_err(sg[key], so[key], 1e-3 * scale)
        assert err <= tol, (key, err)


@pytest.mark.parametrize("law", ["lsd", "hm"])
def test_history_law_tallies_match_oracle(law):
    print(f"[case] tallies {law} random pack 500 spheres, 100 steps", flush=True)
    # sized so that float rounding of the 1e5 summed increments stays near 2e-5
    deck = S.random_deck(500, seed=17, L=9.0, law=law, vmax=30.0)
    deck["sim"].EnergyCalc = True
    dev, ora = make_pair(deck)
    dev.step(100)
    ora.step(100)
    _states_close(deck, dev, ora, 1e-5)
    eg = dev.energy().astype(np.float64)
    eo, _ = ora.energy()
    print(f"{law} tallies (J): device {eg}, oracle {eo}")
    assert eo[0] > 0 and eo[2] > 0
    assert np.allclose(eg, eo, rtol=1e-4, atol=0)
    dev.close()
    ora.close()


def test_polyhedra_mill_tallies_match_oracle():
    print("[case] tallies poly mill 300 cubes, 40 steps", flush=True)
    deck = S.poly_mill_deck(300)
    deck["sim"].EnergyCalc = True
    dev, ora = make_pair(deck)
    dev.drum_rotation(True)
    ora.set_drum_rotation(True)
    dev.set_lifter_rotation(deck["rad_per_step"], deck["lifter_about"])
    dev.step(40)
    ora.run(40, deck["rad_per_step"], deck["lifter_about"])
    _states_close(deck, dev, ora, 1e-5)
    eg = dev.energy().astype(np.float64)
    eo, _ = ora.energy()
    print(f"poly mill tallies (J): device {eg}, oracle {eo}")
    assert np.all(eo > 0)
    assert np.allclose(eg, eo, rtol=1e-4, atol=0)
    dev.close()
    ora.close()


def test_per_particle_hull_tallies_match_oracle():
    print("[case] tallies hull box 600 hulls, 20 steps", flush=True)
    deck = S.hull_box_deck(600, seed=41)
    deck["sim"].EnergyCalc = True
    dev, ora = make_pair(deck)
    dev.step(20)
    ora.step(20)
    _states_close(deck, dev, ora, 1e-5)
    eg = dev.energy().astype(np.float64)
    eo, _ = ora.energy()
    print(f"hull box tallies (J): device {eg}, oracle {eo}")
    assert eo[0] > 0
    assert np.allclose(eg, eo, rtol=1e-4, atol=0)
    dev.close()
    ora.close()
