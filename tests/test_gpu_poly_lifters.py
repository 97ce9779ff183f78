"""Polyhedra against turning lifters on the device (VolumeObject_InteractionParticle
for particle_type 1, Collision_Detection_Polyhedra_VolumeObject.cuh), against
the oracle.

Cubes sit against every face and top edge of the 8 lifters of a drum, so a
vertex, an edge or a face of a cube is inside a lifter from the first step;
the lifters turn before every step on the device (bdem_set_lifter_rotation)
and on the host in the oracle.  Tolerance 1e-5 as for the other polyhedra
tests: the detection and the force law perform the same IEEE operations in
the same order, and differences come from sin/cos in the quaternion update.
"""
import numpy as np
import pytest

import bdem
from bdem import scenarios as S
from helpers import make_pair, rel_err, by_pid, oracle_positions_out

pytestmark = pytest.mark.gpu

TOL = 1e-5


@pytest.mark.parametrize("steps", [1, 40])
def test_poly_mill_steps_match_oracle(steps):
    print(f"[case] poly mill 300 cubes, {steps} steps", flush=True)
    deck = S.poly_mill_deck(300)
    dev, ora = make_pair(deck)
    dev.drum_rotation(True)
    ora.set_drum_rotation(True)
    dev.set_lifter_rotation(deck["rad_per_step"], deck["lifter_about"])
    dev.step(steps)
    ora.run(steps, deck["rad_per_step"], deck["lifter_about"])
    assert ora.lifter_contacts() > 0
    sg, so = by_pid(dev.get_state()), by_pid(ora.state())
    assert np.array_equal(sg["pid"], so["pid"])
    so["pos"] = oracle_positions_out(deck, so)
    for key in ("pos", "vel", "avel", "quat"):
        scale = max(float(np.max(np.abs(so[key]))), 1e-30)
        err = rel_err(sg[key], so[key], 1e-3 * scale)
        print(f"poly mill {steps} steps {key}: max rel err {err:.3e}")
        assert err <= TOL, f"{key}: {err}"
    assert not dev.status() & bdem.lib.WARN_SORT
    dev.close()
    ora.close()
