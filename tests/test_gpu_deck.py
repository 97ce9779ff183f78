"""The reference's project decks, run through the library against the oracle.

Each case loads a shipped deck with the deck reader (bdem_deck_load), sets
a context up from it (bdem_deck_setup) and reads its start file into the
context (bdem_deck_load_state), exactly as the reference's main() does with
KSimulationData, the KDevice constructor and Read_FromFile -- or, for the
decks without a start file, lays the particles on the deck's grid as
Device_Set_SimulationData does; the oracle gets the same deck and state.
Mill decks turn their lifters before every step, as UpdateSim does
(main.cpp:114-121).
"""
import os

import numpy as np
import pytest

from bdem import scenarios as S
from bdem.deck import Deck
from helpers import rel_err, by_pid, oracle_positions_out
from oracle_py import Oracle

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DECKS = os.path.join(ROOT, "tests", "golden", "decks")
TOL = 1e-5

CASES = [
    ("1.Particles_Box", "Spheres_128cm", 100),   # grid fill (Set_Position_DefaultGrid), the run selector's deck
    ("1.Particles_Box", "Poly_128cm", 20),       # grid fill, 8192 dodecahedra
    ("3.BallMill", "Mill_SPH30", 300),
    ("3.BallMill", "Mill_SPH20", 300),
    ("2.Silo", "SPH_30", 300),
    ("4.UnitTests", "2P", 400),
    ("4.UnitTests", "2PS", 400),
    ("4.UnitTests", "Symm", 400),
]


def _load(project, world):
    d = Deck(os.path.join(DECKS, project), world)
    dev = d.device()
    hdr = d.load_state(dev)
    return d, dev, hdr


def test_loaded_state_is_the_start_file():
    d, dev, hdr = _load("3.BallMill", "Mill_SPH30")
    print("[case] load Mill_SPH30", flush=True)
    st = d.read_state()
    g = dev.get_state()
    assert hdr.spec_type == 3
    order = np.argsort(g["pid"])
    ref = np.argsort(st["pid"])
    assert np.array_equal(g["pid"][order], st["pid"][ref])
    for key in ("pos", "vel", "avel"):
        assert np.array_equal(g[key][order], st[key][ref]), key
    assert np.array_equal(dev.energy(), np.zeros(3, np.float32))
    dev.close()


@pytest.mark.parametrize("project,world,steps", CASES)
def test_deck_matches_oracle(project, world, steps):
    print(f"[case] deck {world} {steps} steps", flush=True)
    d, dev, hdr = _load(project, world)
    sc = d.scenario()
    ora = Oracle(sc)
    mill = d.info.sim_type == 1
    if mill:
        dev.drum_rotation(True)
        ora.set_drum_rotation(True)
    dobj = list(d.dobj)
    for s in range(steps):
        if mill and dobj:
            dobj = [S.rotate_dobj(x, sc["lifter_about"], sc["rad_per_step"]) for x in dobj]
            dev.update_dobj(dobj)
            ora.set_dobj(dobj)
        dev.step(1)
        ora.step(1)
    sg, so = by_pid(dev.get_state()), by_pid(ora.state())
    so["pos"] = oracle_positions_out(sc, so)
    assert np.array_equal(sg["pid"], so["pid"])
    for key in ("pos", "vel", "avel"):
        err = rel_err(sg[key], so[key], 1e-3)
        print(f"{world} {key}: max rel err {err:.3e}")
        assert err <= TOL, f"{key}: {err}"
    assert dev.status() & 1 == 0
    dev.close()
    ora.close()
