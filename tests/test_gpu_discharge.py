"""Silo discharge on the GPU against the oracle (Remove_Silo_Discharged_Particles,
GPU_Memory_Transfer.cuh:650-664, every kill_check_interval steps as at
Device/DeviceInterface.cu:2466-2499).

A hopper full of spheres is released with the outlet open and crosses
several discharge checks.  Both sides must kill the same particles at the
same checks; on the step after each check (dead key C + 1, full re-sort)
the sorted keys and cell tables must be bit-exact, and the live states
must agree to 1e-5 (observed bit-exact).
"""
import numpy as np
import pytest

from bdem import scenarios as S
from helpers import make_pair, rel_err, by_pid

pytestmark = pytest.mark.gpu

TOL = 1e-5


@pytest.mark.parametrize("law", ["hm", "ref"])
def test_discharge_matches_oracle_across_checks(law):
    deck = S.silo_deck(3000, seed=5, law=law)
    sim = deck["sim"]
    sim.InitalDelta_t = 2e-5
    sim.Silo_Kill_Height = S.SILO_Y0 - 1.0        # just below the outlet
    deck["ext"].kill_check_interval = 40
    deck["vel"][:, 1] = -150.0                      # released downwards
    # the hatch is the last world object: open it before the first step
    deck["wobj"] = deck["wobj"][:-1]
    print(f"[case] discharge: silo n=3000 law={law}, check every 40 steps", flush=True)
    dev, ora = make_pair(deck)
    killed_prev = 0
    checks_with_kills = 0
    for block in range(24):
        dev.step(40)
        ora.step(40)
        sg, so = dev.get_state(), ora.state()
        dead_g = int(np.count_nonzero(sg["pid"] < 0))
        dead_o = int(np.count_nonzero(so["pid"] < 0))
        kc = dev.kill_count()
        assert kc == dead_o == dead_g, (block, kc, dead_o, dead_g)
        # the same particles: the live P_ID sets agree
        assert np.array_equal(np.sort(sg["pid"][sg["pid"] >= 0]), np.sort(so["pid"][so["pid"] >= 0]))
        if kc > killed_prev:
            checks_with_kills += 1
        killed_prev = kc
        # the step after a check: full re-sort with the dead at key C + 1
        kg, ig = dev.debug_sorted()
        ko, io = ora.sorted()
        assert np.array_equal(kg, ko), block
        assert np.array_equal(ig, io), block
        C = ora.num_cells()
        assert np.all(ko[-dead_o:] == C + 1) if dead_o else True
        cg, eg = dev.debug_cells()
        co, eo = ora.cells()
        assert np.array_equal(cg, co) and np.array_equal(eg, eo), block
        a, b = by_pid(sg), by_pid(so)
        live = b["pid"] >= 0
        for key in ("pos", "vel", "avel"):
            err = rel_err(a[key][live], b[key][live], 1e-3)
            assert err <= TOL, (block, key, err)
        print(f"after {40 * (block + 1)} steps: killed {kc}", flush=True)
    assert checks_with_kills >= 5, checks_with_kills
