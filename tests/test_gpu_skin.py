"""Skin masks: the step kernel tests only the stencil candidates its masks
mark (rows whose cells changed, and every row at an epoch start, are tested
in full).  Its contacts must be the reference stencil's contacts in the
reference order at every step -- through cell changes, re-masked rows and
new epochs.

Each step the library reports the contact list its step kernel used
(bdem_debug_step_contacts); the oracle's stencil traversal gives the list of
the same state (both sides stay bit-identical for the sphere laws, so the
states never drift apart).  The masked step is also checked against the
library's unmasked step (skin 0) over a longer run."""
import numpy as np
import pytest

import bdem
from bdem import scenarios as S
from helpers import make_pair, make_device, by_pid

pytestmark = pytest.mark.gpu

SKIN = 0.05   # cm: cell / 16 for these packs

DECKS = {
    # fast random packs: many cell changes per step, epochs every few steps
    "fast_ref": lambda: S.random_deck(3000, seed=61, L=12.0, vmax=400.0),
    "fast_hm_poly": lambda: S.random_deck(4000, seed=62, L=12.0, law="hm", polydisperse=True, vmax=300.0),
    "slow_lsd": lambda: S.random_deck(4000, seed=63, L=14.0, law="lsd", polydisperse=True, vmax=20.0),
}


@pytest.mark.parametrize("name", sorted(DECKS))
def test_step_contacts_equal_stencil_every_step(name):
    deck = DECKS[name]()
    print(f"[case] skin masks {name}", flush=True)
    dev, ora = make_pair(deck)
    assert dev.skin() == 0.0   # off by default
    dev.set_skin(SKIN)
    assert dev.skin() == pytest.approx(SKIN)
    for step in range(40):
        lo, no = ora.contacts()
        lg, ng = dev.debug_step_contacts()
        ora.step(1)
        assert np.array_equal(ng, no), f"step {step}: counts differ at {np.nonzero(ng != no)[0][:5]}"
        for i in np.nonzero(no)[0]:
            k = min(int(no[i]), 32)
            assert np.array_equal(lg[i, :k], lo[i, :k]), f"step {step}, particle {i}"
    sg, so = by_pid(dev.get_state()), by_pid(ora.state())
    assert np.array_equal(sg["pid"], so["pid"])
    assert np.max(np.abs(sg["pos"] - so["pos"])) <= 1e-5 * max(1.0, np.max(np.abs(so["pos"])))
    assert not dev.status() & bdem.lib.WARN_SORT
    dev.close()
    ora.close()


def test_masks_on_and_off_give_identical_states():
    deck = S.random_deck(6000, seed=64, L=14.0, law="hm", polydisperse=True, vmax=200.0)
    a, b = make_device(deck), make_device(deck)
    a.set_skin(SKIN)
    assert a.skin() > 0.0 and b.skin() == 0.0
    a.step(150)
    b.step(150)
    sa, sb = by_pid(a.get_state()), by_pid(b.get_state())
    for key in ("pid", "pos", "vel", "avel"):
        assert np.array_equal(sa[key], sb[key]), key
    a.close()
    b.close()


def test_aliased_grid_falls_back_to_the_stencil():
    deck = S.random_deck(1500, seed=15, L=10.0, aliased=True)
    dev = make_device(deck)
    dev.set_skin(SKIN)
    assert dev.skin() == 0.0
    dev.close()
