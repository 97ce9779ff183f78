"""Batched independent simulations (include/bdem_batch.h) on the GPU.

Each simulation of a batch must evolve exactly as it does alone in a plain
context with the same SimulationInfo: the same float operations in the same
order, so every position, velocity and spin is bit-identical after every
block of steps.  Checked for the plain-grid and the persistent (queued)
launch shapes, eager steps and HIP-graph replays, a cell factor above 1, and
one simulation against the FP32 oracle directly.
"""
import numpy as np
import pytest

import bdem
from bdem import scenarios as S
from helpers import by_pid, make_device, rel_err
from oracle_py import Oracle

pytestmark = pytest.mark.gpu


def _decks(count=6, seed=9, nmax=6000):
    return S.batch_box_decks(count=count, seed=seed, nmin=300, nmax=nmax)


def _with_sim(deck, sim):
    d = dict(deck)
    d["sim"] = sim
    return d


def _build(decks, **tune):
    batch = bdem.Batch()
    for d in decks:
        dev = make_device(d)
        batch.add(dev)
        dev.close()
    batch.set_tuning(**tune)
    batch.finalize()
    return batch


def _equal(a, b):
    a, b = by_pid(a), by_pid(b)
    assert np.array_equal(a["pid"], b["pid"])
    for k in ("pos", "vel", "avel"):
        assert np.array_equal(a[k], b[k]), k


@pytest.mark.parametrize("tune", [
    dict(threads=256, waves_per_cu=0, grab=1, cell_factor=1.0, use_graph=0),
    dict(threads=64, waves_per_cu=16, grab=4, cell_factor=1.0, use_graph=1),
    dict(threads=128, waves_per_cu=8, grab=1, cell_factor=1.25, use_graph=1),
], ids=["grid-eager", "persistent-graph", "persistent-cell1.25"])
def test_batch_equals_lone_runs(tune):
    decks = _decks()
    print(f"[case] batch of {len(decks)} boxes {[d['sim'].Num_Particles for d in decks]}, tuning {tune}", flush=True)
    batch = _build(decks, **tune)
    lone = [make_device(_with_sim(d, batch.instance_sim(k))) for k, d in enumerate(decks)]
    for block, nsteps in enumerate((1, 7, 40, 152)):
        batch.step(nsteps)
        for dev in lone:
            dev.step(nsteps)
        for k, dev in enumerate(lone):
            _equal(batch.get_state(k), dev.get_state())
        print(f"block {block}: {nsteps} steps, all {len(decks)} simulations identical", flush=True)
    assert batch.status() == 0
    for dev in lone:
        dev.close()
    batch.close()


def test_batch_restart_and_oracle():
    """finalize() restarts every simulation from its initial state; one
    simulation checked against the FP32 restatement directly (the
    restatement's P_IDs are the deck's own)."""
    decks = _decks(count=3, seed=11, nmax=2500)
    print("[case] batch restart + oracle, 3 boxes", flush=True)
    batch = _build(decks, threads=256, waves_per_cu=0, grab=1, cell_factor=1.0, use_graph=1)
    batch.step(30)
    first = [batch.get_state(k) for k in range(3)]
    batch.finalize()
    batch.step(30)
    for k in range(3):
        _equal(batch.get_state(k), first[k])
    ora = Oracle(_with_sim(decks[1], batch.instance_sim(1)), bits=32)
    ora.step(30)
    a, b = by_pid(batch.get_state(1)), by_pid(ora.state())
    for key in ("pos", "vel", "avel"):
        err = rel_err(a[key], b[key], 1e-3)
        print(f"{key}: GPU batch vs FP32 oracle max rel err {err:.2e}", flush=True)
        assert err <= 1e-5, key
    ora.close()
    batch.close()


def test_autotune_picks_a_shape_and_keeps_parity():
    decks = _decks(count=5, seed=13, nmax=4000)
    print("[case] autotune on 5 boxes", flush=True)
    batch = _build(decks)
    best, ms = batch.autotune(steps=5)
    print(f"autotuner chose {best} at {ms:.4f} ms/step", flush=True)
    assert ms > 0
    assert batch.tuning() == best
    lone = [make_device(_with_sim(d, batch.instance_sim(k))) for k, d in enumerate(decks)]
    batch.step(20)
    for k, dev in enumerate(lone):
        dev.step(20)
        _equal(batch.get_state(k), dev.get_state())
        dev.close()
    batch.close()
