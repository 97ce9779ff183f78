"""Shared helpers for the parity tests (test infrastructure)."""
import numpy as np

import bdem
from oracle_py import Oracle


def make_device(deck):
    dev = bdem.Device(deck["wobj"], deck["pobj"], deck["sim"], dobj=deck.get("dobj"))
    if deck.get("ext") is not None:
        dev.set_ext(deck["ext"])
    if deck.get("radius") is not None:
        dev.set_radii(deck["radius"], deck["mass"])
    dev.set_state(deck["pos"], deck["ptype"], deck["pid"], vel=deck["vel"], avel=deck["avel"],
                  quat=deck.get("quat"))
    return dev


def make_pair(deck, bits=32):
    return make_device(deck), Oracle(deck, bits=bits)


def rel_err(a, b, floor):
    """max |a-b| / max(|b|, floor) over all entries."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), floor))) if a.size else 0.0


def by_pid(state):
    """Reorder a state dict by P_ID (dead particles keep their slot order)."""
    order = np.argsort(state["pid"], kind="stable")
    return {k: v[order] for k, v in state.items()}


def oracle_positions_out(deck, state):
    """The positions the library returns for the oracle's state: for
    polyhedra, position - COM of the type in float (Get_SystemState,
    GPU_Memory_Transfer.cuh:504)."""
    pos = np.asarray(state["pos"], np.float32)
    if deck["sim"].particle_type != 1:
        return pos.astype(np.float64)
    com = np.array([[p.COM.x, p.COM.y, p.COM.z] for p in deck["pobj"]], np.float32)
    return (pos - com[state["ptype"]]).astype(np.float64)
