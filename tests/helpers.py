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
    if deck.get("shapes") is not None:
        dev.set_poly_shapes(deck["shapes"])
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
    polyhedra, position - COM of the type (or of the particle's own shape) in
    float (Get_SystemState, GPU_Memory_Transfer.cuh:504)."""
    pos = np.asarray(state["pos"], np.float32)
    if deck["sim"].particle_type != 1:
        return pos.astype(np.float64)
    if deck.get("shapes") is not None:   # per-particle shapes: the COM of each P_ID's row
        com = np.array([[s.COM.x, s.COM.y, s.COM.z] for s in deck["shapes"]], np.float32)
        return (pos - com[np.maximum(state["pid"], 0)]).astype(np.float64)
    com = np.array([[p.COM.x, p.COM.y, p.COM.z] for p in deck["pobj"]], np.float32)
    return (pos - com[state["ptype"]]).astype(np.float64)


def surface_profile_np(pos, radius, lo, hi, bins, axis=0, live=None):
    """The free-surface profile of bdem_get_surface_profile restated in numpy
    float32: bin (int)((u - lo) * inv_w), inv_w = (float)bins / (hi - lo),
    highest y + r per bin, NaN where empty."""
    f = np.float32
    pos = np.asarray(pos, f)
    r = np.broadcast_to(np.asarray(radius, f), (len(pos),))
    inv_w = f(f(bins) / (f(hi) - f(lo)))
    t = (pos[:, axis] - f(lo)) * inv_w
    ok = (t >= 0) & (t < f(bins))
    if live is not None:
        ok &= live
    out = np.full(bins, -np.inf, f)
    np.maximum.at(out, t[ok].astype(np.int64), pos[ok, 1] + r[ok])
    out[np.isneginf(out)] = np.nan
    return out
