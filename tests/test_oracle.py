"""CPU checks that pin the restatement (oracle) independently of itself.

Each test compares the oracle with something that does not share its code:
golden fixtures from brute force and closed forms (tests/golden), numpy
re-implementations of the integer stages, textbook solutions (free fall,
constant spin, Hertz and linear springs, restitution), and FP32 vs FP64.
"""
import json
import math
import os

import numpy as np
import pytest

from bdem import scenarios as S
from bdem import types as T
from oracle_py import Oracle

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")


# ---------------------------------------------------------------- integer stages
def test_broadphase_matches_golden_fixture():
    g = json.load(open(os.path.join(GOLD, "broadphase_small.json")))
    d = g["deck"]
    deck = S.random_deck(d["n"], seed=d["seed"], L=d["L"])
    o = Oracle(deck)
    k, _ = o.hash()
    assert k.tolist() == g["keys"]
    ks, ix = o.sorted()
    assert ks.tolist() == g["sorted_keys"]
    assert ix.tolist() == g["sorted_idx"]
    cs, ce = o.cells()
    for key, (s, e) in g["cells"].items():
        assert (cs[int(key)], ce[int(key)]) == (s, e)
    assert np.count_nonzero(cs != 0xFFFFFFFF) == len(g["cells"])
    lst, nn = o.contacts()
    for i, li in enumerate(g["lists"]):
        assert nn[i] == len(li)
        assert lst[i, :min(len(li), 32)].tolist() == li[:32]


def _np_keys(deck, pos, pid):
    sim = deck["sim"]
    o = np.array([sim.worldOrigin.x, sim.worldOrigin.y, sim.worldOrigin.z], np.float32)
    c = np.array([sim.cellSize.x, sim.cellSize.y, sim.cellSize.z], np.float32)
    n = np.array([sim.num_NNCells.x, sim.num_NNCells.y, sim.num_NNCells.z])
    g = np.floor((pos.astype(np.float32) - o) / c).astype(np.int64)
    C = int(n.prod())
    h = (g[:, 0] + g[:, 2] * n[2] + g[:, 1] * n[0] * n[2]) & 0xFFFFFFFF
    inside = np.all((g >= 0) & (g < n), axis=1) & (h < C)
    keys = np.where(inside, h, C)
    keys = np.where(pid < 0, C + 1, keys)
    return keys.astype(np.uint32)


@pytest.mark.parametrize("aliased", [False, True])
def test_hash_sort_cells_vs_numpy(aliased):
    deck = S.random_deck(2000, seed=5, L=10.0, aliased=aliased)
    # a few dead and escaped particles
    deck["pid"][::97] = -1
    deck["pos"][5] = (500.0, 3.0, 3.0)
    deck["pos"][6] = (-7.0, 3.0, 3.0)
    o = Oracle(deck)
    k, _ = o.hash()
    ref = _np_keys(deck, deck["pos"], deck["pid"])
    assert np.array_equal(k, ref)
    ks, ix = o.sorted()
    order = np.argsort(ref, kind="stable")
    assert np.array_equal(ix, order)
    assert np.array_equal(ks, ref[order])
    cs, ce = o.cells()
    C = o.num_cells()
    sk = ref[order]
    for key in np.unique(sk[sk < C]):
        w = np.nonzero(sk == key)[0]
        assert cs[key] == w[0] and ce[key] == w[-1] + 1
    o.step(1)
    assert o.warnings() & 1  # escaped particles are reported


def test_contact_sets_vs_all_pairs():
    deck = S.random_deck(1500, seed=9, L=9.0, polydisperse=True)
    o = Oracle(deck)
    lst, nn = o.contacts()
    _, ix = o.sorted()
    p = deck["pos"][ix].astype(np.float32)
    r = deck["radius"][ix].astype(np.float32)
    d = p[:, None, :] - p[None, :, :]
    d2 = (d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]) + d[..., 2] * d[..., 2]
    hit = (np.sqrt(d2) - (r[:, None] + r[None, :])) < 0
    np.fill_diagonal(hit, False)
    for i in range(len(p)):
        want = set(np.nonzero(hit[i])[0].tolist())
        got = set(lst[i, :nn[i]].tolist())
        assert nn[i] == len(want) and got == want, i


# ---------------------------------------------------------------- pair law
def _two_sphere_deck(case, law="ref", offset=10.0):
    deck = S.random_deck(2, seed=1, L=20.0, law=law)
    deck["pos"] = np.array([case["pa"], case["pb"]], np.float32) + np.float32(offset)
    deck["vel"] = np.array([case["va"], case["vb"]], np.float32)
    deck["avel"] = np.array([case["wa"], case["wb"]], np.float32)
    return deck


def test_reference_pair_law_matches_golden():
    g = json.load(open(os.path.join(GOLD, "pair_forces.json")))
    for case in g["cases"]:
        deck = _two_sphere_deck(case, offset=0.0)
        for bits, tol in ((64, 1e-9), (32, 2e-5)):
            o = Oracle(deck, bits=bits)
            fpp, tpp, _, _, pid = o.forces()
            a = int(np.nonzero(pid == 0)[0][0])
            m = deck["pobj"][0].mass
            F = fpp[a] - np.array([0.0, -981.0, 0.0]) * np.float32(m)
            scale = max(np.abs(case["F"]).max(), 1e-12)
            assert np.abs(F - case["F"]).max() <= tol * scale
            tsc = max(np.abs(case["T"]).max(), 1e-12)
            assert np.abs(tpp[a] - case["T"]).max() <= tol * max(scale, tsc)


def _mstar_same(m):
    return m / 2.0


def test_lsd_and_hertz_static_contact_closed_forms():
    for law in ("lsd", "hm"):
        case = dict(pa=(0, 0, 0), pb=(0.875, 0, 0), va=(0, 0, 0), vb=(0, 0, 0), wa=(0, 0, 0), wb=(0, 0, 0))
        deck = _two_sphere_deck(case, law=law)
        deck["sim"].Force_Field = T.f3(0, 0, 0)
        o = Oracle(deck, bits=64)
        fpp, tpp, _, _, pid = o.forces()
        a = int(np.nonzero(pid == 0)[0][0])
        F = fpp[a]
        delta = 1.0 - 0.875
        if law == "lsd":
            want = 1e5 * delta
        else:
            E, nu = 1e5, 0.3
            estar = 1.0 / (2 * (1 - nu * nu) / E)
            want = 4.0 / 3.0 * estar * math.sqrt(0.25) * delta ** 1.5
        assert F[0] == pytest.approx(-want, rel=1e-7)  # a sits at -x of b
        assert abs(F[1]) < 1e-9 * want and abs(F[2]) < 1e-9 * want
        assert np.abs(tpp[a]).max() < 1e-9 * want


def test_coulomb_cap_in_sliding_contact():
    # overlapping pair sliding tangentially: after a few steps |Ft| = mu |Fn|
    case = dict(pa=(0, 0, 0), pb=(0.9375, 0, 0), va=(0, 500.0, 0), vb=(0, -500.0, 0), wa=(0, 0, 0), wb=(0, 0, 0))
    deck = _two_sphere_deck(case, law="lsd")
    deck["sim"].Force_Field = T.f3(0, 0, 0)
    deck["sim"].Rotation = False
    o = Oracle(deck, bits=64)
    o.step(10)
    fpp, _, _, _, pid = o.forces()
    a = int(np.nonzero(pid == 0)[0][0])
    st = o.state()
    ia, ib = np.argsort(st["pid"])[:2]
    n = st["pos"][ia] - st["pos"][ib]
    n /= np.linalg.norm(n)
    F = fpp[a]
    fn = abs(np.dot(F, n))
    ft = np.linalg.norm(F - np.dot(F, n) * n)
    assert ft == pytest.approx(0.4 * fn, rel=1e-6)


def _impact_cor(law, e, bits=64, steps=400):
    v0 = 20.0
    case = dict(pa=(0, 0, 0), pb=(1.002, 0, 0), va=(v0, 0, 0), vb=(-v0, 0, 0), wa=(0, 0, 0), wb=(0, 0, 0))
    deck = _two_sphere_deck(case, law=law)
    deck["sim"].Force_Field = T.f3(0, 0, 0)
    if law == "ref":
        deck["pobj"][0] = S.sphere_pobj(0.5, 2500.0, e, 1e5, 0.0)
    o = Oracle(deck, bits=bits)
    o.step(steps)
    st = o.state()
    va = st["vel"][np.argsort(st["pid"])[0]][0]
    return -va / v0


def test_head_on_restitution():
    # reference law: linear spring-dashpot tuned for e' = e - 0.03808
    # (Input/KSimulationData.cpp:893-896), measured over the discrete impact
    assert _impact_cor("ref", 0.6) == pytest.approx(0.6 - 0.03808, abs=0.02)
    assert _impact_cor("lsd", 0.6) == pytest.approx(0.6, abs=0.02)
    # Hertz with Tsuji damping restitutes close to e
    assert _impact_cor("hm", 0.6) == pytest.approx(0.6, abs=0.05)


# ---------------------------------------------------------------- integrator
def test_free_fall_closed_form():
    deck = S.random_deck(1, seed=2, L=200.0)
    deck["pos"] = np.array([[100.0, 150.0, 100.0]], np.float32)
    deck["vel"] = np.array([[3.0, 1.5, -2.0]], np.float32)
    deck["avel"] = np.zeros((1, 3), np.float32)
    o = Oracle(deck, bits=64)
    nsteps, dt, g = 200, float(np.float32(1e-5)), -981.0
    o.step(nsteps)
    st = o.state()
    v0 = deck["vel"][0].astype(np.float64)
    x0 = deck["pos"][0].astype(np.float64)
    vy = v0[1] + nsteps * g * dt
    y = x0[1] + nsteps * v0[1] * dt + g * dt * dt * nsteps * (nsteps + 1) / 2
    assert st["vel"][0, 1] == pytest.approx(vy, rel=1e-12)
    assert st["pos"][0, 1] == pytest.approx(y, rel=1e-12)
    assert st["pos"][0, 0] == pytest.approx(x0[0] + nsteps * v0[0] * dt, rel=1e-12)


def test_constant_spin_quaternion():
    deck = S.random_deck(1, seed=2, L=200.0)
    deck["pos"] = np.array([[100.0, 150.0, 100.0]], np.float32)
    deck["vel"] = np.zeros((1, 3), np.float32)
    w = np.array([3.0, -4.0, 12.0])
    deck["avel"] = np.array([w], np.float32)
    deck["sim"].sphere_orient = True
    o = Oracle(deck, bits=64)
    n, dt = 300, float(np.float32(1e-5))
    o.step(n)
    st = o.state()
    q = st["quat"][0]
    ang = np.linalg.norm(w) * dt * n
    axis = w / np.linalg.norm(w)
    want = np.concatenate([[math.cos(ang / 2)], math.sin(ang / 2) * axis])
    assert np.allclose(q, want, atol=1e-9)
    assert np.linalg.norm(q) == pytest.approx(1.0, abs=1e-12)
    assert np.allclose(st["avel"][0], w)  # no torque: no update, no rolling resistance


def test_sphere_resting_on_floor_overlap():
    deck = S.random_deck(1, seed=2, L=20.0)
    po = deck["pobj"][0]
    deck["pos"] = np.array([[10.0, po.radius - 0.001, 10.0]], np.float32)
    deck["vel"] = np.zeros((1, 3), np.float32)
    deck["avel"] = np.zeros((1, 3), np.float32)
    o = Oracle(deck, bits=64)
    o.step(6000)
    st = o.state()
    delta = po.radius - st["pos"][0, 1]
    assert delta == pytest.approx(po.mass * 981.0 / po.surface_Kn, rel=1e-3)


def test_fp32_tracks_fp64_over_short_horizon():
    deck = S.random_deck(800, seed=31, L=8.0, law="hm", polydisperse=True, vmax=5.0)
    a, b = Oracle(deck, bits=32), Oracle(deck, bits=64)
    a.step(20)
    b.step(20)
    sa, sb = a.state(), b.state()
    oa, ob = np.argsort(sa["pid"]), np.argsort(sb["pid"])
    err = np.abs(sa["pos"][oa] - sb["pos"][ob]).max()
    assert err < 1e-4  # cm, over 20 steps


def test_discharge_kill_marks_and_sentinel():
    deck = S.random_deck(200, seed=3, L=10.0)
    deck["sim"].Kill_Particles = True
    deck["sim"].Silo_Kill_Height = 3.0
    ext = T.ExtParams()
    ext.kill_check_interval = 5
    deck["ext"] = ext
    o = Oracle(deck)
    o.step(5)
    st = o.state()
    dead = st["pid"] < 0
    assert dead.any()
    ms = deck["sim"].max_size
    assert np.allclose(st["pos"][dead], [ms.x, ms.y, ms.z])
    k, _ = o.hash()
    assert np.all(k[dead] == o.num_cells() + 1)


# --------------------------------------------------------------------------
# a9 lifters (VolumeObject_InteractionParticle, DEM_Kernels_Compute.cuh:12-246)
# --------------------------------------------------------------------------
LIFT_X0, LIFT_X1, LIFT_H, LIFT_L = 9.0, 11.0, 2.0, 10.0


def _one_sphere_lifter(pos, vel=(0.0, 0.0, 0.0), bits=64):
    """One sphere, one horizontal lifter (top face y = LIFT_H, top edges along
    z at x = LIFT_X0 and LIFT_X1), no gravity, not a mill (every particle is
    checked), box walls far away."""
    deck = S.random_deck(1, seed=2, L=20.0)
    deck["sim"].Force_Field = T.f3(0.0, 0.0, 0.0)
    x0, x1, h, L = LIFT_X0, LIFT_X1, LIFT_H, LIFT_L
    deck["dobj"] = [S.lifter_dobj([(x0, 0, 0), (x0, 0, L), (x1, 0, L), (x1, 0, 0),
                                   (x0, h, 0), (x0, h, L), (x1, h, L), (x1, h, 0)],
                                  ((x0 + x1) / 2, h / 2, L / 2), 2.0, attached=False)]
    deck["pos"] = np.array([pos], np.float32)
    deck["vel"] = np.array([vel], np.float32)
    deck["avel"] = np.zeros((1, 3), np.float32)
    return deck, Oracle(deck, bits=bits)


def test_lifter_face_contact_closed_form():
    """Sphere at rest pressed onto the top face by delta: F = Dsurface_Kn * delta
    along the face normal, no damping (v = 0), no tangential force, no torque."""
    po = S.random_deck(1, seed=2, L=20.0)["pobj"][0]
    r, delta = po.radius, 0.05
    deck, o = _one_sphere_lifter((10.0, LIFT_H + r - delta, 5.0))
    o.step(1)
    st = o.state()
    dt = float(np.float32(1e-5))
    assert st["vel"][0, 1] == pytest.approx(po.Dsurface_Kn * delta / po.mass * dt, rel=1e-6)
    assert abs(st["vel"][0, 0]) < 1e-12 and abs(st["vel"][0, 2]) < 1e-12
    assert np.all(st["avel"][0] == 0.0)
    assert o.lifter_types()[0] == 1


def test_lifter_edge_contact_closed_form():
    """Sphere diagonally off the top edge at x = LIFT_X0: neither face holds the
    projected point, the edge does.  F = Dsurface_Kn * delta along (-1, 1, 0)/sqrt 2,
    and no spin (contact point = centre)."""
    po = S.random_deck(1, seed=2, L=20.0)["pobj"][0]
    r, delta = po.radius, 0.05
    a = (r - delta) / math.sqrt(2.0)
    deck, o = _one_sphere_lifter((LIFT_X0 - a, LIFT_H + a, 5.0))
    o.step(1)
    st = o.state()
    dt = float(np.float32(1e-5))
    dv = po.Dsurface_Kn * delta / po.mass * dt / math.sqrt(2.0)
    assert st["vel"][0, 0] == pytest.approx(-dv, rel=1e-5)
    assert st["vel"][0, 1] == pytest.approx(dv, rel=1e-5)
    assert np.all(st["avel"][0] == 0.0)
    assert o.lifter_types()[0] == 2


def test_lifter_edge_memory_skips_faces():
    """Above the top face next to the edge both tests hold.  With no remembered
    contact the face wins (normal +y, depth delta); after an edge contact the
    faces are skipped (Collision_Detection_Spheres.cuh:121) and the edge normal
    acts with depth r - |p - edge|."""
    po = S.random_deck(1, seed=2, L=20.0)["pobj"][0]
    r, delta, b = po.radius, 0.05, 0.1
    p = (LIFT_X0 + b, LIFT_H + r - delta, 5.0)
    dt = float(np.float32(1e-5))
    k = po.Dsurface_Kn / po.mass * dt
    deck, o = _one_sphere_lifter(p)
    o.step(1)
    v = o.state()["vel"][0]
    assert v[1] == pytest.approx(k * delta, rel=1e-6) and abs(v[0]) < 1e-12
    deck, o = _one_sphere_lifter(p)
    o.set_lifter_types([2])
    o.step(1)
    v = o.state()["vel"][0]
    d = math.hypot(b, r - delta)
    assert v[0] == pytest.approx(k * (r - d) * b / d, rel=1e-5)
    assert v[1] == pytest.approx(k * (r - d) * (r - delta) / d, rel=1e-5)
    assert o.lifter_types()[0] == 2


def test_lifter_attached_to_drum_moves_with_it():
    """An attached lifter moves with the drum: a sphere resting on it in a
    rotating mill feels tangential friction along the lifter velocity
    cross((0,0,-MillRotVel), cp - (R,R,0)), zero for a fixed lifter."""
    deck = S.mill_deck(n=40, seed=3)
    deck = S.place_at_lifters(deck)
    o_rot = Oracle(deck, bits=64)
    fixed = dict(deck)
    fixed["dobj"] = [S.T.DOBJ.from_buffer_copy(d) for d in deck["dobj"]]
    for d in fixed["dobj"]:
        d.is_attached_DWorld = False
    o_fix = Oracle(fixed, bits=64)
    o_rot.step(1)
    o_fix.step(1)
    a, b = o_rot.state(), o_fix.state()
    va = a["vel"][np.argsort(a["pid"])]
    vb = b["vel"][np.argsort(b["pid"])]
    # P_ID 2 sits on lifter 0's top face: tangential push along +x
    # (MillRotVel < 0, contact point below the drum centre)
    assert va[2, 0] - vb[2, 0] > 0.0
