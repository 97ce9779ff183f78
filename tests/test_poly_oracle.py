"""CPU checks that pin the polyhedra part of the restatement.

Closed forms: two cubes in face-face, edge-face, vertex-face and edge-edge
contact at a known depth give that depth, the face normal and the contact
point by hand, and the linear spring force; swapping the two reverses the
force exactly.  An independent separation test (the convex hull of the
Minkowski difference, scipy) decides whether random pairs intersect: no
separated pair may report a contact.  The reference's detector (vertex-in-
polyhedron, then edge-edge with a validity check it marks "not robust",
Collision_Detection_Polyhedra.cuh:640-660) misses some edge-through-face
crossings; the tests bound that rate rather than assert zero.  The inverse
inertia product is checked against the box tensor by hand, free spin against
the closed-form quaternion, and resting contact on the floor against the
vertex-plane depths.
"""
import math

import numpy as np
import pytest

from bdem import scenarios as S
from oracle_py import Oracle

I = np.array([1, 0, 0, 0], np.float32)
S2 = math.sqrt(2.0) / 2.0
S3 = math.sqrt(3.0) / 2.0
BIG = (S.CUBE_VERTS * 2.0, S.CUBE_FACES)
D = 0.01
KN = 1e5


def _q_corner_to_x():
    """Rotation taking the cube corner (1,1,1)/sqrt(3) onto +x."""
    u = np.ones(3) / math.sqrt(3.0)
    axis = np.cross(u, [1.0, 0.0, 0.0])
    axis /= np.linalg.norm(axis)
    h = 0.5 * math.acos(u[0])
    return np.array([math.cos(h), *(math.sin(h) * axis)], np.float32)


# name: (A orientation, B centre x offset, B orientation, B shape, class, contact point x)
CASES = {
    "face-face": (I, 0.5 + 1.0 - D, I, BIG, 3, 0.5),
    "edge-face": (S.axis_quat((0, 0, 1), 45), S2 + 1.0 - D, I, BIG, 2, S2),
    "vertex-face": (_q_corner_to_x(), S3 + 1.0 - D, I, BIG, 1, S3),
    "edge-edge": (S.axis_quat((0, 0, 1), 45), 2 * S2 - D, S.axis_quat((0, 1, 0), 45), None, 4, S2),
}


def _pair(name):
    qa, dx, qb, shb, _, _ = CASES[name]
    return S.poly_pair_deck((5, 5, 5), (5 + dx, 5, 5), qa, qb, shape_b=shb, gravity=False)


@pytest.mark.parametrize("name", sorted(CASES))
@pytest.mark.parametrize("bits", [64, 32])
def test_cube_contact_closed_forms(name, bits):
    _, _, _, _, cls, cpx = CASES[name]
    o = Oracle(_pair(name), bits=bits)
    r = o.poly_pair(0, 1)
    tol = 1e-9 if bits == 64 else 2e-6
    assert r["hit"] and r["cls"] == cls
    want = D + (1e-5 if cls == 4 else 0.0)   # edge-edge adds 1e-5 (Collision_Detection_Polyhedra.cuh:672)
    assert abs(r["dis"] - want) < 2e-6 + tol
    assert np.allclose(r["n"], [-1, 0, 0], atol=1e-6)
    assert np.allclose(r["cp"], [5 + cpx, 5, 5], atol=2e-6)
    # at rest: linear spring only, along the normal (Collision_Response.cuh:374-494)
    assert np.allclose(r["F"], -KN * r["dis"] * np.array([1, 0, 0]), rtol=1e-6, atol=1e-6)
    s = o.poly_pair(1, 0)
    assert s["hit"] and s["cls"] == cls and s["dis"] == r["dis"]
    assert np.array_equal(s["F"], -r["F"])
    o.close()


def test_face_face_without_moment():
    """Face-face contact acts at the centroid of the penetrating vertices:
    aligned cubes feel no torque."""
    o = Oracle(_pair("face-face"), bits=64)
    assert np.all(o.poly_pair(0, 1)["T"] == 0.0)


@pytest.mark.parametrize("depth,speed,expect", [
    (0.2, 0.0, 0.025),     # > 0.15 r: contact error, reset to 0.025 r
    (0.06, 10.0, 0.003),   # 0.05 r .. 0.15 r, slow: 0.003 r
    (0.06, 80.0, 0.010),   # fast: 0.010 r (the first speed branch wins)
])
def test_depth_clamps(depth, speed, expect):
    """DEM_Kernels_Compute.cuh:918-950: large depths are replaced by fixed
    fractions of A's bounding radius, along the line of centres."""
    r0 = math.sqrt(3.0) / 2.0
    deck = S.poly_pair_deck((5, 5, 5), (5 + 1.5 - depth * r0, 5, 5), I, I, shape_b=BIG,
                            va=(speed, 0, 0), gravity=False)
    o = Oracle(deck, bits=64)
    r = o.poly_pair(0, 1)
    R = deck["pobj"][0].radius
    assert r["hit"]
    assert r["dis"] == pytest.approx(np.float32(expect) * R, rel=1e-6)
    assert np.allclose(r["n"], [-1, 0, 0])
    assert np.allclose(r["cp"], deck["pos"][0])


def _rot(q, v):
    w, x, y, z = (float(c) for c in q)
    R = np.array([[w * w + x * x - y * y - z * z, 2 * (x * y - w * z), 2 * (x * z + w * y)],
                  [2 * (x * y + w * z), w * w - x * x + y * y - z * z, 2 * (y * z - w * x)],
                  [2 * (x * z - w * y), 2 * (y * z + w * x), w * w - x * x - y * y + z * z]])
    return v @ R.T


def _gap(VA, VB):
    """Signed separation of two convex vertex sets: > 0 separated (a lower
    bound of the distance), < 0 intersecting, from the Minkowski difference."""
    from scipy.spatial import ConvexHull
    M = (VA[:, None, :] - VB[None, :, :]).reshape(-1, 3)
    return float(np.max(ConvexHull(M).equations[:, 3]))


@pytest.mark.parametrize("kind", ["cube", "hull"])
def test_random_pairs_against_minkowski_separation(kind):
    shape = (S.CUBE_VERTS, S.CUBE_FACES) if kind == "cube" else S.hull_shape(31, size=1.7)
    rng = np.random.default_rng(5)
    q = S.random_quats(91, 600)
    counts = dict(sep=0, sep_hit=0, pen=0, pen_hit=0)
    for t in range(300):
        dirn = rng.normal(size=3)
        dirn /= np.linalg.norm(dirn)
        pa = np.array([5.0, 5.0, 5.0])
        deck = S.poly_pair_deck(pa, pa, q[2 * t], q[2 * t + 1], shape_a=shape, gravity=False)
        R = deck["pobj"][0].radius
        deck["pos"][1] = (pa + dirn * rng.uniform(0.5, 1.0) * 2 * R).astype(np.float32)
        P = deck["pobj"][0]
        loc = np.array([[P.Local_Vertex[k].x, P.Local_Vertex[k].y, P.Local_Vertex[k].z]
                        for k in range(P.num_vertex)])
        VA = _rot(q[2 * t], loc) + deck["pos"][0]
        VB = _rot(q[2 * t + 1], loc) + deck["pos"][1]
        g = _gap(VA, VB)
        o = Oracle(deck, bits=64)
        hit = o.poly_pair(0, 1)["hit"]
        o.close()
        if g > 1e-3:
            counts["sep"] += 1
            counts["sep_hit"] += hit
        elif g < -1e-3:
            counts["pen"] += 1
            counts["pen_hit"] += hit
    print(kind, counts)
    assert counts["sep"] > 50 and counts["pen"] > 50
    assert counts["sep_hit"] == 0
    assert counts["pen_hit"] >= 0.93 * counts["pen"]


def test_angular_acc_tensor_product():
    """Get_Angular_Acc forms IC = RQ I0 and then IN = RQ^T IC
    (Device/Utilities.cuh:511-560): RQ^T RQ = 1, so the stored tensor acts
    unrotated whatever the orientation.  A 1 x 2 x 3 box turned 90 degrees
    about z therefore gets alpha_x = Ixx^-1 tau_x (not Iyy^-1), and a random
    orientation gives I0^-1 tau to rounding."""
    V = S.CUBE_VERTS * np.array([1.0, 2.0, 3.0])
    deck = S.poly_pair_deck((3, 5, 5), (8, 5, 5), S.axis_quat((0, 0, 1), 90), I,
                            shape_a=(V, S.CUBE_FACES), gravity=False)
    o = Oracle(deck, bits=64)
    Iinv = np.array(deck["pobj"][0].InertiaT[:]).reshape(3, 3)
    m = deck["pobj"][0].mass
    assert Iinv[0, 0] == pytest.approx(12.0 / (13.0 * m), rel=1e-6)   # box: Ixx = m (b^2 + c^2) / 12
    assert Iinv[1, 1] == pytest.approx(12.0 / (10.0 * m), rel=1e-6)
    a = o.angular_acc(0, [2.0, 0.0, 0.0])
    assert a == pytest.approx([2.0 * Iinv[0, 0], 0.0, 0.0], rel=1e-6, abs=1e-6 * Iinv[0, 0])
    o.close()
    for q in S.random_quats(4, 4):
        deck["quat"][0] = q
        for bits, tol in ((64, 3e-7), (32, 2e-6)):   # q is stored in float: |q| = 1 to 1e-7
            o = Oracle(deck, bits=bits)
            tau = np.array([0.3, -1.2, 0.7])
            assert o.angular_acc(0, tau) == pytest.approx(Iinv @ tau, rel=tol)
            o.close()


def test_free_spin_quaternion():
    """No contact, no torque: omega stays, q_n = dq^n q_0."""
    w = np.array([3.0, -4.0, 12.0])
    deck = S.poly_pair_deck((3, 5, 5), (8, 5, 5), I, I, wa=w, gravity=False)
    o = Oracle(deck, bits=64)
    n, dt = 300, float(np.float32(1e-5))
    o.step(n)
    st = o.state()
    k = int(np.nonzero(st["pid"] == 0)[0][0])
    ang = np.linalg.norm(w) * dt * n
    want = np.concatenate([[math.cos(ang / 2)], math.sin(ang / 2) * w / np.linalg.norm(w)])
    assert np.allclose(st["quat"][k], want, atol=1e-9)
    assert np.allclose(st["avel"][k], w)


def test_cube_on_floor_faces_and_vertices():
    """Collision_Detection_Vertcies_Surface: a flat cube sunk by d into the
    floor is a face contact of depth d at its centre (no moment); a cube on
    one corner is a vertex contact at the projected corner."""
    d = 0.004
    deck = S.poly_pair_deck((5, 0.5 - d, 5), (9, 5, 9), I, I, gravity=False)
    o = Oracle(deck, bits=64)
    F, T = o.poly_walls(0)
    kn_s = deck["pobj"][0].surface_Kn
    assert F == pytest.approx([0.0, kn_s * d, 0.0], rel=1e-5, abs=1e-6)
    assert np.all(T == 0.0)
    o.close()
    q = _q_corner_to_x()                      # corner along +x ...
    q2 = S.axis_quat((0, 0, 1), -90)          # ... turned to -y
    w_, x_, y_, z_ = q2
    a = q
    qq = np.array([w_ * a[0] - x_ * a[1] - y_ * a[2] - z_ * a[3],
                   w_ * a[1] + x_ * a[0] + y_ * a[3] - z_ * a[2],
                   w_ * a[2] - x_ * a[3] + y_ * a[0] + z_ * a[1],
                   w_ * a[3] + x_ * a[2] - y_ * a[1] + z_ * a[0]], np.float32)
    deck = S.poly_pair_deck((5, S3 - d, 5), (9, 5, 9), qq, I, gravity=False)
    o = Oracle(deck, bits=64)
    F, T = o.poly_walls(0)
    assert F[1] == pytest.approx(kn_s * d, rel=1e-4)
    assert abs(F[0]) < 1e-3 and abs(F[2]) < 1e-3
    assert np.linalg.norm(T) < 1e-3 * np.linalg.norm(F)   # the corner is under the centre


def test_poly_pack_fp32_tracks_fp64():
    deck = S.poly_deck(216, seed=33, L=9.0, shape="mixed", vmax=20.0)
    a, b = Oracle(deck, bits=32), Oracle(deck, bits=64)
    a.step(20)
    b.step(20)
    sa, sb = a.state(), b.state()
    oa, ob = np.argsort(sa["pid"]), np.argsort(sb["pid"])
    assert np.array_equal(sa["pid"][oa], sb["pid"][ob])
    assert np.all(np.isfinite(sa["pos"]))
    err = np.abs(sa["pos"][oa] - sb["pos"][ob]).max()
    assert err < 1e-3, err


# ---- polyhedra against a lifter (VolumeObject_InteractionParticle, particle_type 1) ----

def _qmul(a, b):
    w1, x1, y1, z1 = (float(c) for c in a)
    w2, x2, y2, z2 = (float(c) for c in b)
    return np.array([w1 * w2 - x1 * x2 - y1 * y2 - z1 * z2, w1 * x2 + x1 * w2 + y1 * z2 - z1 * y2,
                     w1 * y2 - x1 * z2 + y1 * w2 + z1 * x2, w1 * z2 + x1 * y2 - y1 * x2 + z1 * w2], np.float32)


def _lifter_case(centre, q):
    """A unit cube at rest (no gravity) near one lifter: x 3..7, y 0..1.5,
    z 1..9 (top face y = 1.5, side faces x = 3 and x = 7); the second cube of
    the pair deck is far away."""
    deck = S.poly_pair_deck(centre, (9, 8, 9), q, I, gravity=False)
    V = [(3, 0, 1), (3, 0, 9), (7, 0, 9), (7, 0, 1), (3, 1.5, 1), (3, 1.5, 9), (7, 1.5, 9), (7, 1.5, 1)]
    deck["dobj"] = [S.lifter_dobj(V, (5, 0.75, 5), 4.0, attached=False)]
    return deck, deck["dobj"][0]


CORNER_DOWN = _qmul(S.axis_quat((0, 0, 1), -90), _q_corner_to_x())   # corner (1,1,1)/sqrt(3) to -y


@pytest.mark.parametrize("name,centre,q,face", [
    ("face-face", (5, 1.5 + 0.5 - D, 5), I, 2),
    ("edge-face", (5, 1.5 + S2 - D, 5), S.axis_quat((0, 0, 1), 45), 2),
    ("vertex-face", (5, 1.5 + S3 - D, 5), CORNER_DOWN, 2),
    ("side vertex-face", (3 - S3 + D, 1.0, 5), _q_corner_to_x(), 0),
])
@pytest.mark.parametrize("bits", [64, 32])
def test_cube_against_lifter_closed_forms(name, centre, q, face, bits):
    deck, L = _lifter_case(centre, q)
    n = np.array([L.faces[face].normal.x, L.faces[face].normal.y, L.faces[face].normal.z])
    o = Oracle(deck, bits=bits)
    fl, tl = o.lifter_forces()
    kn = deck["pobj"][0].Dsurface_Kn
    # at rest: the lifter spring alone, Dsurface_Kn * depth along the face normal, no moment
    assert fl[0] == pytest.approx(kn * D * n, rel=2e-4, abs=1e-3)
    assert np.linalg.norm(tl[0]) < 1e-4 * np.linalg.norm(fl[0])
    assert np.all(fl[1] == 0.0)
    o.close()


def test_cube_clear_of_the_lifter_feels_nothing():
    deck, _ = _lifter_case((5, 1.5 + 0.5 + 0.02, 5), I)
    o = Oracle(deck, bits=32)
    fl, tl = o.lifter_forces()
    assert np.all(fl == 0.0) and np.all(tl == 0.0)
    o.close()
