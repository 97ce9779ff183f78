"""GPU parity for per-particle polyhedron shapes (bdem_set_poly_shapes):
every particle its own convex hull of 8 to 16 points, against the oracle
with the same table.

Contact lists are bit-exact; forces and states within relative 1e-5, as for
the per-type polyhedra (the detection, the pair law, the walls, the lifters
and the tensor integrator perform the oracle's IEEE operations in its order;
differences come from sin/cos in the quaternion update).  A two-type deck
run through a per-P_ID table that holds each particle's type geometry must
equal the per-type run bit for bit.
"""
import numpy as np
import pytest

import bdem
from bdem import scenarios as S
from bdem import types as T
from helpers import make_device, make_pair, rel_err, by_pid, oracle_positions_out

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _compare_states(name, deck, dev, ora):
    sg, so = by_pid(dev.get_state()), by_pid(ora.state())
    assert np.array_equal(sg["pid"], so["pid"])
    so["pos"] = oracle_positions_out(deck, so)
    for key in ("pos", "vel", "avel", "quat"):
        scale = max(float(np.max(np.abs(so[key]))), 1e-30)
        err = rel_err(sg[key], so[key], 1e-3 * scale)
        print(f"{name} {key}: max rel err {err:.3e}")
        assert err <= TOL, f"{name} {key}: {err}"


@pytest.fixture(scope="module")
def hull_box():
    print("[case] hull box 600 per-particle hulls", flush=True)
    return S.hull_box_deck(600, seed=41)


def test_hull_contact_lists_and_forces(hull_box):
    dev, ora = make_pair(hull_box)
    assert dev.num_poly_shapes() == (600, True)
    lg, ng = dev.debug_contacts()
    lo, no = ora.contacts()
    assert np.array_equal(ng, no)
    assert no.sum() > 0
    for i in np.nonzero(no)[0]:
        k = min(int(no[i]), 32)
        assert np.array_equal(lg[i, :k], lo[i, :k]), f"particle {i}"
    g = dev.debug_forces()
    o = ora.forces()
    assert np.array_equal(g[4], o[4])
    assert float(np.abs(o[0]).max()) > 0.0
    for a, b, what in zip(g[:4], o[:4], ("F_pp", "T_pp", "F_wall", "T_wall")):
        scale = max(float(np.max(np.abs(b))), 1e-30)
        err = rel_err(a, b, 1e-3 * scale)
        print(f"hull box {what}: max rel err {err:.3e}")
        assert err <= TOL, f"{what}: {err}"
    dev.close()
    ora.close()


@pytest.mark.parametrize("kmin,kmax", [(4, 6), (17, 18)])
def test_smallest_and_largest_hulls_match_oracle(kmin, kmax):
    """Hulls of 4-6 points (tetrahedra and up) and of 17-18 points, whose 32
    faces fill the shape table's face rows (48 edges): the lane-per-face and
    lane-per-vertex tests of the pair kernels at their extremes.  Contact
    lists bit-exact, forces and 20-step states within 1e-5."""
    print(f"[case] hull box 300 hulls of {kmin}-{kmax} points", flush=True)
    deck = S.hull_box_deck(300, seed=43, kmin=kmin, kmax=kmax)
    if kmax == 18:
        assert max(s.num_faces for s in deck["shapes"]) == T.SHAPE_MAX_FACE
    dev, ora = make_pair(deck)
    lg, ng = dev.debug_contacts()
    lo, no = ora.contacts()
    assert np.array_equal(ng, no)
    assert no.sum() > 0
    for i in np.nonzero(no)[0]:
        k = min(int(no[i]), 32)
        assert np.array_equal(lg[i, :k], lo[i, :k]), f"particle {i}"
    g = dev.debug_forces()
    o = ora.forces()
    assert np.array_equal(g[4], o[4])
    assert float(np.abs(o[0]).max()) > 0.0
    for a, b, what in zip(g[:4], o[:4], ("F_pp", "T_pp", "F_wall", "T_wall")):
        scale = max(float(np.max(np.abs(b))), 1e-30)
        err = rel_err(a, b, 1e-3 * scale)
        print(f"hulls of {kmin}-{kmax} points {what}: max rel err {err:.3e}")
        assert err <= TOL, f"{what}: {err}"
    dev.step(20)
    ora.step(20)
    _compare_states(f"hulls of {kmin}-{kmax} points, 20 steps", deck, dev, ora)
    dev.close()
    ora.close()


@pytest.mark.parametrize("steps", [1, 20])
def test_hull_box_steps_match_oracle(hull_box, steps):
    dev, ora = make_pair(hull_box)
    dev.step(steps)
    ora.step(steps)
    _compare_states(f"hull box {steps} steps", hull_box, dev, ora)
    assert not dev.status() & bdem.lib.WARN_SORT
    dev.close()
    ora.close()


@pytest.mark.parametrize("steps", [1, 40])
def test_hull_mill_with_lifters_matches_oracle(steps):
    print(f"[case] hull mill 600 hulls with lifters, {steps} steps", flush=True)
    deck = S.hull_mill_deck(600, seed=23, R=8.0, L=4.0, lifter_w=1.0, lifter_h=1.2, rmin=0.25, rmax=0.4,
                            dt=1e-5, fill_scale=0.8)
    rad = np.array([s.radius for s in deck["shapes"]], np.float32)
    deck = S.place_at_lifters(dict(deck, radius=0.5 * rad), seed=24, overlap=0.0)
    deck["radius"] = None
    dev, ora = make_pair(deck)
    dev.drum_rotation(True)
    ora.set_drum_rotation(True)
    dev.set_lifter_rotation(deck["rad_per_step"], deck["lifter_about"])
    dev.step(steps)
    ora.run(1, deck["rad_per_step"], deck["lifter_about"])
    assert ora.lifter_contacts() > 30          # hulls inside every lifter face and top edge
    if steps > 1:
        ora.run(steps - 1, deck["rad_per_step"], deck["lifter_about"])
    _compare_states(f"hull mill {steps} steps", deck, dev, ora)
    assert not dev.status() & bdem.lib.WARN_SORT
    dev.close()
    ora.close()


def test_per_pid_table_of_type_rows_equals_type_table():
    print("[case] mixed poly deck, type table vs per-P_ID copy, 20 steps", flush=True)
    deck = S.poly_deck(1000, seed=54, L=16.0, shape="mixed", spacing=1.25)
    rows = [S.shape_from_pobj(deck["pobj"][t]) for t in deck["ptype"]]
    d2 = dict(deck, shapes=(T.PolyShape * len(rows))(*rows))
    a, b = make_device(deck), make_device(d2)
    assert a.num_poly_shapes() == (2, False) and b.num_poly_shapes() == (1000, True)
    a.step(20)
    b.step(20)
    sa, sb = a.get_state(), b.get_state()
    for key in ("pos", "vel", "avel", "quat", "pid"):
        assert np.array_equal(sa[key], sb[key]), key
    a.close()
    b.close()


def test_shape_table_rejections_keep_the_context():
    deck = S.hull_box_deck(64, seed=44, L=8.0)
    dev = make_device(deck)
    short = (T.PolyShape * 10)(*list(deck["shapes"])[:10])
    with pytest.raises(bdem.BdemError):
        dev.set_poly_shapes(short)          # live P_IDs up to 63 have no row
    bad = (T.PolyShape * 64)(*deck["shapes"])
    bad[5].num_edges = 65
    with pytest.raises(bdem.BdemError):
        dev.set_poly_shapes(bad)
    assert dev.num_poly_shapes() == (64, True)
    dev.step(2)
    assert not dev.status() & bdem.lib.WARN_SORT
    dev.close()


def test_separating_face_memory_shrinks_the_pair_queue_only():
    """The face that separated a listed pair is tried first at the next step:
    fewer pairs reach the wave-per-pair kernel, and the states are those of
    a context whose memory starts empty at every step (the oracle)."""
    print("[case] hull box 600, separating-face memory", flush=True)
    deck = S.hull_box_deck(600, seed=41)
    dev, ora = make_pair(deck)
    dev.step(1)
    listed1, wave1 = dev.debug_poly_queues()
    dev.step(1)
    listed2, wave2 = dev.debug_poly_queues()
    print(f"step 1: {listed1} listed, {wave1} to the wave kernel; step 2: {listed2}, {wave2}")
    assert wave2 < wave1
    ora.step(2)
    _compare_states("hull box memory 2 steps", deck, dev, ora)
    dev.close()
    ora.close()


def test_use_hist_is_accepted_and_changes_nothing():
    print("[case] mixed poly deck with use_hist, 20 steps", flush=True)
    deck = S.poly_deck(1000, seed=54, L=16.0, shape="mixed", spacing=1.25)
    sim2 = T.SimulationInfo.from_buffer_copy(deck["sim"])
    sim2.use_hist = True
    a, b = make_device(deck), make_device(dict(deck, sim=sim2))
    a.step(20)
    b.step(20)
    sa, sb = a.get_state(), b.get_state()
    for key in ("pos", "vel", "avel", "quat", "pid"):
        assert np.array_equal(sa[key], sb[key]), key
    a.close()
    b.close()
