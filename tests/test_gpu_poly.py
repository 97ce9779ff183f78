"""GPU parity for polyhedra (particle_type 1): the rotate / list / pair /
finish kernels against the CPU restatement.

Contact lists are bit-exact.  Forces and states use relative tolerance 1e-5;
the detection logic, the pair law, the walls and the tensor integrator
perform the same IEEE operations in the same order as the oracle, so
differences come only from sin/cos in the quaternion update (device vs host
libm) and what they feed.
"""
import math

import numpy as np
import pytest

import bdem
from bdem import scenarios as S
from helpers import make_pair, rel_err, by_pid, oracle_positions_out

pytestmark = pytest.mark.gpu

TOL = 1e-5
R_CUBE = math.sqrt(3.0) / 2.0

DECKS = {
    "cube_random": lambda: S.poly_deck(1000, seed=51, L=16.0, shape="cube", spacing=2 * R_CUBE * 0.8),
    "cubes2_faces": lambda: S.poly_deck(1000, seed=52, L=18.0, shape="cubes2", tilt=0.5, spacing=1.19,
                                        jitter=0.01),
    "hull": lambda: S.poly_deck(1000, seed=53, L=16.0, shape="hull", size=1.7, spacing=1.2),
    "mixed": lambda: S.poly_deck(1000, seed=54, L=16.0, shape="mixed", spacing=1.25),
}


@pytest.fixture(params=sorted(DECKS))
def pair(request):
    print(f"[case] poly deck={request.param}", flush=True)
    deck = DECKS[request.param]()
    dev, ora = make_pair(deck)
    yield request.param, deck, dev, ora
    dev.close()
    ora.close()


def test_poly_contact_lists_bit_exact(pair):
    name, deck, dev, ora = pair
    lg, ng = dev.debug_contacts()
    lo, no = ora.contacts()
    assert np.array_equal(ng, no)
    assert no.sum() > 0
    for i in np.nonzero(no)[0]:
        k = min(int(no[i]), 32)
        assert np.array_equal(lg[i, :k], lo[i, :k]), f"particle {i}"


def test_poly_forces_match(pair):
    name, deck, dev, ora = pair
    g = dev.debug_forces()
    o = ora.forces()
    assert np.array_equal(g[4], o[4])
    for a, b, what in zip(g[:4], o[:4], ("F_pp", "T_pp", "F_wall", "T_wall")):
        scale = max(float(np.max(np.abs(b))), 1e-30)
        err = rel_err(a, b, 1e-3 * scale)
        exact = float(np.mean(a.astype(np.float64) == b))
        print(f"{name} {what}: max rel err {err:.3e}, bit-exact fraction {exact:.4f}")
        assert err <= TOL, f"{name} {what}: {err}"


@pytest.mark.parametrize("steps", [1, 20])
def test_poly_step_states_match(pair, steps):
    name, deck, dev, ora = pair
    dev.step(steps)
    ora.step(steps)
    sg = by_pid(dev.get_state())
    so = by_pid(ora.state())
    assert np.array_equal(sg["pid"], so["pid"])
    so["pos"] = oracle_positions_out(deck, so)
    for key in ("pos", "vel", "avel", "quat"):
        scale = max(float(np.max(np.abs(so[key]))), 1e-30)
        err = rel_err(sg[key], so[key], 1e-3 * scale)
        exact = float(np.mean(sg[key].astype(np.float64) == so[key]))
        print(f"{name} {steps} steps {key}: max rel err {err:.3e}, bit-exact {exact:.4f}")
        assert err <= TOL, f"{name} {key}: {err}"
    assert not dev.status() & bdem.lib.WARN_SORT


I = np.array([1, 0, 0, 0], np.float32)
S2 = math.sqrt(2.0) / 2.0


@pytest.mark.parametrize("name,qa,dx,qb,big", [
    ("face-face", I, 1.49, I, True),
    ("edge-face", S.axis_quat((0, 0, 1), 45), S2 + 0.99, I, True),
    ("edge-edge", S.axis_quat((0, 0, 1), 45), 2 * S2 - 0.01, S.axis_quat((0, 1, 0), 45), False),
])
def test_poly_closed_form_pairs(name, qa, dx, qb, big):
    """Two cubes 0.01 into each other: the GPU pair force equals the
    oracle's, which the CPU suite pins to the hand-computed depth."""
    deck = S.poly_pair_deck((5, 5, 5), (5 + dx, 5, 5), qa, qb,
                            shape_b=(S.CUBE_VERTS * 2, S.CUBE_FACES) if big else None, gravity=False)
    dev, ora = make_pair(deck)
    g = dev.debug_forces()
    o = ora.forces()
    for a, b in zip(g[:2], o[:2]):
        assert rel_err(a, b, 1e-3 * max(float(np.abs(b).max()), 1e-30)) <= TOL
    assert float(np.abs(o[0]).max()) > 500.0   # Kn * 0.01 = 1000 on each cube
    dev.close()
    ora.close()


def _rot_np(q, v):
    """Turn v by the unit quaternion q = (w, x, y, z): the rotation matrix of
    q, as the CPU closed-form tests use it."""
    w, x, y, z = (np.float64(c) for c in q)
    R = np.array([[w * w + x * x - y * y - z * z, 2 * (x * y - w * z), 2 * (x * z + w * y)],
                  [2 * (x * y + w * z), w * w - x * x + y * y - z * z, 2 * (y * z - w * x)],
                  [2 * (x * z - w * y), 2 * (y * z + w * x), w * w - x * x - y * y + z * z]])
    return v @ R.T


def test_point_cloud_is_the_turned_shapes():
    """bdem_get_poly_vertices: every vertex is the type's local vertex turned
    by the particle's orientation plus its centre of mass."""
    print("[case] poly point cloud, mixed deck 20 steps", flush=True)
    deck = DECKS["mixed"]()
    dev = bdem.Device(deck["wobj"], deck["pobj"], deck["sim"])
    dev.set_state(deck["pos"], deck["ptype"], deck["pid"], vel=deck["vel"], avel=deck["avel"], quat=deck["quat"])
    dev.step(20)
    st = dev.get_state()
    verts, nv, pid = dev.poly_vertices()
    assert np.array_equal(pid, st["pid"])
    pobj = deck["pobj"]
    worst = 0.0
    for i in range(len(pid)):
        p = pobj[st["ptype"][i]]
        assert nv[i] == p.num_vertex
        com = np.array([p.COM.x, p.COM.y, p.COM.z])
        centre = st["pos"][i].astype(np.float64) + com          # the library returns position - COM
        local = np.array([[p.Local_Vertex[k].x, p.Local_Vertex[k].y, p.Local_Vertex[k].z]
                          for k in range(p.num_vertex)])
        q = st["quat"][i].astype(np.float64)
        want = _rot_np(q / np.linalg.norm(q), local) + centre
        worst = max(worst, float(np.abs(verts[i, :nv[i]] - want).max()))
        assert np.all(verts[i, nv[i]:] == 0.0)
    print(f"point cloud: max |device - numpy| {worst:.3e} cm")
    assert worst < 2e-5 * max(16.0, 1.0)     # coordinates up to the box size, 16 cm
    with pytest.raises(bdem.lib.BdemError):
        dev.poly_vertices(max_vertex=int(nv.max()) - 1)
    dev.close()
