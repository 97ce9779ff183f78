"""Per-particle polyhedron shapes on the CPU: the hull builder
(bdem_poly_shapes_from_points) against scipy's Quickhull and closed forms,
and the oracle's shape-table indexing against its per-type path.

The builder is checked by what the geometry fixes: scipy's hull volume and
area for random point sets, Euler's V - E + F = 2, every vertex behind every
face plane, outward normals, and the cube's mass, bounding radius and
inverse inertia 6 / (m a^2) in closed form.  The oracle's per-particle table
is pinned by running a two-type deck once through the type table and once
through a per-P_ID table holding each particle's type geometry: both must
give the same states bit for bit after the sort has permuted the slots (a
row looked up by slot instead of by P_ID fails this).
"""
import math

import numpy as np
import pytest

import bdem
from bdem import scenarios as S
from bdem import types as T
from oracle_py import Oracle


def _f3(v):
    return np.array([v.x, v.y, v.z], np.float64)


def _random_sets(seed, count, kmin=4, kmax=16):
    rng = np.random.default_rng(seed)
    sets = []
    for _ in range(count):
        k = int(rng.integers(kmin, kmax + 1))
        v = rng.normal(size=(k, 3)) * rng.uniform(0.5, 2.0, 3)
        sets.append(v.astype(np.float32))
    return sets


def test_hulls_match_quickhull_volume_area_and_topology():
    from scipy.spatial import ConvexHull
    sets = _random_sets(3, 300)
    shapes = bdem.lib.poly_shapes_from_points(sets, 2700.0)
    for pts, s in zip(sets, shapes):
        h = ConvexHull(pts.astype(np.float64))
        vol = s.mass / (2700.0 * 1e-6)
        assert abs(vol - h.volume) <= 2e-5 * h.volume
        area = sum(s.face[f].area for f in range(s.num_faces))
        assert abs(area - h.area) <= 2e-5 * h.area
        assert s.num_vertex == len(h.vertices)
        assert s.num_vertex - s.num_edges + s.num_faces == 2
        com = _f3(s.COM)
        V = np.array([_f3(s.Local_Vertex[k]) for k in range(s.num_vertex)]) + com
        assert abs(s.radius - np.max(np.linalg.norm(V - com, axis=1))) <= 1e-6 * s.radius
        for f in range(s.num_faces):
            F = s.face[f]
            n, c = _f3(F.normal), _f3(F.centroid)
            assert abs(np.linalg.norm(n) - 1.0) < 1e-6
            assert np.dot(n, c - com) > 0.0                       # outward
            assert np.all((V - c) @ n <= 1e-5 * s.radius)        # every vertex behind the plane
            order = [F.vertex_Order[q] for q in range(F.num_vertex)]
            # counter-clockwise seen from outside: consecutive cross products along n
            for q in range(F.num_vertex):
                a, b = V[order[q]] - c, V[order[(q + 1) % F.num_vertex]] - c
                assert np.dot(np.cross(a, b), n) > 0.0
        edges = {tuple(sorted((s.edge[e].point_vindex[0], s.edge[e].point_vindex[1]))) for e in range(s.num_edges)}
        assert len(edges) == s.num_edges


def test_cube_mass_radius_and_inertia_closed_form():
    a = 2.0
    corners = np.array([[x, y, z] for x in (-1, 1) for y in (-1, 1) for z in (-1, 1)], np.float64)
    pts = (corners * (a / 2.0) + np.array([5.0, -3.0, 1.0])).astype(np.float32)
    s = bdem.lib.poly_shapes_from_points([pts], 2500.0)[0]
    m = 2500.0 * a ** 3 * 1e-6
    assert s.num_vertex == 8 and s.num_faces == 6 and s.num_edges == 12
    assert all(s.face[f].num_vertex == 4 for f in range(6))     # coplanar points merged into quads
    assert abs(s.mass - m) <= 1e-6 * m
    assert abs(s.radius - a * math.sqrt(3.0) / 2.0) <= 1e-6
    assert np.allclose(_f3(s.COM), [5.0, -3.0, 1.0])
    Iinv = np.array(s.InertiaT[:]).reshape(3, 3)
    assert np.allclose(Iinv, np.eye(3) * 6.0 / (m * a * a), rtol=1e-5, atol=1e-3)
    assert abs(s.vbound) == 0.0 or s.vbound >= s.radius       # vbound is the library's to set


def test_builder_drops_interior_points_and_rejects_degenerate_sets():
    pts = np.vstack([S.CUBE_VERTS, [[0.1, 0.2, -0.3]]]).astype(np.float32)
    s = bdem.lib.poly_shapes_from_points([pts], 1000.0)[0]
    assert s.num_vertex == 8
    flat = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0], [0.5, 0.2, 0]], np.float32)
    with pytest.raises(bdem.BdemError):
        bdem.lib.poly_shapes_from_points([flat], 1000.0)
    with pytest.raises(bdem.BdemError):
        bdem.lib.poly_shapes_from_points([flat[:3]], 1000.0)


def test_hull_points_follow_the_configuration():
    pts, k, r = S.hull_points(5, 2000, 8, 16, 0.15, 0.30)
    assert k.min() == 8 and k.max() == 16
    assert np.all((r >= 0.15) & (r <= 0.30))
    shapes = bdem.lib.poly_shapes_from_points((pts, k), 2700.0)
    rad = np.array([s.radius for s in shapes])
    assert np.allclose(rad, r, rtol=1e-5)                     # bounding radius about the vertex mean


def _per_pid_copy(deck):
    d = dict(deck)
    rows = [S.shape_from_pobj(deck["pobj"][t]) for t in deck["ptype"]]
    d["shapes"] = (T.PolyShape * len(rows))(*rows)
    return d


@pytest.mark.parametrize("shape", ["mixed", "cubes2"])
def test_oracle_per_pid_table_equals_type_table(shape):
    deck = S.poly_deck(216, seed=61, L=12.0, shape=shape, spacing=1.25 if shape == "mixed" else 1.19,
                       tilt=None if shape == "mixed" else 0.5)
    a = Oracle(deck)
    b = Oracle(_per_pid_copy(deck))
    a.step(15)
    b.step(15)
    sa, sb = a.state(), b.state()
    assert not np.array_equal(sa["pid"], np.arange(len(sa["pid"])))   # the sort has permuted the slots
    for key in ("pos", "vel", "avel", "quat", "pid"):
        assert np.array_equal(sa[key], sb[key]), key
    assert a.contacts()[1].sum() > 0
    a.close()
    b.close()


def test_oracle_rows_follow_pid():
    """A table whose rows differ per particle: shifting every particle's row
    by one P_ID changes the run (the rows are read, and per particle)."""
    deck = S.hull_box_deck(64, seed=43, L=8.0, overlap=0.3)
    base = Oracle(deck)
    assert base.contacts()[1].sum() > 0
    base.step(10)
    s0 = base.state()
    d2 = dict(deck)
    rows = list(deck["shapes"])
    d2["shapes"] = (T.PolyShape * len(rows))(*(rows[1:] + rows[:1]))
    other = Oracle(d2)
    other.step(10)
    assert not np.array_equal(s0["pos"], other.state()["pos"])
    base.close()
    other.close()
