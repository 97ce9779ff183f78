"""Project decks and state files (include/bdem_deck.h), on the CPU.

The fixtures under tests/golden/decks are the reference's own project data
files (Projects/, unchanged; Start_SPH_60 and Start_SPH_90 are left out for
size).  The checks pin the reader against what the files and the geometry
fix: values read from the text by an independent parse, closed forms of the
shapes (volume of a regular dodecahedron, a cube), the drum and lifter
rotation, and the layout of the files the reference itself wrote.
"""
import math
import os
import re

import numpy as np
import pytest

import bdem
from bdem import scenarios as S
from bdem import types as T
from bdem.deck import Deck, read_ksim, read_state, write_state

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DECKS = os.path.join(ROOT, "tests", "golden", "decks")

# World files and whether they give Delta_t in section 2 (the layout the
# reference parser rejects).
WORLDS = {
    ("1.Particles_Box", "Poly_128cm"): 0, ("1.Particles_Box", "Spheres_128cm"): 0,
    ("2.Silo", "SPH_30"): 1, ("2.Silo", "SPH_60"): 1, ("2.Silo", "SPH_90"): 0,
    ("3.BallMill", "Mill_SPH20"): 1, ("3.BallMill", "Mill_SPH30"): 0, ("3.BallMill", "Mill_SPH30_Fill"): 1,
    ("3.BallMill", "Mill_SPH30_NoLifter"): 1,
    ("4.UnitTests", "2P"): 0, ("4.UnitTests", "2PS"): 1, ("4.UnitTests", "Surface"): 1,
    ("4.UnitTests", "Symm"): 1,
}

f32 = np.float32


def deck(project, world):
    return Deck(os.path.join(DECKS, project), world)


def v3(f):
    return np.array([f.x, f.y, f.z], np.float64)


def labelled(path, label):
    """The numbers after `label` in a deck file (an independent parse)."""
    toks = open(path).read().split()
    out = []
    for k, t in enumerate(toks):
        if t == label:
            out.append(float(toks[k + 1]))
    return out


@pytest.mark.parametrize("project,world", sorted(WORLDS))
def test_every_shipped_world_loads(project, world):
    d = deck(project, world)
    assert d.info.dt_in_setup == WORLDS[(project, world)]
    text = open(os.path.join(DECKS, project, "World", world + ".World")).read()
    dt = float(re.search(r"Delta_t\(s\):\s*(\S+)", text).group(1))
    assert d.sim.InitalDelta_t == float(f32(dt))
    n = int(re.search(r"Num_Particles:\s*(\d+)", text).group(1))
    assert d.sim.Num_Particles == n == sum(d.info.num_per_pobj[:d.info.num_pobj])
    # cells: ceil((World_Size - Origin) / Cell_Size), main.cpp:564-566
    o = [float(x) for x in re.search(r"Origin:\s*(\S+)\s+(\S+)\s+(\S+)", text).groups()]
    w = [float(x) for x in re.search(r"World_Size:\s*(\S+)\s+(\S+)\s+(\S+)", text).groups()]
    c = [float(x) for x in re.search(r"Cell_Size:\s*(\S+)\s+(\S+)\s+(\S+)", text).groups()]
    cells = [math.ceil((w[k] - o[k]) / c[k]) for k in range(3)]
    assert [d.sim.num_NNCells.x, d.sim.num_NNCells.y, d.sim.num_NNCells.z] == cells
    assert d.sim.max_size.x == pytest.approx(cells[0] * c[0])
    assert d.sim.Integration_Type == 0 and not d.sim.use_symmetry and not d.sim.use_hist


def test_run_selector():
    assert read_ksim(os.path.join(DECKS, "Simulation.KSim")) == ("1.Particles_Box", "Spheres_128cm", 0, 0, 0)


def test_missing_deck_reports_the_file():
    with pytest.raises(bdem.lib.BdemError, match="cannot open"):
        deck("3.BallMill", "NoSuchWorld")


def test_ball_mill_deck():
    d = deck("3.BallMill", "Mill_SPH30")
    s, i = d.sim, d.info
    assert (i.sim_type, i.force_spec, i.particle_type) == (1, 1, 0)
    assert s.isMillSim and s.EnergyCalc and s.Simulation_Type == 1
    # drum: diameter 90, height 15.10, both caps, axis z
    cyl = d.wobj[0].cyl_geo
    assert d.wobj[0].surface_type == T.SURFACE_CYLINDER
    assert cyl.radius == 45.0 and cyl.height == float(f32(15.10))
    assert cyl.has_top_cap and cyl.has_bot_cap
    assert list(v3(cyl.AXIS)) == [1, 1, 0]
    assert list(v3(cyl.center_bot_cap)) == [45, 45, 0]
    assert np.allclose(v3(cyl.center_top_cap), [45, 45, 15.1], atol=1e-5)
    # Ball_5cm: diameter 5, density 8000 -> mass 4/3 pi r^3 rho 1e-6 (kg, cm)
    p = d.pobj[0]
    pf = os.path.join(DECKS, "3.BallMill", "ParticleObjects", "Ball_5cm.POBJ")
    rho = labelled(pf, "Density:")[0]
    assert p.radius == 2.5
    assert p.mass == float(f32(4.0 / 3.0 * math.pi * 2.5 ** 3 * rho * 1e-6))
    assert p.InertiaT[0] == pytest.approx(1.0 / (0.4 * 2.5 * 2.5 * p.mass), rel=1e-6)
    # spring constants: the file's KN, COR and friction through the KN rules
    cor = labelled(pf, "COR:")
    kn = labelled(pf, "KN:")
    q = S.sphere_pobj(2.5, rho, cor[0], kn[0], 0.25)
    q.mass = p.mass
    q.surface_Kn, q.Dsurface_Kn = kn[1], kn[2]
    q = bdem.lib.derive_spring_constants([q], bdem.lib.SPRING_KN, [cor[0]], [cor[1]], [cor[2]])[0]
    for a, b in ((p.PP[0].Kn, q.PP[0].Kn), (p.PP[0].Cn, q.PP[0].Cn), (p.surface_Cn, q.surface_Cn),
                 (p.Dsurface_Cn, q.Dsurface_Cn), (p.surface_Kn, kn[1]), (p.Dsurface_Kn, kn[2])):
        assert a == pytest.approx(b, rel=1e-7)
    assert p.PP[0].Fric_kinetic == float(f32(0.25)) and p.surface_Fric_static == float(f32(0.30))
    # one revolution at 22 RPM and dt 1e-5: 272,727 steps of RadPerStep
    assert i.mill_rpm == 22.0 and i.power_revs == 1 and i.total_revs == 5
    assert i.steps_per_rev == 272727
    assert abs(i.steps_per_rev * i.rad_per_step + 2 * math.pi) < 2e-5
    assert i.rad_per_step < 0
    omega = 22 * 2 * math.pi / 60
    assert s.MillRotVel == pytest.approx(-omega * 45 * 0.01, rel=1e-6)
    assert i.total_time == pytest.approx(5 * 272727 * 1e-5, rel=1e-6)
    # lifters: 8 copies of Square_Lifter16 turned by 2 pi k / 8 about (45, 45)
    assert len(d.dobj) == 8 and i.dobj_types == 1 and i.dobj_repeat == 8
    about = np.array([45.0, 45.0, 0.0])
    base = np.array([v3(d.dobj[0].vertex[k]) for k in range(8)])
    lf = os.path.join(DECKS, "3.BallMill", "LifterObjects", "Square_Lifter16.DOBJ")
    toks = open(lf).read().split()
    k0 = toks.index("NUM_VERTEX:")
    file_verts = np.array([[float(t) for t in toks[k0 + 3 + 4 * j:k0 + 6 + 4 * j]] for j in range(8)])
    assert np.array_equal(base, file_verts)
    for k in range(8):
        a = 2 * math.pi * k / 8
        rot = np.array([[math.cos(a), math.sin(a), 0], [-math.sin(a), math.cos(a), 0], [0, 0, 1]])
        want = (base - about) @ rot.T + about
        got = np.array([v3(d.dobj[k].vertex[j]) for j in range(8)])
        assert np.allclose(got, want, atol=2e-5), k
        com = (v3(d.dobj[0].COM) - about) @ rot.T + about
        assert np.allclose(v3(d.dobj[k].COM), com, atol=2e-5)
        assert d.dobj[k].is_attached_DWorld and not d.dobj[k].is_translating
        for j in range(d.dobj[k].num_faces):
            f = d.dobj[k].faces[j]
            n = v3(f.normal)
            assert abs(np.linalg.norm(n) - 1) < 1e-6
            pts = got[[f.vertex_Order[q] for q in range(4)]]
            assert np.abs((pts - pts[0]) @ n).max() < 1e-4     # normal of the face's plane
            assert f.area == pytest.approx(4.0 * 15.0, rel=1e-5)   # every face is 4 x 15 cm
    # DObject_CylinderBoundR = R - 2 COM.y of the first lifter, cent = (45, 45, 0)
    assert s.DObject_CylinderBoundR == 45.0 - 2 * 2.0
    assert list(v3(s.cent)) == [45, 45, 0]


def test_silo_deck_with_delta_t_in_setup():
    d = deck("2.Silo", "SPH_30")
    s, i = d.sim, d.info
    assert i.dt_in_setup == 1 and s.InitalDelta_t == float(f32(1e-5))
    assert i.sim_type == 2 and not s.isMillSim
    assert [d.info.num_per_pobj[k] for k in range(2)] == [1425, 525]
    assert list(v3(s.Force_Field)) == [0, -980, 0]
    # the hatch is the last world object; the kill plane 5 cm below it
    assert i.hatch_height == 30.0 and s.Silo_Kill_Height == 25.0
    assert i.measure_flow == 1 and i.flow_steps == 8696 and i.kill_particle == 0 and not s.Kill_Particles
    assert i.total_time == 4.0
    assert s.Roll_Res == float(f32(0.01)) and s.Rotation
    assert len(d.dobj) == 2 and d.dobj[0].boundR == float(f32(0.072))
    # the plane surfaces agree with an independent float32 derivation
    for w in d.wobj:
        V = [tuple(v3(w.vertex[k])) for k in range(w.num_vertex)]
        orders = [[w.surfaces[j].vertex_Order[q] for q in range(w.surfaces[j].num_vertex)]
                  for j in range(w.num_surfaces)]
        ref = S.plane_wobj(V, orders)
        for j in range(w.num_surfaces):
            a, b = w.surfaces[j], ref.surfaces[j]
            assert np.allclose(v3(a.normal), v3(b.normal), atol=1e-6)
            assert np.allclose(v3(a.centroid), v3(b.centroid), atol=1e-5)
            assert a.area == pytest.approx(b.area, rel=1e-6)


def test_silo_deck_without_output_format_and_kill():
    d = deck("2.Silo", "SPH_90")
    assert d.info.write_steps == 1000 and d.info.output_format == 3
    assert d.info.kill_particle == 1 and d.sim.Kill_Particles


def test_dodecahedron_deck():
    """Poly_1cm: a regular dodecahedron given about the origin plus an
    offset (Format 1, comma-separated faces)."""
    d = deck("1.Particles_Box", "Poly_128cm")
    p = d.pobj[0]
    assert (p.num_vertex, p.num_faces) == (20, 12)
    assert all(p.face[j].num_vertex == 5 for j in range(12))
    pf = os.path.join(DECKS, "1.Particles_Box", "ParticleObjects", "Poly_1cm.POBJ")
    off = [float(x) for x in re.search(r"OffSet:\s*(\S+)\s+(\S+)\s+(\S+)", open(pf).read()).groups()]
    com = v3(p.COM)
    assert np.allclose(com, off, atol=1e-6)
    V = np.array([v3(p.vertex[k]) for k in range(20)])
    R = np.linalg.norm(V - com, axis=1)
    assert p.radius == pytest.approx(R.max(), rel=1e-6)
    # edge length of the regular dodecahedron with this circumradius
    a = 4 * R.mean() / (math.sqrt(3) * (1 + math.sqrt(5)))
    vol = (15 + 7 * math.sqrt(5)) / 4 * a ** 3
    rho = labelled(pf, "DENSITY:")[0]
    assert p.mass == pytest.approx(rho * vol * 1e-6, rel=2e-4)
    face_area = 0.25 * math.sqrt(5 * (5 + 2 * math.sqrt(5))) * a * a
    for j in range(12):
        f = p.face[j]
        n = v3(f.normal)
        assert np.dot(n, v3(f.centroid) - com) > 0          # outward
        assert f.area == pytest.approx(face_area, rel=2e-4)
    # inertia: diagonal 10 read, inverted, then scaled by 1/(0.4 r^2 m)
    s = 1.0 / (0.4 * p.radius * p.radius * p.mass)
    assert p.InertiaT[0] == pytest.approx(0.1 * s, rel=1e-6) and p.InertiaT[1] == 0.0
    assert p.Local_Vertex[3].x == pytest.approx(p.vertex[3].x - p.COM.x, abs=1e-6)


def test_cube_deck():
    d = deck("4.UnitTests", "2P")
    p = d.pobj[0]
    assert (p.num_vertex, p.num_faces, p.num_edges) == (8, 6, 12)
    assert np.allclose(v3(p.COM), [0.28, 0.28, 0.28], atol=1e-6)
    pf = os.path.join(DECKS, "4.UnitTests", "ParticleObjects", "Cube_1cm.POBJ")
    rho = labelled(pf, "DENSITY:")[0]
    assert p.mass == pytest.approx(rho * 0.56 ** 3 * 1e-6, rel=1e-5)
    for j in range(6):
        assert p.face[j].area == pytest.approx(0.56 * 0.56, rel=1e-5)
        assert abs(abs(np.dot(v3(p.face[j].normal), v3(p.face[j].centroid) - v3(p.COM))) - 0.28) < 1e-5


def test_unit_test_volume_object_without_tolerance():
    d = deck("4.UnitTests", "Surface")
    assert len(d.dobj) == 1 and all(d.dobj[0].faces[j].Tol == float(f32(0.10)) for j in range(3))


# ---- state files ------------------------------------------------------------

def _rows(path, skip, n):
    return np.loadtxt(path, skiprows=skip, max_rows=n, dtype=np.float64)


def test_mill_start_file():
    d = deck("3.BallMill", "Mill_SPH30")
    st = d.read_state()
    h = st["header"]
    assert (h.spec_type, h.rev_num, h.rev_complete, h.time, h.rot_angle) == (3, 0, 0, 0.0, 0.0)
    assert st["rows"] == 242
    assert sorted(st["pid"]) == list(range(242)) and not st["ptype"].any()
    raw = _rows(d.start_path, 3, 242)
    assert np.array_equal(st["pid"], raw[:, 0].astype(np.int32))
    assert np.array_equal(st["pos"], raw[:, 2:5].astype(np.float32))
    assert np.array_equal(st["vel"], raw[:, 5:8].astype(np.float32))
    assert np.array_equal(st["avel"], raw[:, 8:11].astype(np.float32))


def test_silo_start_file_columns():
    """SpecType 4: pos, vel, acc, angular velocity, then |v|, |w|, |a| in
    m/s; the magnitudes fix which columns are which."""
    d = deck("2.Silo", "SPH_30")
    st = d.read_state()
    assert st["header"].spec_type == 4 and st["rows"] == 1950
    raw = _rows(d.start_path, 3, 1950)
    for key, col in (("vel", 14), ("avel", 15), ("acc", 16)):
        mag = np.linalg.norm(st[key].astype(np.float64), axis=1) * 0.01
        assert np.allclose(mag, raw[:, col], rtol=2e-5, atol=1e-9), key


def test_polyhedron_start_file():
    d = deck("4.UnitTests", "2P")
    st = d.read_state()
    assert st["rows"] == 2 and st["header"].spec_type == 3
    assert list(st["pos"][0]) == [5, 6, 10] and list(st["vel"][0]) == [800, 0, 0]
    assert list(st["quat"][0]) == [1, 0.5, 0, 0] and list(st["avel"][1]) == [0, 0, 1000]


def test_writer_reproduces_the_reference_rows(tmp_path):
    """Reading StartMill_SPH30 and writing it back as SpecType 3 gives the
    reference writer's row text, character for character."""
    d = deck("3.BallMill", "Mill_SPH30")
    st = d.read_state()
    out = tmp_path / "state"
    write_state(str(out), st["header"], 1, 0, 0, st["pos"], st["ptype"], st["pid"], vel=st["vel"],
                avel=st["avel"])
    ours = open(out).read().splitlines()
    theirs = open(d.start_path).read().splitlines()
    assert ours[0].rstrip() == theirs[0].rstrip() and ours[1] == theirs[1]
    assert [r.rstrip() for r in ours[3:245]] == [r.rstrip() for r in theirs[3:245]]


@pytest.mark.parametrize("spec", [1, 2, 3, 4])
@pytest.mark.parametrize("ptype_kind", [0, 1])
@pytest.mark.parametrize("sim_type", [0, 1, 2])
def test_state_round_trip(tmp_path, spec, ptype_kind, sim_type):
    rng = np.random.default_rng(spec * 10 + ptype_kind + 100 * sim_type)
    n = 50
    pos = rng.uniform(-50, 50, (n, 3)).astype(np.float32)
    vel = rng.normal(0, 30, (n, 3)).astype(np.float32)
    avel = rng.normal(0, 300, (n, 3)).astype(np.float32)
    acc = rng.normal(0, 1e4, (n, 3)).astype(np.float32)
    q = rng.normal(0, 1, (n, 4)).astype(np.float32)
    ptype = rng.integers(0, 3, n).astype(np.uint32)
    pid = rng.permutation(n).astype(np.int32)
    pid[7] = -1                                   # removed particles are not written
    h = T.StateHeader()
    h.spec_type, h.time, h.rev_num, h.rev_complete, h.rot_angle, h.num_rem = spec, 1.25, 3, 1, -0.5, 17
    for k in range(5):
        h.energy[k] = 0.1 * (k + 1)
    path = str(tmp_path / "s")
    write_state(path, h, sim_type, 0, ptype_kind, pos, ptype, pid, vel=vel, avel=avel, acc=acc,
                quat=q if ptype_kind else None)
    keep = pid != -1
    st = read_state(path, sim_type, 0, ptype_kind, n)
    assert st["rows"] == n - 1
    g = st["header"]
    assert g.spec_type == spec and g.time == 1.25
    if sim_type == 1:
        assert (g.rev_num, g.rev_complete, g.rot_angle) == (3, 1, -0.5)
        assert np.allclose(list(g.energy), [0.1, 0.2, 0.3, 0.4, 0.5])
    if sim_type == 2:
        assert g.num_rem == 17
    assert np.array_equal(st["pid"], pid[keep]) and np.array_equal(st["ptype"], ptype[keep])
    tol = dict(rtol=5e-6, atol=1e-30)
    assert np.allclose(st["pos"], pos[keep], **tol)
    if spec >= 2:
        assert np.allclose(st["vel"], vel[keep], **tol) and np.allclose(st["avel"], avel[keep], **tol)
    if spec == 4:
        assert np.allclose(st["acc"], acc[keep], **tol)
    if ptype_kind:
        assert np.allclose(st["quat"], q[keep], **tol)


def test_deck_scenario_matches_the_state_file():
    d = deck("3.BallMill", "Mill_SPH30")
    sc = d.scenario()
    assert sc["pos"].shape == (242, 3) and len(sc["dobj"]) == 8
    assert sc["lifter_about"] == (45.0, 45.0, 0.0) and sc["rad_per_step"] == d.info.rad_per_step
    assert sc["quat"] is None and sc["sim"].Num_Particles == 242


# ---- grid fills (decks without a start file) ------------------------------

def test_default_grid_fill_is_the_lattice():
    d = deck("1.Particles_Box", "Spheres_128cm")
    st = d.initial_state()
    g = d.init
    n = d.sim.Num_Particles
    assert (g.num.x, g.num.y, g.num.z) == (32, 32, 32) and n == 32768
    pitch = 2 * d.pobj[0].radius + np.array([g.space.x, g.space.y, g.space.z], np.float64)
    k = np.arange(n)
    ijk = np.stack([k % 32, k // (32 * 32), (k // 32) % 32], axis=1)   # x inner, z middle, y outer
    want = np.array([g.start.x, g.start.y, g.start.z]) + ijk * pitch
    assert np.allclose(st["pos"], want, atol=2e-4)
    assert np.array_equal(st["pid"], k) and not st["ptype"].any()
    assert np.all(st["vel"] == np.float32([0.1, 0.0, 0.1]))


def _random_grid_deck(tmp_path, project, world, spec):
    """A copy of a shipped project whose World file asks for GridType 1."""
    import shutil
    dst = tmp_path / project
    shutil.copytree(os.path.join(DECKS, project), dst)
    wf = dst / "World" / (world + ".World")
    text = open(wf).read()
    text = re.sub(r"(GridType\(1=ran\):\s*)0", r"\g<1>1", text, count=1)
    for key, val in spec.items():
        text = re.sub(key, val, text, count=1)
    open(wf, "w").write(text)
    return Deck(str(dst), world)


def _libc_random_grid(n, nx, ny, nz, n_types, counts, orient):
    """Set_Position_Random_Grid's draws (Device/DeviceInterface.cu:435-575)
    with the C library's own rand(), seeded as an unseeded program starts."""
    import ctypes
    libc = ctypes.CDLL("libc.so.6")
    libc.srand(1)
    rmax = np.float32(2147483647)
    quats, types, ids = [], [], []
    sel, cnt = -1, [0] * n_types
    for k in range(nx * ny * nz):
        if orient:
            w = np.float32(np.float32(np.float32(libc.rand()) / rmax) * 30.0)
            x, y, z = (np.float32(np.float32(libc.rand()) / rmax) for _ in range(3))
            quats.append((w, x, y, z))
        sel = 0 if sel + 1 >= n_types else sel + 1
        pty = sel
        if cnt[pty] >= counts[pty]:
            pty = next(i for i in range(n_types) if cnt[i] <= counts[i] - 1)
        cnt[pty] += 1
        types.append(pty)
        ids.append(k)
    types, ids = np.array(types[:n]), np.array(ids[:n])
    for _ in range(n):
        a = libc.rand() % (n - 1)
        b = libc.rand() % (n - 1)
        types[a], types[b] = types[b], types[a]
        ids[a], ids[b] = ids[b], ids[a]
    return quats[:n], types, ids


def test_random_grid_fill_matches_the_c_library_draws(tmp_path):
    d = _random_grid_deck(tmp_path, "1.Particles_Box", "Poly_128cm", {})
    st = d.initial_state()
    n = d.sim.Num_Particles
    quats, types, ids = _libc_random_grid(n, 16, 32, 16, 1, [n], True)
    assert np.array_equal(st["ptype"], types) and np.array_equal(st["pid"], ids)
    assert sorted(st["pid"]) == list(range(n))
    q = st["quat"].astype(np.float64)
    assert np.allclose(np.linalg.norm(q, axis=1), 1.0, atol=1e-6)
    # make_quaterion(theta deg, axis) with the header's pi 3.14152, then normalised
    for k in (0, 1, 2, 777, n - 1):
        w, x, y, z = (float(v) for v in quats[k])
        th = np.float32(np.float32(2.0 * 3.14152 / 360.0) * np.float32(w))
        c, s = np.cos(th / 2), np.sin(th / 2)
        raw = np.array([c, s * x, s * y, s * z])
        assert np.allclose(q[k], raw / np.linalg.norm(raw), atol=2e-6)
    # a turn of up to 30 degrees about an unnormalised axis in [0,1]^3, the
    # quaternion normalised afterwards: the turn reaches 2 atan(sqrt(3) tan 15 deg)
    angles = 2 * np.degrees(np.arccos(np.clip(q[:, 0], -1, 1)))
    assert angles.max() <= 2 * np.degrees(np.arctan(np.sqrt(3) * np.tan(np.radians(15.0)))) + 1e-3
    assert angles.mean() > 10.0


def test_random_grid_fill_alternates_types(tmp_path):
    d = _random_grid_deck(tmp_path, "2.Silo", "SPH_30",
                          {r"Read_FromFile:\s*1\s*\n\s*FileName:\s*\S+":
                           "Read_FromFile: 0\nGridType(1=ran): 1\nGridSize(X,Y,Z): 13 30 5\n"
                           "Start(X,Y,Z): 0.55 30.60 0.55\nSpace: 0.05 0.05 0.20\nVelocity: 0.0 -50.0 25.0"})
    st = d.initial_state()
    n = d.sim.Num_Particles
    _, types, ids = _libc_random_grid(n, 13, 30, 5, 2, [1425, 525], False)
    assert np.array_equal(st["ptype"], types) and np.array_equal(st["pid"], ids)
    assert np.bincount(st["ptype"]).tolist() == [1425, 525]


def test_spec5_rows_are_position_plus_file_vertices(tmp_path):
    """SpecType 5 (Output_SystemState, Input/FileIO.h:253-263): P_ID, type,
    position, then every vertex of the type as vertex + position, the file
    frame's vertices without the orientation, as the reference writes them."""
    rng = np.random.default_rng(5)
    pobj = [S.poly_pobj(S.CUBE_VERTS * 1.5, S.CUBE_FACES, 2500, 0.5, 1e5, 0.4),
            S.poly_pobj(*S.hull_shape(3), 2500, 0.5, 1e5, 0.4)]
    n = 40
    pos = rng.uniform(-20, 20, (n, 3)).astype(np.float32)
    ptype = rng.integers(0, 2, n).astype(np.uint32)
    pid = np.arange(n, dtype=np.int32)
    pid[3] = -1
    h = T.StateHeader()
    h.spec_type, h.time = 5, 0.5
    path = str(tmp_path / "s5")
    write_state(path, h, 0, 0, 1, pos, ptype, pid, quat=rng.normal(0, 1, (n, 4)).astype(np.float32), pobj=pobj)
    rows = [ln.split() for ln in open(path).read().splitlines()[3:] if ln.strip()]
    assert len(rows) == n - 1
    for row, i in zip(rows, [k for k in range(n) if pid[k] != -1]):
        assert int(row[0]) == pid[i] and int(row[1]) == ptype[i]
        assert np.allclose([float(x) for x in row[2:5]], pos[i], rtol=1e-5, atol=1e-5)
        p = pobj[ptype[i]]
        want = np.array([[p.vertex[k].x, p.vertex[k].y, p.vertex[k].z] for k in range(p.num_vertex)],
                        np.float32) + pos[i]
        got = np.array([float(x) for x in row[5:]]).reshape(-1, 3)
        assert got.shape == want.shape and np.allclose(got, want, rtol=1e-5, atol=1e-5)
