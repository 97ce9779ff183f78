"""Seeded synthetic decks in the reference's units (cm, kg, s).

Builds the object tables a reference caller would pass to
Device_Set_SimulationData (Device/DeviceInterface.h:16-19) plus the initial
particle state:

  box_deck   8,192 monodisperse spheres in a closed 50 cm cube (box fill)
  silo_deck  1,048,576 polydisperse spheres in a cylindrical silo over a
             30-degree conical hopper with a closed 6 cm outlet
  random_deck  small dense random packs for parity tests

Derived quantities follow the reference's host code: wall normals, areas and
centroids as ReadWObject computes them (Input/KSimulationData.cpp:841-875),
sphere mass 4/3 pi r^3 rho 1e-6 (Input/KSimulationData.cpp:1096), damping
from restitution (Set_SpringConstantsParticleKN / Surface_KN,
Input/KSimulationData.cpp:950-1045), inverse inertia 1/(0.4 r^2 m)
(main.cpp:615-616), grid size ceil((worldSize - origin)/cell)
(main.cpp:563-569), max_size (main.cpp:820-822).

Random numbers: splitmix64, one stream per quantity, uniform float
(z >> 40) * 2**-24.
"""
import math

import numpy as np

from . import types as T

GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)

STREAM_POS, STREAM_JITTER, STREAM_VEL, STREAM_RADIUS, STREAM_SHAPE, STREAM_MATERIAL = range(6)


def splitmix64(seed, stream, n):
    """n consecutive splitmix64 outputs of stream `stream` of `seed`."""
    with np.errstate(over="ignore"):
        base = np.uint64((seed * 64 + stream) & 0xFFFFFFFFFFFFFFFF)
        z = base + GOLDEN * (np.arange(1, n + 1, dtype=np.uint64))
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return z


def uniform(seed, stream, n, lo=0.0, hi=1.0):
    z = splitmix64(seed, stream, n)
    u = (z >> np.uint64(40)).astype(np.float64) * (2.0 ** -24)
    return lo + (hi - lo) * u


# ---------------------------------------------------------------------------
# reference host derivations
# ---------------------------------------------------------------------------
def sphere_mass(radius_cm, density):
    """Input/KSimulationData.cpp:1096 (double, stored as float)."""
    return np.float32((4.0 / 3.0) * 2.0 * math.asin(1.0) * radius_cm ** 3 * density * 1e-6)


def cn_from_cor(kn, m, cor):
    """Damping from restitution, Input/KSimulationData.cpp:950-1045 (double)."""
    if cor >= 1.0:
        return 0.0
    if 0.30 <= cor <= 0.90:
        cor -= 0.038080
    pi = 2.0 * math.asin(1.0)
    return -2.0 * math.log(cor) * math.sqrt(kn * m) / math.sqrt(math.log(cor) ** 2 + pi * pi)


def inertia_inv(radius, mass):
    """main.cpp:615-616 in float arithmetic."""
    r = np.float32(radius)
    m = np.float32(mass)
    return np.float32(np.float32(1.0) / (np.float32(np.float32(np.float32(0.40) * r) * r) * m))


def sphere_pobj(radius, density, cor, kn, mu, cor_s=None, kn_s=None, mu_s=None):
    """A sphere POBJ as ReadPObject_Sphere + Set_SpringConstants*_KN build it
    (ForceSpec 1, same-type pair: reduced mass m/2, Input/KSimulationData.cpp:958)."""
    p = T.POBJ()
    p.radius = float(radius)
    p.mass = float(sphere_mass(radius, density))
    m = float(np.float32(p.mass))
    kn_s = kn if kn_s is None else kn_s
    cor_s = cor if cor_s is None else cor_s
    mu_s = mu if mu_s is None else mu_s
    p.PP[0].Kn = kn
    p.PP[0].Cn = cn_from_cor(float(np.float32(kn)), 1.0 / (1.0 / m + 1.0 / m), cor)
    p.PP[0].Fric_static = mu
    p.PP[0].Fric_kinetic = mu
    p.surface_Kn = kn_s
    p.surface_Cn = cn_from_cor(float(np.float32(kn_s)), m, cor_s)
    p.surface_Fric_static = mu_s
    p.surface_Fric_kinetic = mu_s
    p.Dsurface_Kn = kn_s
    p.Dsurface_Cn = p.surface_Cn
    p.Dsurface_Fric_static = mu_s
    p.Dsurface_Fric_kinetic = mu_s
    p.InertiaT[0] = float(inertia_inv(p.radius, p.mass))
    return p


def plane_wobj(vertices, surfaces, tol=1.0):
    """A planar WOBJ as ReadWObject derives it (Input/KSimulationData.cpp:712-875):
    quad normal = normalise(cross(V2, V1)), V1 = v1-v0, V2 = v2-v1; area = sum of
    the two triangle halves; centroid = (1/nv) * sum, all in float arithmetic.
    `surfaces` are 0-based vertex orders."""
    f = np.float32
    w = T.WOBJ()
    w.surface_type = T.SURFACE_PLANE
    w.num_vertex = len(vertices)
    V = np.asarray(vertices, dtype=np.float32)
    for k, v in enumerate(V):
        w.vertex[k] = T.f3(*v)
    w.num_surfaces = len(surfaces)
    for j, order in enumerate(surfaces):
        s = w.surfaces[j]
        s.Tol = tol
        s.num_vertex = len(order)
        total = np.zeros(3, np.float32)
        for q, o in enumerate(order):
            s.vertex_Order[q] = o
            total = (total + V[o]).astype(np.float32)
        if len(order) == 3:
            V1 = V[order[2]] - V[order[0]]
            V2 = V[order[1]] - V[order[2]]
            n = np.cross(V2, V1).astype(np.float32)
            ln = f(np.sqrt(f(np.dot(n, n))))
            area = f(f(0.5) * ln)
        else:
            V1 = (V[order[1]] - V[order[0]]).astype(np.float32)
            V2 = (V[order[2]] - V[order[1]]).astype(np.float32)
            V3 = (V[order[3]] - V[order[2]]).astype(np.float32)
            V4 = (V[order[0]] - V[order[3]]).astype(np.float32)
            n = np.cross(V2, V1).astype(np.float32)
            ln = f(np.sqrt(f(np.dot(n, n))))
            n2 = np.cross(V3, V4).astype(np.float32)
            area = f(f(f(0.5) * ln) + f(f(0.5) * f(np.sqrt(f(np.dot(n2, n2))))))
        nn = (n / ln).astype(np.float32)
        s.normal = T.f3(*nn)
        s.area = float(area)
        c = (f(1.0) / f(len(order))) * total
        s.centroid = T.f3(*c.astype(np.float32))
    return w


def infinite_plane_wobj(point, normal):
    """One infinite plane (area == -1 selects Collision_Detection_Sphere_Plane,
    DEM_Kernels_Compute.cuh:319-330); vertex[0] is the plane point."""
    w = T.WOBJ()
    w.surface_type = T.SURFACE_PLANE
    w.num_vertex = 1
    w.vertex[0] = T.f3(*point)
    w.num_surfaces = 1
    s = w.surfaces[0]
    s.num_vertex = 0
    s.area = -1.0
    s.normal = T.f3(*normal)
    s.centroid = T.f3(*point)
    s.Tol = 1.0
    return w


def cylinder_wobj(radius, height, axis_normal, offset, top=False, bot=False):
    """A macro cylinder as ReadWObject builds it (Input/KSimulationData.cpp:772-835):
    radius = diameter/2, AXIS = (1,1,1) - n_bot, c_bot = offset + R*(1,1,1 - n_bot)."""
    w = T.WOBJ()
    w.surface_type = T.SURFACE_CYLINDER
    w.num_surfaces = 1
    g = w.cyl_geo
    g.radius = float(radius)
    g.height = float(height)
    g.has_top_cap = top
    g.has_bot_cap = bot
    nb = np.asarray(axis_normal, np.float32)
    g.normal_bot_cap = T.f3(*nb)
    g.normal_top_cap = T.f3(*(-nb))
    axis = (np.ones(3, np.float32) - nb).astype(np.float32)
    g.AXIS = T.f3(*axis)
    cb = (np.asarray(offset, np.float32) + np.float32(radius) * axis).astype(np.float32)
    g.center_bot_cap = T.f3(*cb)
    ct = (cb + nb * np.float32(height)).astype(np.float32)
    g.center_top_cap = T.f3(*ct)
    return w


def sim_info(n, dt, gravity, origin, world_size, cell, rotation=True, roll_res=0.0,
             kill=False, kill_height=0.0, sim_type=0):
    s = T.SimulationInfo()
    s.particle_type = 0
    s.Num_Particles = int(n)
    s.Num_ParticleObjects = 1
    s.Rotation = bool(rotation)
    s.InitalDelta_t = float(dt)
    s.Force_Field = T.f3(*gravity)
    s.Roll_Res = float(roll_res)
    s.worldOrigin = T.f3(*origin)
    s.cellSize = T.f3(cell, cell, cell)
    total = np.asarray(world_size, np.float32) - np.asarray(origin, np.float32)
    ncell = [int(math.ceil(float(np.float32(t) / np.float32(cell)))) for t in total]
    s.num_NNCells = T.Int3(*ncell)
    s.max_size = T.f3(*(np.asarray(ncell, np.float32) * np.float32(cell)))
    s.Kill_Particles = bool(kill)
    s.Silo_Kill_Height = float(kill_height)
    s.Simulation_Type = sim_type
    s.Integration_Type = 0
    return s


def _deck(sim, wobj, pobj, pos, vel, ptype=None, pid=None, radius=None, mass=None, ext=None,
          name=""):
    n = len(pos)
    return dict(
        name=name, sim=sim, wobj=wobj, pobj=pobj, ext=ext,
        pos=np.ascontiguousarray(pos, np.float32),
        vel=np.ascontiguousarray(vel, np.float32),
        avel=np.zeros((n, 3), np.float32),
        ptype=np.zeros(n, np.uint32) if ptype is None else ptype.astype(np.uint32),
        pid=np.arange(n, dtype=np.int32) if pid is None else pid.astype(np.int32),
        radius=None if radius is None else radius.astype(np.float32),
        mass=None if mass is None else mass.astype(np.float32),
    )


def box_walls(L):
    """Six quads of a closed cube [0, L]^3 with the vertex orders of the
    reference's Container_128 deck (Projects/1.Particles_Box/SurfaceObjects/
    Container_128.WOBJ), normals pointing inside."""
    verts = [(0, 0, 0), (0, 0, L), (L, 0, L), (L, 0, 0),
             (0, L, 0), (0, L, L), (L, L, L), (L, L, 0)]
    quads = [(3, 2, 1, 0), (4, 0, 1, 5), (5, 1, 2, 6), (6, 2, 3, 7), (7, 3, 0, 4), (4, 5, 6, 7)]
    return plane_wobj(verts, quads, tol=1.0)


def lsd_ext(kn, e, mu, kt_over_kn=2.0 / 7.0):
    x = T.ExtParams()
    x.contact_law = T.LAW_LSD_COULOMB
    for t in range(8):
        m = x.material[t]
        m.kn, m.kt_over_kn, m.restitution, m.mu = kn, kt_over_kn, e, mu
    return x


def box_deck(law="ref", seed=1, nx=32, ny=16, nz=16, L=50.0, radius=0.5, density=2500.0,
             kn=1e5, e=0.6, mu=0.4, dt=1e-5, spacing=1.5, jitter=0.05, vmax=10.0):
    """Box fill: a seeded jittered lattice in the upper half of a closed cube.
    Defaults are the 8,192-sphere configuration (r 5 mm, rho 2500, 0.5 m box,
    kn 1e5, e 0.6, mu 0.4, dt 1e-5, cell 2r, origin -1)."""
    n = nx * ny * nz
    gi = np.indices((nx, ny, nz)).reshape(3, -1).T.astype(np.float64)
    x0 = np.array([1.75, 26.0, 1.75])
    pos = x0 + gi * spacing
    pos += uniform(seed, STREAM_JITTER, 3 * n, -jitter, jitter).reshape(n, 3)
    vel = uniform(seed, STREAM_VEL, 3 * n, -vmax, vmax).reshape(n, 3)
    cell = 2.0 * radius
    sim = sim_info(n, dt, (0.0, -981.0, 0.0), (-1.0, -1.0, -1.0), (L, L, L), cell,
                   rotation=True, roll_res=0.0)
    pobj = [sphere_pobj(radius, density, e, kn, mu)]
    ext = None if law == "ref" else lsd_ext(kn, e, mu)
    return _deck(sim, [box_walls(L)], pobj, pos, vel, ext=ext, name=f"box{n}-{law}")


STREAM_COUNT = 6


def batch_box_decks(count=512, seed=4, nmin=1024, nmax=65536, rmin=0.2, rmax=0.6, packing=0.3,
                    density=2500.0, dt=1e-5, vmax=10.0, knmin=5e4, knmax=5e5, emin=0.3, emax=0.9,
                    mumin=0.2, mumax=0.8):
    """Independent box fills for a batch: counts log-uniform in [nmin, nmax],
    radii U[rmin, rmax] per particle, each box edge L = (N 4/3 pi E[r^3] /
    packing)^(1/3), a seeded random non-overlapping fill of [0, L]^3 (larger
    spheres first), velocities U(+-vmax), and per box kn log-uniform in
    [knmin, knmax], restitution U[emin, emax], friction U[mumin, mumax];
    linear spring-dashpot with Coulomb friction, kt = 2/7 kn and tangential
    history; hash cell 2 rmax of the box's own radii."""
    from . import lib as L
    u = uniform(seed, STREAM_COUNT, count)
    counts = np.rint(np.exp(math.log(nmin) + (math.log(nmax) - math.log(nmin)) * u)).astype(np.int64)
    mat = uniform(seed, STREAM_MATERIAL, 3 * count).reshape(count, 3)
    kn = np.exp(math.log(knmin) + (math.log(knmax) - math.log(knmin)) * mat[:, 0])
    e = emin + (emax - emin) * mat[:, 1]
    mu = mumin + (mumax - mumin) * mat[:, 2]
    er3 = (rmax ** 4 - rmin ** 4) / (4.0 * (rmax - rmin))
    decks = []
    for i in range(count):
        n = int(counts[i])
        sub = seed * 1000003 + i
        rad = uniform(sub, STREAM_RADIUS, n, rmin, rmax).astype(np.float32)
        rad = np.sort(rad)[::-1].copy()
        Lb = float((n * (4.0 / 3.0) * math.pi * er3 / packing) ** (1.0 / 3.0))
        pos = L.fill_random(rad, (0.0, 0.0, 0.0), (Lb, Lb, Lb), seed=sub * 64 + STREAM_POS, gap=0.01)
        vel = uniform(sub, STREAM_VEL, 3 * n, -vmax, vmax).reshape(n, 3)
        # sphere_mass in double, rounded once to float, for every radius at once
        mass = ((4.0 / 3.0) * 2.0 * math.asin(1.0) * rad.astype(np.float64) ** 3 * density * 1e-6).astype(np.float32)
        cell = float(2.0 * rad.max())
        sim = sim_info(n, dt, (0.0, -981.0, 0.0), (-1.0, -1.0, -1.0), (Lb, Lb, Lb), cell, rotation=True,
                       roll_res=0.0)
        pobj = [sphere_pobj(float(rad.max()), density, float(e[i]), float(kn[i]), float(mu[i]))]
        ext = lsd_ext(float(kn[i]), float(e[i]), float(mu[i]))
        ext.per_particle_radius = 1
        decks.append(_deck(sim, [box_walls(Lb)], pobj, pos, vel, ext=ext, radius=rad, mass=mass,
                           name=f"batchbox{i}-{n}"))
        decks[-1]["params"] = dict(kn=float(kn[i]), e=float(e[i]), mu=float(mu[i]), L=Lb)
    return decks


def collapse_deck(nx=6, ny=14, nz=19, L=20.0, seed=4, radius=0.5, density=2500.0, kn=1e5, e=0.3, mu=0.6,
                  roll_res=0.001, dt=1e-5, spacing=1.05, jitter=0.02):
    """Column collapse: a jittered lattice of spheres against the x = 0 wall
    of a closed cube [0, L]^3, released at rest.  It slumps into a ramp that
    falls away from the wall; the ramp's slope is the angle of repose."""
    n = nx * ny * nz
    gi = np.indices((nx, ny, nz)).reshape(3, -1).T.astype(np.float64)
    pos = np.array([radius + 0.05] * 3) + gi * spacing
    pos += uniform(seed, STREAM_JITTER, 3 * n, -jitter, jitter).reshape(n, 3)
    vel = np.zeros((n, 3))
    sim = sim_info(n, dt, (0.0, -981.0, 0.0), (-1.0, -1.0, -1.0), (L, L, L), 2.0 * radius, rotation=True,
                   roll_res=roll_res)
    pobj = [sphere_pobj(radius, density, e, kn, mu)]
    return _deck(sim, [box_walls(L)], pobj, pos, vel, name=f"collapse{n}")


def random_deck(n, seed=7, L=12.0, law="ref", radius=0.5, polydisperse=False, rmin=0.2, rmax=0.4,
                density=2500.0, kn=1e5, e=0.6, mu=0.4, dt=1e-5, vmax=50.0, wmax=20.0,
                cell=None, origin=-1.0, aliased=False):
    """Dense random pack (overlaps allowed) in a closed box, random v and omega:
    exercises every contact and wall branch in one step."""
    pos = uniform(seed, STREAM_POS, 3 * n, 0.6, L - 0.6).reshape(n, 3)
    vel = uniform(seed, STREAM_VEL, 3 * n, -vmax, vmax).reshape(n, 3)
    avel = uniform(seed, STREAM_VEL + 8, 3 * n, -wmax, wmax).reshape(n, 3)
    rmax_ = rmax if polydisperse else radius
    cell = 2.0 * rmax_ if cell is None else cell
    ws = (L, L, L) if not aliased else (L, L, L * 0.5)
    sim = sim_info(n, dt, (0.0, -981.0, 0.0), (origin, origin, origin), ws, cell, rotation=True,
                   roll_res=0.001)
    pobj = [sphere_pobj(radius, density, e, kn, mu)]
    ext = None
    rad = mass = None
    if law != "ref" or polydisperse:
        ext = T.ExtParams() if law == "ref" else lsd_ext(kn, e, mu)
        if law == "hm":
            ext = hm_ext(1e5, 0.3, e, mu)
    if polydisperse:
        ext.per_particle_radius = 1
        rad = uniform(seed, STREAM_RADIUS, n, rmin, rmax).astype(np.float32)
        mass = np.array([sphere_mass(float(r), density) for r in rad], np.float32)
    d = _deck(sim, [box_walls(L)], pobj, pos, vel, ext=ext, radius=rad, mass=mass,
              name=f"random{n}-{law}")
    d["avel"] = avel.astype(np.float32)
    return d


def hm_ext(E, nu, e, mu):
    x = T.ExtParams()
    x.contact_law = T.LAW_HERTZ_MINDLIN
    for t in range(8):
        m = x.material[t]
        m.youngs, m.poisson, m.restitution, m.mu = E, nu, e, mu
    return x


# ---------------------------------------------------------------------------
# silo
# ---------------------------------------------------------------------------
SILO_R = 25.0          # cylinder radius, cm
SILO_H = 120.0         # cylinder height, cm
SILO_OUTLET_R = 3.0    # outlet radius (60 mm outlet), cm
SILO_HALF_ANGLE = 30.0  # hopper half-angle from the axis, degrees
SILO_AXIS_XZ = 26.0    # axis position in x and z
SILO_Y0 = 10.0         # outlet (hatch) height


def silo_geometry():
    hop_h = (SILO_R - SILO_OUTLET_R) / math.tan(math.radians(SILO_HALF_ANGLE))
    return dict(y0=SILO_Y0, y1=SILO_Y0 + hop_h, top=SILO_Y0 + hop_h + SILO_H, hopper_height=hop_h)


def silo_deck(n=1048576, seed=2, rmin=0.2, rmax=0.4, density=2600.0, E=1e5, nu=0.3, e=0.5,
              mu=0.5, dt=5e-6, fill_spacing=None, law="hm", fill="fcc", packing=0.25):
    """Silo discharge deck.  Cylinder R 25 cm (axis y) over a 30-degree conical
    hopper down to a 6 cm outlet closed by a hatch plane (the last world
    object, so bdem_open_hatch removes it like Device_OpenHatch).  Radii
    U[rmin, rmax] from the seed, zero velocity.  Initial state (fill):
    "fcc", an FCC lattice of spacing 2*rmax + gap filling the hopper and the
    cylinder bottom-up; "random", the seeded random non-overlapping fill
    (bdem_fill_random, gap 0.01 cm) of a column of the cylinder above the
    hopper tall enough for a solid fraction `packing`.  Hertz-Mindlin E in
    kg cm^-1 s^-2 (1e7 Pa = 1e5)."""
    g = silo_geometry()
    rad = uniform(seed, STREAM_RADIUS, n, rmin, rmax).astype(np.float32)
    if fill == "random":
        from . import lib as L
        vol = float(np.sum((4.0 / 3.0) * math.pi * rad.astype(np.float64) ** 3))
        H = vol / (packing * math.pi * (SILO_R - rmax) ** 2) * 1.02 + 2.0 * rmax
        lo = (SILO_AXIS_XZ - SILO_R, g["y1"] + 0.01, SILO_AXIS_XZ - SILO_R)
        hi = (SILO_AXIS_XZ + SILO_R, g["y1"] + 0.01 + H, SILO_AXIS_XZ + SILO_R)
        pos = L.fill_random(rad, lo, hi, seed=seed * 64 + STREAM_POS, gap=0.01,
                            cyl=(SILO_AXIS_XZ, SILO_AXIS_XZ, SILO_R - 0.01))
    else:
        d = 2.0 * rmax + 0.01 if fill_spacing is None else fill_spacing
        pos = _fcc_fill(n, d, rmax, g)
    mass = np.array((4.0 / 3.0) * math.pi * rad.astype(np.float64) ** 3 * density * 1e-6, np.float32)
    top = float(pos[:, 1].max()) + 2.0 * rmax
    cell = 2.0 * rmax
    world = (2 * SILO_AXIS_XZ, top, 2 * SILO_AXIS_XZ)
    kill_h = SILO_Y0 - 5.0  # main.cpp:743, hatch height - 5
    sim = sim_info(n, dt, (0.0, -981.0, 0.0), (0.0, 0.0, 0.0), world, cell, rotation=True,
                   roll_res=0.0, kill=True, kill_height=kill_h, sim_type=2)
    cyl = cylinder_wobj(SILO_R, g["top"] - g["y1"], (0.0, 1.0, 0.0),
                        (SILO_AXIS_XZ - SILO_R, 0.0, SILO_AXIS_XZ - SILO_R))
    hatch = infinite_plane_wobj((SILO_AXIS_XZ, SILO_Y0, SILO_AXIS_XZ), (0.0, 1.0, 0.0))
    pobj = [sphere_pobj(rmax, density, e, 1e5, mu)]
    ext = hm_ext(E, nu, e, mu) if law == "hm" else T.ExtParams()
    ext.per_particle_radius = 1
    ext.num_cones = 1
    c = ext.cone[0]
    c.base = T.f3(SILO_AXIS_XZ, 0.0, SILO_AXIS_XZ)
    c.axis = T.f3(0.0, 1.0, 0.0)
    c.s0, c.s1 = SILO_Y0, g["y1"]
    c.r0, c.r1 = SILO_OUTLET_R, SILO_R
    vel = np.zeros((n, 3), np.float32)
    return _deck(sim, [cyl, hatch], pobj, pos, vel, ext=ext, radius=rad, mass=mass,
                 name=f"silo{n}")


def _fcc_fill(n, d, rmax, g):
    """First n sites (bottom-up) of an FCC lattice of nearest-neighbour
    distance d that lie inside the hopper and cylinder with clearance rmax."""
    a = d * math.sqrt(2.0)  # cubic cell edge
    basis = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]]) * a
    cx = cz = SILO_AXIS_XZ
    half = SILO_R
    nxz = int(math.ceil(2 * half / a)) + 1
    layer_h = a
    out = []
    count = 0
    y = SILO_Y0 + rmax + 0.01
    k = 0
    tan = math.tan(math.radians(SILO_HALF_ANGLE))
    cos = math.cos(math.radians(SILO_HALF_ANGLE))
    ii, jj = np.meshgrid(np.arange(nxz), np.arange(nxz), indexing="ij")
    while count < n:
        for b in basis:
            yy = y + k * layer_h + b[1]
            xs = cx - half + ii * a + b[0]
            zs = cz - half + jj * a + b[2]
            rho = np.sqrt((xs - cx) ** 2 + (zs - cz) ** 2)
            if yy < g["y1"]:
                rcone = SILO_OUTLET_R + (yy - SILO_Y0) * tan
                ok = rho <= rcone - rmax / cos - 0.01
            else:
                ok = rho <= SILO_R - rmax - 0.01
            sel = np.stack([xs[ok], np.full(ok.sum(), yy), zs[ok]], axis=1)
            out.append(sel)
            count += len(sel)
        k += 1
    pts = np.concatenate(out)
    pts = pts[np.lexsort((pts[:, 2], pts[:, 0], pts[:, 1]))]
    return pts[:n].astype(np.float32)


# --------------------------------------------------------------------------
# Rotating drum with lifters (ball-mill geometry)
# --------------------------------------------------------------------------
def _quad_face(s, V, order, tol):
    """Centroid, normal and area of a quad face as ReadDObject derives them
    (Input/KSimulationData.cpp:1706-1760): n = normalise(cross(V2, V1)),
    area = |cross(V2,V1)|/2 + |cross(V3,V4)|/2, all in float arithmetic."""
    f = np.float32
    s.Tol = tol
    s.num_vertex = len(order)
    total = np.zeros(3, np.float32)
    for q, o in enumerate(order):
        s.vertex_Order[q] = o
        total = (total + V[o]).astype(np.float32)
    V1 = (V[order[1]] - V[order[0]]).astype(np.float32)
    V2 = (V[order[2]] - V[order[1]]).astype(np.float32)
    V3 = (V[order[3]] - V[order[2]]).astype(np.float32)
    V4 = (V[order[0]] - V[order[3]]).astype(np.float32)
    n = np.cross(V2, V1).astype(np.float32)
    ln = f(np.sqrt(f(np.dot(n, n))))
    n2 = np.cross(V3, V4).astype(np.float32)
    s.area = float(f(f(f(0.5) * ln) + f(f(0.5) * f(np.sqrt(f(np.dot(n2, n2)))))))
    s.normal = T.f3(*(n / ln).astype(np.float32))
    s.centroid = T.f3(*((f(1.0) / f(len(order))) * total).astype(np.float32))


def lifter_dobj(verts, com, bound_r, tol=0.1, attached=True):
    """A rectangular lifter with the layout of Projects/3.BallMill/LifterObjects/
    Square_Lifter16.DOBJ: 8 vertices (base y0 at the drum wall, top y1), the
    two side faces and the top face, and the two top edges along the axis."""
    d = T.DOBJ()
    d.btype = 1
    d.Axis = T.f3(1.0, 1.0, 0.0)
    d.COM = T.f3(*com)
    d.boundR = float(bound_r)
    V = np.asarray(verts, np.float32)
    d.num_vertex = 8
    for k in range(8):
        d.vertex[k] = T.f3(*V[k])
    d.num_faces = 3
    for j, order in enumerate([(0, 4, 5, 1), (2, 6, 7, 3), (6, 5, 4, 7)]):
        _quad_face(d.faces[j], V, order, tol)
    d.num_edges = 2
    for k, (a, b) in enumerate([(4, 5), (6, 7)]):
        d.edge[k].point_vindex[0] = a
        d.edge[k].point_vindex[1] = b
    d.is_attached_DWorld = bool(attached)
    d.is_translating = False
    return d


def rotate_dobj(d, about, angle):
    """Rotate a lifter about the z axis through `about` and re-derive its faces,
    as Rotate_VolObject_Inc does each step (Input/Processing_Mills.h:7-83):
    x' = cos*x + sin*y, y' = -sin*x + cos*y, in float arithmetic."""
    f = np.float32
    c, s = f(math.cos(angle)), f(math.sin(angle))
    ab = np.asarray(about, np.float32)
    V = np.array([[d.vertex[k].x, d.vertex[k].y, d.vertex[k].z] for k in range(d.num_vertex)], np.float32)
    out = T.DOBJ.from_buffer_copy(d)
    for k in range(d.num_vertex):
        p = (V[k] - ab).astype(np.float32)
        q = np.array([f(f(c * p[0]) + f(s * p[1])), f(f(-s * p[0]) + f(c * p[1])), p[2]], np.float32)
        V[k] = (q + ab).astype(np.float32)
        out.vertex[k] = T.f3(*V[k])
    for j in range(d.num_faces):
        order = [d.faces[j].vertex_Order[q] for q in range(d.faces[j].num_vertex)]
        _quad_face(out.faces[j], V, order, d.faces[j].Tol)
    p = (np.array([d.COM.x, d.COM.y, d.COM.z], np.float32) - ab).astype(np.float32)
    q = np.array([f(f(c * p[0]) + f(s * p[1])), f(f(-s * p[0]) + f(c * p[1])), p[2]], np.float32)
    out.COM = T.f3(*(q + ab).astype(np.float32))
    return out


def mill_deck(n=3000, seed=3, R=15.0, L=10.0, n_lifters=8, lifter_w=1.5, lifter_h=1.5, rpm=30.0,
              law="ref", radius=0.5, polydisperse=False, rmin=0.4, rmax=0.6, density=2700.0,
              kn=2e5, e=0.4, mu=0.6, dt=1e-5, vmax=20.0):
    """Drum of radius R and length L (axis z, centre (R, R)) with n_lifters
    square lifters attached to the wall, rotating at `rpm`; spheres on a
    jittered lattice in the lower half of the drum.  Lifter l is the bottom
    lifter rotated by 2 pi l / n_lifters about the axis (Repeat_Lifters,
    Input/KSimulationData.cpp:1962-2028)."""
    f = np.float32
    rmax_ = rmax if polydisperse else radius
    cell = 2.0 * rmax_
    x0, x1 = R - lifter_w / 2, R + lifter_w / 2
    base = lifter_dobj([(x0, 0, 0), (x0, 0, L), (x1, 0, L), (x1, 0, 0),
                        (x0, lifter_h, 0), (x0, lifter_h, L), (x1, lifter_h, L), (x1, lifter_h, 0)],
                       (R, lifter_h / 2, L / 2), max(lifter_w, lifter_h))
    about = (R, R, 0.0)
    dobj = [base] + [rotate_dobj(base, about, 2.0 * math.pi * k / n_lifters) for k in range(1, n_lifters)]
    # spheres: lattice inside the drum below its centre, clear of the lifters
    d = 2.0 * rmax_ * 1.05
    pts = []
    inner = R - lifter_h - rmax_ * 1.5
    g = np.arange(R - inner + rmax_, R + inner, d)
    zs = np.arange(rmax_ * 1.2, L - rmax_ * 1.2, d)
    for y in np.arange(R - inner + rmax_, R + 0.8 * inner, d):
        for x in g:
            if (x - R) ** 2 + (y - R) ** 2 > (inner - rmax_) ** 2:
                continue
            for z in zs:
                pts.append((x, y, z))
                if len(pts) == n:
                    break
            if len(pts) == n:
                break
        if len(pts) == n:
            break
    if len(pts) < n:
        raise ValueError(f"drum holds only {len(pts)} of {n} spheres on this lattice")
    pos = np.asarray(pts, np.float64)
    pos += uniform(seed, STREAM_POS, 3 * n, -0.02, 0.02).reshape(n, 3)
    vel = uniform(seed, STREAM_VEL, 3 * n, -vmax, vmax).reshape(n, 3)
    vel[:, 2] = 0.0
    sim = sim_info(n, dt, (0.0, -981.0, 0.0), (-1.0, -1.0, -1.0), (2 * R + 1, 2 * R + 1, L + 1), cell,
                   rotation=True, roll_res=0.0, sim_type=1)
    sim.isMillSim = True
    sim.Rotating_WObject = False
    omega = rpm * 2.0 * math.pi / 60.0
    sim.MillRotVel = float(f(-omega))
    sim.DObject_CylinderBoundR = float(f(f(R) - f(2.0) * f(base.COM.y)))
    sim.cent = T.f3(R, R, 0.0)
    pobj = [sphere_pobj(radius, density, e, kn, mu)]
    ext = None
    rad = mass = None
    if law != "ref" or polydisperse:
        ext = T.ExtParams() if law == "ref" else lsd_ext(kn, e, mu)
        if law == "hm":
            ext = hm_ext(1e5, 0.3, e, mu)
    if polydisperse:
        ext.per_particle_radius = 1
        rad = uniform(seed, STREAM_RADIUS, n, rmin, rmax).astype(np.float32)
        mass = np.array([sphere_mass(float(r), density) for r in rad], np.float32)
    drum = cylinder_wobj(R, L, (0.0, 0.0, 1.0), (0.0, 0.0, 0.0), top=True, bot=True)
    deck = _deck(sim, [drum], pobj, pos, vel, ext=ext, radius=rad, mass=mass, name=f"mill{n}-{law}")
    deck["dobj"] = dobj
    deck["lifter_about"] = about
    deck["rad_per_step"] = -omega * dt
    return deck


def drum_deck(n=20000, seed=5, R=15.0, L=10.0, rpm=30.0, rmin=0.15, rmax=0.25, density=2500.0, kn=1e5,
              e=0.5, mu=0.4, roll_res=0.001, dt=1e-5, vmax=0.0, spacing=1.05):
    """Smooth rotating drum (radius R, length L, axis z, centre (R, R), end
    plates, no lifters) at `rpm`, reference pair law with per-particle radii
    U[rmin, rmax] on a jittered lattice in the lower part of the drum: the
    calibration workload.  The drum's wall constants come from (kn, e, mu)
    and stay fixed; the particle pair constants and rolling resistance are
    the calibrated parameters.  The lattice pitch is spacing x 2 rmax
    (below 1: neighbours start in contact)."""
    f = np.float32
    cell = 2.0 * rmax
    d = 2.0 * rmax * spacing
    pts = []
    inner = R - rmax * 1.5
    xs = np.arange(R - inner + rmax, R + inner, d)
    zs = np.arange(rmax * 1.2, L - rmax * 1.2, d)
    for y in np.arange(R - inner + rmax, R + inner, d):
        for x in xs:
            if (x - R) ** 2 + (y - R) ** 2 > (inner - rmax) ** 2:
                continue
            for z in zs:
                pts.append((x, y, z))
                if len(pts) == n:
                    break
            if len(pts) == n:
                break
        if len(pts) == n:
            break
    if len(pts) < n:
        raise ValueError(f"drum holds only {len(pts)} of {n} spheres on this lattice")
    pos = np.asarray(pts, np.float64) + uniform(seed, STREAM_JITTER, 3 * n, -0.01, 0.01).reshape(n, 3)
    vel = uniform(seed, STREAM_VEL, 3 * n, -vmax, vmax).reshape(n, 3) if vmax > 0 else np.zeros((n, 3))
    rad = uniform(seed, STREAM_RADIUS, n, rmin, rmax).astype(np.float32)
    mass = ((4.0 / 3.0) * 2.0 * math.asin(1.0) * rad.astype(np.float64) ** 3 * density * 1e-6).astype(np.float32)
    sim = sim_info(n, dt, (0.0, -981.0, 0.0), (-1.0, -1.0, -1.0), (2 * R + 1, 2 * R + 1, L + 1), cell,
                   rotation=True, roll_res=roll_res, sim_type=1)
    sim.isMillSim = True
    sim.Rotating_WObject = False
    omega = rpm * 2.0 * math.pi / 60.0
    sim.MillRotVel = float(f(-omega))
    sim.cent = T.f3(R, R, 0.0)
    pobj = [sphere_pobj(0.5 * (rmin + rmax), density, e, kn, mu)]
    ext = T.ExtParams()
    ext.per_particle_radius = 1
    drum = cylinder_wobj(R, L, (0.0, 0.0, 1.0), (0.0, 0.0, 0.0), top=True, bot=True)
    deck = _deck(sim, [drum], pobj, pos, vel, ext=ext, radius=rad, mass=mass, name=f"drum{n}")
    deck["theta"] = (mu, e, kn, roll_res)
    return deck


def place_at_lifters(deck, seed=5, overlap=0.05):
    """Move the first spheres of a mill deck onto every lifter: one on each
    face and one at each top edge, overlapping by `overlap`, at seeded
    positions along the lifter.  Exercises face and edge contacts from the
    first step."""
    pos = deck["pos"].copy()
    rad = deck["radius"] if deck.get("radius") is not None else np.full(len(pos), deck["pobj"][0].radius, np.float32)
    k = 0
    u = uniform(seed, STREAM_JITTER, 64 * len(deck["dobj"]), 0.2, 0.8)
    for D in deck["dobj"]:
        V = np.array([[D.vertex[q].x, D.vertex[q].y, D.vertex[q].z] for q in range(D.num_vertex)])
        for j in range(D.num_faces):
            F = D.faces[j]
            n = np.array([F.normal.x, F.normal.y, F.normal.z])
            c = np.array([F.centroid.x, F.centroid.y, F.centroid.z])
            c[2] = V[:, 2].min() + u[k] * (V[:, 2].max() - V[:, 2].min())
            pos[k] = c + n * (rad[k] - overlap)
            k += 1
        ntop = np.array([D.faces[2].normal.x, D.faces[2].normal.y, D.faces[2].normal.z])
        for e, side in zip(range(D.num_edges), (0, 1)):
            a, b = D.edge[e].point_vindex[0], D.edge[e].point_vindex[1]
            ns = np.array([D.faces[side].normal.x, D.faces[side].normal.y, D.faces[side].normal.z])
            dirn = (ntop + ns) / np.linalg.norm(ntop + ns)
            p = V[a] + u[k] * (V[b] - V[a])
            pos[k] = p + dirn * (rad[k] - overlap)
            k += 1
    out = dict(deck)
    out["pos"] = pos.astype(np.float32)
    return out


# --------------------------------------------------------------------------
# Polyhedra (particle_type 1)
# --------------------------------------------------------------------------
CUBE_VERTS = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1],
                       [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]], np.float64) * 0.5
# outward counter-clockwise quads (0-based)
CUBE_FACES = [[0, 3, 2, 1], [4, 5, 6, 7], [0, 1, 5, 4], [2, 3, 7, 6], [1, 2, 6, 5], [0, 4, 7, 3]]


def _edges_of(faces):
    seen, out = set(), []
    for f in faces:
        for q in range(len(f)):
            a, b = f[q], f[(q + 1) % len(f)]
            key = (min(a, b), max(a, b))
            if key not in seen:
                seen.add(key)
                out.append((a, b))
    return out


def _tet_inertia(a, b, c):
    """Inertia tensor (about the origin, unit density) and volume of the
    tetrahedron (0, a, b, c), from the covariance of a simplex."""
    vol = np.dot(a, np.cross(b, c)) / 6.0
    S = np.zeros((3, 3))
    P = [a, b, c]
    for x in P:
        S += np.outer(x, x)
    s = a + b + c
    S += np.outer(s, s)
    cov = vol / 20.0 * S
    return vol, np.trace(cov) * np.eye(3) - cov


def poly_pobj(verts, faces, density, cor, kn, mu, cor_s=None, kn_s=None, mu_s=None, tol=0.1):
    """A polyhedron POBJ as ReadPObject_Poly + Make_Particle_LocalInfo build
    it (Input/KSimulationData.cpp:1180-1530): COM = vertex mean, Local_Vertex
    = vertex - COM, bounding radius = largest |Local_Vertex|, face centroid =
    vertex mean of the face, face normal = normalise(cross(v1-v2, v2-v0)) for
    triangles and normalise(cross(v2-v3, v3-v0)) for polygons, area = the
    centroid fan, all in float arithmetic.  `faces` are 0-based outward CCW
    vertex orders.  Mass = density * volume * 1e-6 and the stored tensor is
    the inverse of the inertia tensor about the COM (kg cm^2); the reference
    reads a normalised tensor from its deck, so the tensor here is input data.
    Triangles get their own area (the reference's reader leaves the previous
    face's value there)."""
    f = np.float32
    V = np.asarray(verts, np.float32)
    nv = len(V)
    assert nv <= 32 and len(faces) <= 32
    p = T.POBJ()
    p.num_vertex = nv
    total = np.zeros(3, np.float32)
    for k in range(nv):
        p.vertex[k] = T.f3(*V[k])
        total = (total + V[k]).astype(np.float32)
    com = (total * f(f(1.0) / f(nv))).astype(np.float32)
    p.COM = T.f3(*com)
    loc = (V - com).astype(np.float32)
    rad = f(0.0)
    for k in range(nv):
        p.Local_Vertex[k] = T.f3(*loc[k])
        ln = f(np.sqrt(f(np.dot(loc[k], loc[k]))))
        rad = max(rad, ln)
    p.radius = float(rad)
    p.num_faces = len(faces)
    vol = 0.0
    Itot = np.zeros((3, 3))
    cd = V.astype(np.float64) - com.astype(np.float64)
    for j, order in enumerate(faces):
        s = p.face[j]
        s.num_vertex = len(order)
        s.Tol = tol
        tot = np.zeros(3, np.float32)
        for q, o in enumerate(order):
            s.vertex_Order[q] = o
            tot = (tot + V[o]).astype(np.float32)
        c = ((f(1.0) / f(len(order))) * tot).astype(np.float32)
        s.centroid = T.f3(*c)
        if len(order) == 3:
            n = np.cross((V[order[1]] - V[order[2]]).astype(np.float32),
                         (V[order[2]] - V[order[0]]).astype(np.float32)).astype(np.float32)
        else:
            n = np.cross((V[order[2]] - V[order[3]]).astype(np.float32),
                         (V[order[3]] - V[order[0]]).astype(np.float32)).astype(np.float32)
        ln = f(np.sqrt(f(np.dot(n, n))))
        s.normal = T.f3(*(n / ln).astype(np.float32))
        area = f(0.0)
        for q in range(len(order)):
            a = (V[order[q]] - c).astype(np.float32)
            b = (V[order[(q + 1) % len(order)]] - c).astype(np.float32)
            x = np.cross(a, b).astype(np.float32)
            area = f(area + f(f(0.5) * f(np.sqrt(f(np.dot(x, x))))))
        s.area = float(area)
        cc = c.astype(np.float64) - com.astype(np.float64)
        for q in range(len(order)):
            dv, Iv = _tet_inertia(cd[order[q]], cd[order[(q + 1) % len(order)]], cc)
            vol += dv
            Itot += Iv
    edges = _edges_of(faces)
    assert len(edges) <= 32
    p.num_edges = len(edges)
    for k, (a, b) in enumerate(edges):
        p.edge[k].point_vindex[0] = a
        p.edge[k].point_vindex[1] = b
    p.mass = float(f(density * vol * 1e-6))
    Iinv = np.linalg.inv(Itot * density * 1e-6)
    for g in range(9):
        p.InertiaT[g] = float(f(Iinv.flat[g]))
    m = float(f(p.mass))
    kn_s = kn if kn_s is None else kn_s
    cor_s = cor if cor_s is None else cor_s
    mu_s = mu if mu_s is None else mu_s
    p.PP[0].Kn = kn
    p.PP[0].Cn = cn_from_cor(float(f(kn)), 1.0 / (1.0 / m + 1.0 / m), cor)
    p.PP[0].Fric_static = mu
    p.PP[0].Fric_kinetic = mu
    p.surface_Kn = kn_s
    p.surface_Cn = cn_from_cor(float(f(kn_s)), m, cor_s)
    p.surface_Fric_static = mu_s
    p.surface_Fric_kinetic = mu_s
    p.Dsurface_Kn = kn_s
    p.Dsurface_Cn = p.surface_Cn
    p.Dsurface_Fric_static = mu_s
    p.Dsurface_Fric_kinetic = mu_s
    return p


def cross_pair_params(pobj, kn, e, mu):
    """PP[tb] of every type for every partner type, as Set_SpringConstants_KN
    derives them (Input/KSimulationData.cpp:950-1045): reduced mass m/2 for
    the same type, mA mB/(mA+mB) across types."""
    for a, pa in enumerate(pobj):
        for b, pb in enumerate(pobj):
            ma, mb = float(np.float32(pa.mass)), float(np.float32(pb.mass))
            pa.PP[b].Kn = kn
            pa.PP[b].Cn = cn_from_cor(float(np.float32(kn)), 1.0 / (1.0 / ma + 1.0 / mb), e)
            pa.PP[b].Fric_static = mu
            pa.PP[b].Fric_kinetic = mu


def hull_shape(seed, nv=12, aspect=(1.0, 0.8, 0.6), size=1.0):
    """A convex polyhedron: the hull of nv seeded points on an ellipsoid,
    with outward CCW triangles (nv <= 12 keeps E = 3nv-6 <= 32)."""
    from scipy.spatial import ConvexHull
    u = uniform(seed, STREAM_SHAPE, 2 * nv, 0.0, 1.0).reshape(nv, 2)
    th = np.arccos(1.0 - 2.0 * u[:, 0])
    ph = 2.0 * math.pi * u[:, 1]
    P = np.stack([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)], 1)
    P = P * np.asarray(aspect) * (0.5 * size)
    h = ConvexHull(P)
    keep = sorted(set(h.vertices.tolist()))
    remap = {v: k for k, v in enumerate(keep)}
    V = P[keep]
    faces = []
    for simp, eq in zip(h.simplices, h.equations):
        a, b, c = (P[s] for s in simp)
        if np.dot(np.cross(b - a, c - a), eq[:3]) < 0:
            simp = simp[[0, 2, 1]]
        faces.append([remap[s] for s in simp])
    return V, faces


def random_quats(seed, n):
    """Uniform random unit quaternions (w, x, y, z) (Shoemake's method)."""
    u = uniform(seed, STREAM_SHAPE + 16, 3 * n, 0.0, 1.0).reshape(n, 3).astype(np.float64)
    a, b = np.sqrt(1 - u[:, 0]), np.sqrt(u[:, 0])
    t1, t2 = 2 * math.pi * u[:, 1], 2 * math.pi * u[:, 2]
    q = np.stack([b * np.cos(t2), a * np.sin(t1), a * np.cos(t1), b * np.sin(t2)], 1)
    q = q / np.linalg.norm(q, axis=1, keepdims=True)
    return q.astype(np.float32)


def axis_quat(axis, deg):
    a = np.asarray(axis, np.float64)
    a = a / np.linalg.norm(a)
    h = math.radians(deg) * 0.5
    return np.array([math.cos(h), *(math.sin(h) * a)], np.float32)


def tilt_quats(seed, n, max_deg):
    """Rotations by a seeded angle in [0, max_deg] about seeded axes."""
    u = uniform(seed, STREAM_SHAPE + 24, 4 * n, 0.0, 1.0).reshape(n, 4).astype(np.float64)
    ax = np.stack([np.sqrt(1 - (2 * u[:, 0] - 1) ** 2) * np.cos(2 * math.pi * u[:, 1]),
                   np.sqrt(1 - (2 * u[:, 0] - 1) ** 2) * np.sin(2 * math.pi * u[:, 1]),
                   2 * u[:, 0] - 1], 1)
    h = np.radians(max_deg) * u[:, 2] * 0.5
    return np.concatenate([np.cos(h)[:, None], np.sin(h)[:, None] * ax], 1).astype(np.float32)


def poly_deck(n, seed=31, L=12.0, shape="cube", size=1.0, density=2500.0, kn=1e5, e=0.6, mu=0.4,
              dt=1e-5, vmax=20.0, wmax=10.0, spacing=None, overlap=0.04, rotation=True, tilt=None,
              jitter=0.05):
    """Polyhedra (particle_type 1) on a jittered lattice in a closed box,
    random orientations, velocities and spins: neighbours' bounding spheres
    overlap, so face, edge and vertex contacts of every kind occur.
    shape = "cube" (8 V, 6 F, 12 E), "hull" (12-vertex convex hull, 20 F,
    30 E), "mixed" (cube and hull) or "cubes2" (cubes of side 1 and 1.4);
    with two types they alternate as a 3-D checkerboard.  tilt = None draws uniform
    orientations; a number keeps every body within that many degrees of the
    lattice axes, so that faces meet faces."""
    if shape == "cube":
        shapes = [(CUBE_VERTS * size, CUBE_FACES)]
    elif shape == "hull":
        shapes = [hull_shape(seed, size=size)]
    elif shape == "cubes2":
        shapes = [(CUBE_VERTS * size, CUBE_FACES), (CUBE_VERTS * (1.4 * size), CUBE_FACES)]
    else:
        shapes = [(CUBE_VERTS * size, CUBE_FACES), hull_shape(seed, size=size)]
    pobj = [poly_pobj(V, F, density, e, kn, mu) for V, F in shapes]
    cross_pair_params(pobj, kn, e, mu)
    rmax = max(p.radius for p in pobj)
    d = spacing if spacing is not None else 2.0 * rmax * (1.0 - overlap) * 0.6
    m = int(math.floor((L - 2 * rmax - 0.2) / d)) + 1
    assert m ** 3 >= n, "box too small for the lattice"
    gi = np.indices((m, m, m)).reshape(3, -1).T[:n].astype(np.float64)
    pos = rmax + 0.1 + gi * d
    pos += uniform(seed, STREAM_JITTER, 3 * n, -jitter * d, jitter * d).reshape(n, 3)
    vel = uniform(seed, STREAM_VEL, 3 * n, -vmax, vmax).reshape(n, 3)
    avel = uniform(seed, STREAM_VEL + 8, 3 * n, -wmax, wmax).reshape(n, 3)
    cell = 2.0 * rmax
    sim = sim_info(n, dt, (0.0, -981.0, 0.0), (-1.0, -1.0, -1.0), (L, L, L), cell,
                   rotation=rotation, roll_res=0.0)
    sim.particle_type = 1
    sim.Num_ParticleObjects = len(pobj)
    ptype = (gi.sum(axis=1).astype(np.int64) % len(pobj)).astype(np.uint32)
    out = _deck(sim, [box_walls(L)], pobj, pos, vel, ptype=ptype, name=f"poly{n}-{shape}")
    out["avel"] = avel.astype(np.float32)
    out["quat"] = random_quats(seed, n) if tilt is None else tilt_quats(seed, n, tilt)
    return out


def poly_pair_deck(pa, pb, qa, qb, shape_a=(CUBE_VERTS, CUBE_FACES), shape_b=None, va=(0, 0, 0),
                   vb=(0, 0, 0), wa=(0, 0, 0), wb=(0, 0, 0), kn=1e5, e=0.6, mu=0.4, L=12.0, gravity=True):
    """Two polyhedra in an otherwise empty box (type 0 = A's shape, type 1 =
    B's shape when given), for closed-form contact checks."""
    shapes = [shape_a] if shape_b is None else [shape_a, shape_b]
    pobj = [poly_pobj(V, F, 2500.0, e, kn, mu) for V, F in shapes]
    cross_pair_params(pobj, kn, e, mu)
    rmax = max(p.radius for p in pobj)
    sim = sim_info(2, 1e-5, (0.0, -981.0 if gravity else 0.0, 0.0), (-1.0, -1.0, -1.0), (L, L, L),
                   2.0 * rmax, rotation=True, roll_res=0.0)
    sim.particle_type = 1
    sim.Num_ParticleObjects = len(pobj)
    ptype = np.array([0, 0 if shape_b is None else 1], np.uint32)
    d = _deck(sim, [box_walls(L)], pobj, np.array([pa, pb], np.float64), np.array([va, vb], np.float64),
              ptype=ptype, name="poly-pair")
    d["avel"] = np.array([wa, wb], np.float32)
    d["quat"] = np.array([qa, qb], np.float32)
    return d


def hull_points(seed, n, kmin=8, kmax=16, rmin=0.15, rmax=0.30):
    """Point sets for per-particle hulls: k_p ~ U{kmin..kmax} points uniform
    on the unit sphere (seeded), scaled so that the largest distance from
    their mean (the body's COM as the reference's reader takes it) is
    r_p ~ U[rmin, rmax].  Returns (pts [n, kmax, 3] float32, k [n] int32,
    r [n])."""
    k = (kmin + np.floor(uniform(seed, STREAM_SHAPE, n, 0.0, 1.0) * (kmax - kmin + 1))).astype(np.int32)
    k = np.minimum(k, kmax)
    u = uniform(seed, STREAM_SHAPE + 32, 2 * n * kmax, 0.0, 1.0).reshape(n, kmax, 2).astype(np.float64)
    z = 1.0 - 2.0 * u[..., 0]
    ph = 2.0 * math.pi * u[..., 1]
    rho = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    P = np.stack([rho * np.cos(ph), rho * np.sin(ph), z], -1)
    live = np.arange(kmax)[None, :] < k[:, None]
    P = np.where(live[..., None], P, 0.0)
    mean = P.sum(1) / k[:, None]
    d = np.where(live, np.linalg.norm(P - mean[:, None, :], axis=-1), 0.0).max(1)
    r = uniform(seed, STREAM_RADIUS, n, rmin, rmax).astype(np.float64)
    P = np.where(live[..., None], P * (r / d)[:, None, None], 0.0)
    return P.astype(np.float32), k, r


def shape_from_pobj(p):
    """The geometry of a POBJ row as a PolyShape row (what the library makes
    of a per-type table)."""
    s = T.PolyShape()
    s.radius, s.mass, s.COM = p.radius, p.mass, p.COM
    s.num_vertex, s.num_edges, s.num_faces = p.num_vertex, p.num_edges, p.num_faces
    for k in range(32):
        s.Local_Vertex[k] = p.Local_Vertex[k]
        s.edge[k] = p.edge[k]
        s.face[k] = p.face[k]
    for g in range(9):
        s.InertiaT[g] = p.InertiaT[g]
    return s


def _poly_type_for_shapes(shapes, density, kn, e, mu):
    """Type 0 of a per-particle-shape deck: the constants every pair and wall
    contact uses (POBJ::PP and the surface constants), damping from the
    restitution at the mean particle mass (the reference keeps one constant
    per type pair, Set_SpringConstants_KN, Input/KSimulationData.cpp:950-1045).
    Its own geometry (a cube of the mean bounding radius) is not read once
    the shape table is set."""
    rbar = float(np.mean([sh.radius for sh in shapes]))
    p = poly_pobj(CUBE_VERTS * (rbar / math.sqrt(3.0)), CUBE_FACES, density, e, kn, mu)
    m = float(np.float32(np.mean([sh.mass for sh in shapes])))
    p.PP[0].Cn = cn_from_cor(float(np.float32(kn)), 0.5 * m, e)
    p.surface_Cn = cn_from_cor(float(np.float32(kn)), m, e)
    p.Dsurface_Cn = p.surface_Cn
    return p


def hull_mill_deck(n=200000, seed=21, R=30.0, L=20.0, n_lifters=8, lifter_w=2.0, lifter_h=3.0,
                   speed=0.6, rmin=0.3, rmax=0.6, kmin=8, kmax=16, density=2700.0, kn=2e5, e=0.4, mu=0.6,
                   dt=2e-6, gap=0.02, vmax=0.0, fill_scale=0.5):
    """Per-particle convex hulls in a lifter drum: n hulls of kmin..kmax
    seeded points on a sphere (hull_points, bounding radius U[rmin, rmax]
    about the COM), each its own row of the shape table (P_ID p -> row p),
    in a drum of radius R and length L (axis z, centre (R, R), end plates)
    with n_lifters rectangular lifters (lifter_w wide, lifter_h high)
    turning at `speed` x the critical speed sqrt(g / R).  Start: random
    non-overlapping bounding spheres (bdem_fill_random, largest first, gap
    apart) filling the drum inside the lifter tips, random orientations;
    the spheres placed have fill_scale x the bounding radius.  200,000
    bounding spheres of U[0.3, 0.6] cm would need 1.9 times the drum's
    volume, so the default places half-radius spheres (23 % of the space
    inside the lifter tips): the hulls then start in contact with about 13
    neighbours each and the charge spreads from there."""
    from .lib import fill_random, poly_shapes_from_points
    f = np.float32
    pts, k, _ = hull_points(seed, n, kmin, kmax, rmin, rmax)
    shapes = poly_shapes_from_points((pts, k), density)
    rad = np.array([sh.radius for sh in shapes], np.float32)
    x0, x1 = R - lifter_w / 2, R + lifter_w / 2
    base = lifter_dobj([(x0, 0, 0), (x0, 0, L), (x1, 0, L), (x1, 0, 0),
                        (x0, lifter_h, 0), (x0, lifter_h, L), (x1, lifter_h, L), (x1, lifter_h, 0)],
                       (R, lifter_h / 2, L / 2), max(lifter_w, lifter_h))
    about = (R, R, 0.0)
    dobj = [base] + [rotate_dobj(base, about, 2.0 * math.pi * q / n_lifters) for q in range(1, n_lifters)]
    # fill in a frame whose vertical axis is the drum axis: (X, Y, Z) = (x, z, y)
    inner = R - lifter_h - 0.05
    order = np.argsort(-rad, kind="stable")
    placed = fill_random(rad[order] * np.float32(fill_scale), (R - inner, 0.05, R - inner), (R + inner, L - 0.05, R + inner), seed,
                         gap=gap, cyl=(R, R, inner), max_tries=1000000)
    pos = np.zeros((n, 3), np.float32)
    pos[order] = placed[:, [0, 2, 1]]
    vel = uniform(seed, STREAM_VEL, 3 * n, -vmax, vmax).reshape(n, 3) if vmax > 0 else np.zeros((n, 3))
    cell = 2.0 * float(rad.max())
    sim = sim_info(n, dt, (0.0, -981.0, 0.0), (-1.0, -1.0, -1.0), (2 * R + 1, 2 * R + 1, L + 1), cell,
                   rotation=True, roll_res=0.0, sim_type=1)
    sim.particle_type = 1
    sim.Num_ParticleObjects = 1
    sim.isMillSim = True
    sim.Rotating_WObject = False
    omega = speed * math.sqrt(981.0 / R)
    sim.MillRotVel = float(f(-omega))
    sim.DObject_CylinderBoundR = float(f(f(R) - f(2.0) * f(base.COM.y)))
    sim.cent = T.f3(R, R, 0.0)
    pobj = [_poly_type_for_shapes(shapes, density, kn, e, mu)]
    drum = cylinder_wobj(R, L, (0.0, 0.0, 1.0), (0.0, 0.0, 0.0), top=True, bot=True)
    deck = _deck(sim, [drum], pobj, pos, vel, name=f"hullmill{n}")
    deck["shapes"] = shapes
    deck["dobj"] = dobj
    deck["lifter_about"] = about
    deck["rad_per_step"] = -omega * dt
    deck["rpm"] = omega * 60.0 / (2.0 * math.pi)
    deck["quat"] = random_quats(seed, n)
    deck["avel"] = np.zeros((n, 3), np.float32)
    return deck


def hull_box_deck(n=600, seed=41, L=14.0, rmin=0.5, rmax=0.9, kmin=8, kmax=16, density=2500.0, kn=1e5,
                  e=0.6, mu=0.4, dt=1e-5, vmax=20.0, wmax=10.0, overlap=0.25, jitter=0.05):
    """Per-particle hulls (hull_points) on a jittered lattice in a closed box
    with random orientations, velocities and spins; the lattice pitch is
    2 rmax (1 - overlap), so bounding spheres overlap and vertex, edge and
    face contacts of every kind occur from the first step."""
    from .lib import poly_shapes_from_points
    pts, k, _ = hull_points(seed, n, kmin, kmax, rmin, rmax)
    shapes = poly_shapes_from_points((pts, k), density)
    rb = max(sh.radius for sh in shapes)
    d = 2.0 * rb * (1.0 - overlap)
    m = int(math.floor((L - 2 * rb - 0.2) / d)) + 1
    assert m ** 3 >= n, "box too small for the lattice"
    gi = np.indices((m, m, m)).reshape(3, -1).T[:n].astype(np.float64)
    pos = rb + 0.1 + gi * d
    pos += uniform(seed, STREAM_JITTER, 3 * n, -jitter * d, jitter * d).reshape(n, 3)
    vel = uniform(seed, STREAM_VEL, 3 * n, -vmax, vmax).reshape(n, 3)
    avel = uniform(seed, STREAM_VEL + 8, 3 * n, -wmax, wmax).reshape(n, 3)
    sim = sim_info(n, dt, (0.0, -981.0, 0.0), (-1.0, -1.0, -1.0), (L, L, L), 2.0 * rb, rotation=True,
                   roll_res=0.0)
    sim.particle_type = 1
    sim.Num_ParticleObjects = 1
    pobj = [_poly_type_for_shapes(shapes, density, kn, e, mu)]
    out = _deck(sim, [box_walls(L)], pobj, pos, vel, name=f"hullbox{n}")
    out["shapes"] = shapes
    out["avel"] = avel.astype(np.float32)
    out["quat"] = random_quats(seed, n)
    return out


def poly_mill_deck(n=300, seed=11, size=1.0, gap=0.6, R=15.0, L=10.0, n_lifters=8, lifter_w=1.5,
                   lifter_h=1.5, rpm=30.0, density=2700.0, kn=2e5, e=0.4, mu=0.6, dt=1e-5, vmax=10.0):
    """Cubes of side `size` in the drum of mill_deck (8 attached lifters),
    randomly oriented; the first ones sit against every lifter face and top
    edge with their centre `gap`*size from it, so a vertex, an edge or a face
    of the cube is inside the lifter (a randomly turned cube reaches 0.5 to
    0.87 of its side along any direction)."""
    rb = 0.5 * math.sqrt(3.0) * size
    deck = mill_deck(n=n, seed=seed, R=R, L=L, n_lifters=n_lifters, lifter_w=lifter_w, lifter_h=lifter_h,
                     rpm=rpm, radius=rb, density=density, kn=kn, e=e, mu=mu, dt=dt, vmax=vmax)
    pobj = [poly_pobj(CUBE_VERTS * size, CUBE_FACES, density, e, kn, mu)]
    cross_pair_params(pobj, kn, e, mu)
    deck["pobj"] = pobj
    deck["sim"].particle_type = 1
    deck["sim"].Num_ParticleObjects = 1
    deck["quat"] = random_quats(seed, n)
    deck["avel"] = np.zeros((n, 3), np.float32)
    placed = place_at_lifters(dict(deck, radius=np.full(n, gap * size, np.float32)), seed=seed + 1, overlap=0.0)
    deck["pos"] = placed["pos"]
    deck["name"] = f"polymill{n}"
    return deck
