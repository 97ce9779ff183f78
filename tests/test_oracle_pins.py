"""Pins of the oracle's wall responses and history laws (CPU).

Each check compares the restatement with something that does not share its
code: a brute-force nearest-point search on the cone surface, the textbook
Hertz-Mindlin / Cundall-Strack forms (Hertz 1882, Mindlin-Deresiewicz 1953,
Tsuji et al. 1992), equilibria of a sphere at rest on a wall, the sliding
block (deceleration mu g), frame invariance of the moving drum wall, and the
exact restitution of an undamped-clamp linear oscillator.  A sign, factor-2
or wrong-radius mistake in any of these functions fails at least one test.

Reference lines restated by the functions pinned here:
  static wall   Collision_Response.cuh:4-109
  moving wall   Collision_Response.cuh:116-233, DEM_Kernels_Compute.cuh:449-452
  end plates    Collision_Response.cuh:504-526 (and the overwrite of
                DEM_Kernels_Compute.cuh:539, 585)
  cone, history laws: extensions with no reference counterpart.
"""
import math

import numpy as np
import pytest
from scipy.optimize import minimize

from bdem import scenarios as S
from bdem import types as T
from oracle_py import Oracle

G = 981.0
DT = float(np.float32(1e-5))


def one_sphere(pos, wobj, vel=(0, 0, 0), avel=(0, 0, 0), law="ref", r=0.5, dt=1e-5, gravity=(0, -G, 0),
               rotation=True, e=0.6, mu=0.4, kn=1e5, E=1e5, nu=0.3, L=40.0, cones=(), mu_s=None):
    sim = S.sim_info(1, dt, gravity, (-1.0, -1.0, -1.0), (L, L, L), 2.0 * r, rotation=rotation)
    pobj = [S.sphere_pobj(r, 2500.0, e, kn, mu, mu_s=mu_s)]
    if law == "ref":
        ext = T.ExtParams() if cones else None
    elif law == "lsd":
        ext = S.lsd_ext(kn, e, mu)
    else:
        ext = S.hm_ext(E, nu, e, mu)
    if cones:
        ext.num_cones = len(cones)
        for k, c in enumerate(cones):
            ext.cone[k] = c
    d = S._deck(sim, wobj, pobj, np.array([pos]), np.array([vel]), ext=ext)
    d["avel"] = np.array([avel], np.float32)
    return d


def far_plane():
    # a floor far below every test position: keeps the WOBJ table non-empty
    return S.infinite_plane_wobj((0.0, -30.0, 0.0), (0.0, 1.0, 0.0))


def make_cone(base, axis, s0, s1, r0, r1):
    c = T.Cone()
    c.base = T.f3(*base)
    c.axis = T.f3(*axis)
    c.s0, c.s1, c.r0, c.r1 = s0, s1, r0, r1
    return c


def estar_gstar(E, nu):
    """Textbook effective moduli of two identical bodies: 1/E* = 2(1-nu^2)/E,
    G = E/(2(1+nu)), 1/G* = 2(2-nu)/G."""
    Gm = E / (2.0 * (1.0 + nu))
    return E / (2.0 * (1.0 - nu * nu)), Gm / (2.0 * (2.0 - nu))


def beta_of(e):
    return math.log(e) / math.sqrt(math.log(e) ** 2 + math.pi ** 2)


def wall_force(deck):
    o = Oracle(deck, bits=64)
    _, _, fw, tw, _ = o.forces()
    return fw[0], tw[0]


# ------------------------------------------------------------------- cone wall
CONE = dict(base=(20.0, 0.0, 20.0), axis=(0.0, 1.0, 0.0), s0=5.0, s1=15.0, r0=3.0, r1=9.0)


def cone_brute_force(p, c, r):
    """Nearest point of the frustum's lateral surface by dense sampling of
    (t, phi) and a bounded local refinement; returns depth r - dist and the
    unit normal from that point to the centre (None when dist >= r)."""
    base, ax = np.array(c["base"]), np.array(c["axis"])
    u = np.cross(ax, [1.0, 0.0, 0.0] if abs(ax[0]) < 0.9 else [0.0, 0.0, 1.0])
    u /= np.linalg.norm(u)
    w = np.cross(ax, u)

    def surf(t, ph):
        s = c["s0"] + t * (c["s1"] - c["s0"])
        rr = c["r0"] + t * (c["r1"] - c["r0"])
        return base + s * ax + rr * (np.cos(ph) * u + np.sin(ph) * w)

    t = np.linspace(0.0, 1.0, 801)
    ph = np.linspace(0.0, 2 * np.pi, 721)
    tt, pp = np.meshgrid(t, ph, indexing="ij")
    s = c["s0"] + tt * (c["s1"] - c["s0"])
    rr = c["r0"] + tt * (c["r1"] - c["r0"])
    pts = base + s[..., None] * ax + rr[..., None] * (np.cos(pp)[..., None] * u + np.sin(pp)[..., None] * w)
    d2 = np.sum((pts - p) ** 2, axis=-1)
    k = np.unravel_index(np.argmin(d2), d2.shape)
    res = minimize(lambda x: np.sum((surf(x[0], x[1]) - p) ** 2), [t[k[0]], ph[k[1]]],
                   method="L-BFGS-B", bounds=[(0.0, 1.0), (None, None)], options=dict(ftol=1e-30, gtol=1e-14))
    q = surf(*res.x)
    dist = float(np.linalg.norm(p - q))
    if dist >= r:
        return None
    return r - dist, (p - q) / dist


@pytest.mark.parametrize("s,rho,phi", [
    (10.0, 5.5, 0.3),    # inside the hopper, on the slanted wall
    (10.0, 6.5, 2.0),    # outside the frustum
    (15.3, 9.2, 4.0),    # above the wide rim: nearest point on the rim circle
    (4.7, 2.8, 5.5),     # below the narrow rim
    (12.0, 7.0, 1.0),    # inside, deeper
])
def test_cone_contact_matches_brute_force(s, rho, phi):
    c = CONE
    p = np.array(c["base"]) + np.array([rho * math.cos(phi), s, rho * math.sin(phi)])
    p = p.astype(np.float32).astype(np.float64)        # the position the deck stores
    r = 0.5
    want = cone_brute_force(p, c, r)
    assert want is not None
    deck = one_sphere(tuple(p), [far_plane()], cones=[make_cone(**c)])
    kn_s = deck["pobj"][0].surface_Kn
    F, _ = wall_force(deck)
    depth, n = want
    # at rest (v = 0, omega = 0) the static response is Kn_s * depth * n; the
    # search fixes the distance to ~1e-12 and the nearest point to ~1e-7
    assert np.linalg.norm(F) == pytest.approx(kn_s * depth, rel=1e-9)
    assert np.allclose(F / np.linalg.norm(F), n, rtol=0, atol=1e-6)


def test_cone_far_and_axis_give_no_contact():
    c = CONE
    for p in [(23.0, 10.0, 20.0), (20.0, 10.0, 20.0), (20.0, 30.0, 20.0)]:
        deck = one_sphere(p, [far_plane()], cones=[make_cone(**c)])
        F, _ = wall_force(deck)
        assert np.all(F == 0.0), p


@pytest.mark.parametrize("law", ["ref", "hm"])
def test_sphere_at_rest_on_a_flat_cone(law):
    """A cone with s0 == s1 is an annular disc: a sphere settles on it with
    the floor's overlap, m g / Kn_s (reference law) or the Hertz overlap
    (3 m g / (4 E* sqrt r))^(2/3)."""
    r = 0.5
    cone = make_cone((20.0, 0.0, 20.0), (0.0, 1.0, 0.0), 5.0, 5.0, 0.0, 15.0)
    deck = one_sphere((26.0, 5.0 + r, 20.0), [far_plane()], law=law, cones=[cone], rotation=False)
    m = float(deck["pobj"][0].mass)
    o = Oracle(deck, bits=64)
    o.step(20000)
    st = o.state()
    depth = 5.0 + r - st["pos"][0, 1]
    if law == "ref":
        want = m * G / deck["pobj"][0].surface_Kn
    else:
        es, _ = estar_gstar(1e5, 0.3)
        want = (3.0 * m * G / (4.0 * es * math.sqrt(r))) ** (2.0 / 3.0)
    assert depth == pytest.approx(want, rel=2e-4)
    assert np.abs(st["vel"][0]).max() < 1e-4


def test_sphere_sticks_on_an_inclined_cone():
    """Hertz-Mindlin with friction: a sphere (no rotation) on the outside of a
    cone whose surface is inclined 15 degrees (tan 15 < mu) comes to rest with
    normal overlap (3 m g cos(15) / (4 E* sqrt r))^(2/3) and a tangential force
    m g sin(15) up the slope; the wall force balances gravity."""
    r, E, nu, mu = 0.5, 1e5, 0.3, 0.5
    incl = math.radians(15.0)
    # axis y, apex up: radius shrinks with s; slope angle from horizontal = incl
    s0, s1, r0 = 0.0, 10.0, 10.0 / math.tan(incl)
    cone = make_cone((20.0, 0.0, 20.0), (0.0, 1.0, 0.0), s0, s1, r0, 0.0)
    # touching point on the meridian at t = 0.5, sphere centre one radius out along the normal
    sm, rm = 5.0, r0 / 2.0
    nrm = np.array([math.sin(incl), math.cos(incl)])    # (radial, axial) outward normal
    rho, y = rm + r * nrm[0], sm + r * nrm[1]
    deck = one_sphere((20.0 + rho, y, 20.0), [far_plane()], law="hm", cones=[cone], rotation=False, mu=mu,
                      E=E, nu=nu, L=60.0)
    m = float(deck["pobj"][0].mass)
    o = Oracle(deck, bits=64)
    o.step(20000)
    st = o.state()
    p = st["pos"][0]
    # distance from the meridian line through (s0, r0)-(s1, 0)
    rho_n, y_n = math.hypot(p[0] - 20.0, p[2] - 20.0), p[1]
    dist = (rho_n - rm) * nrm[0] + (y_n - sm) * nrm[1]
    es, _ = estar_gstar(E, nu)
    want = (3.0 * m * G * math.cos(incl) / (4.0 * es * math.sqrt(r))) ** (2.0 / 3.0)
    assert r - dist == pytest.approx(want, rel=1e-3)
    _, _, fw, _, _ = o.forces()
    assert fw[0] == pytest.approx([0.0, m * G, 0.0], abs=2e-4 * m * G)
    assert np.abs(st["vel"][0]).max() < 1e-4


# ------------------------------------------------------------- history-law probe
def _probe_oracle(law, e=0.5, mu=0.5, kn=1e5, E=1e5, nu=0.3):
    deck = one_sphere((10.0, 10.0, 10.0), [far_plane()], law=law, e=e, mu=mu, kn=kn, E=E, nu=nu)
    return Oracle(deck, bits=64)


def test_hertz_mindlin_stiffness_damping_and_normal_force():
    E, nu, e = 1e5, 0.3, 0.5
    o = _probe_oracle("hm", e=e, E=E, nu=nu)
    es, gs = estar_gstar(E, nu)
    b = beta_of(e)
    n = np.array([0.0, 1.0, 0.0])
    rstar, mstar, delta = 0.25, 6e-4, 2e-3
    sq = math.sqrt(rstar * delta)
    fn0 = 4.0 / 3.0 * es * math.sqrt(rstar) * delta ** 1.5
    kt = 8.0 * gs * sq
    sn, st = 2.0 * es * sq, 8.0 * gs * sq
    gn = -2.0 * math.sqrt(5.0 / 6.0) * b * math.sqrt(sn * mstar)
    gt = -2.0 * math.sqrt(5.0 / 6.0) * b * math.sqrt(st * mstar)
    tol = 2e-6
    # static overlap: Hertz normal force only
    F, xi = o.history_law(n, delta, (0, 0, 0), rstar, mstar)
    assert F == pytest.approx([0.0, fn0, 0.0], rel=tol)
    # a stored spring with no slip: F_t = -kt xi (the spring is kept as it is)
    x0 = np.array([3e-5, 0.0, -4e-5])
    F, xi = o.history_law(n, delta, (0, 0, 0), rstar, mstar, xi=x0)
    assert F[[0, 2]] == pytest.approx(-kt * x0[[0, 2]], rel=tol)
    assert xi == pytest.approx(x0, rel=1e-12)
    # first step of a sliding contact: xi = vt dt, F_t = -(kt dt + gt) vt
    vt = np.array([2.0, 0.0, 0.0])
    F, xi = o.history_law(n, delta, vt, rstar, mstar)
    assert xi == pytest.approx(vt * DT, rel=1e-12)
    assert F[0] == pytest.approx(-(kt * DT + gt) * vt[0], rel=tol)
    assert F[1] == pytest.approx(fn0, rel=tol)
    # approach speed: Tsuji normal damping -gn vn on top of Hertz
    vn = -3.0
    F, _ = o.history_law(n, delta, (0, vn, 0), rstar, mstar)
    assert F[1] == pytest.approx(fn0 - gn * vn, rel=tol)
    # a separating contact whose damping exceeds the elastic part: no pull
    F, _ = o.history_law(n, delta, (0, 1e4, 0), rstar, mstar)
    assert F[1] == 0.0


def test_coulomb_cap_and_spring_slide():
    o = _probe_oracle("hm", mu=0.5)
    n = np.array([0.0, 0.0, 1.0])
    rstar, mstar, delta = 0.25, 6e-4, 2e-3
    es, gs = estar_gstar(1e5, 0.3)
    fn = 4.0 / 3.0 * es * math.sqrt(rstar) * delta ** 1.5
    kt = 8.0 * gs * math.sqrt(rstar * delta)
    big = np.array([1.0, -1.0, 0.0])   # far beyond mu fn / kt
    F, xi = o.history_law(n, delta, (0, 0, 0), rstar, mstar, xi=big)
    ft = F[:2]
    # the law's constants are stored in float: 1e-6 against the textbook values
    assert np.linalg.norm(ft) == pytest.approx(0.5 * fn, rel=1e-6)
    assert ft / np.linalg.norm(ft) == pytest.approx(-big[:2] / np.linalg.norm(big[:2]), rel=1e-12)
    # the spring slides back onto the cone: kt |xi| = mu fn
    assert xi[:2] == pytest.approx(-ft / kt, rel=1e-6)


def test_stick_slip_and_unloading_sequence():
    """Constant slip velocity u: the spring grows by u dt per step while
    |F_t| < mu F_n, then the force stays on the cap; reversing u unloads the
    spring linearly from the cap (every step kt u dt), with the damping term
    changing sign once."""
    o = _probe_oracle("lsd", e=0.5, mu=0.3, kn=1e5)
    kn = 1e5
    kt = kn * 2.0 / 7.0
    n = np.array([1.0, 0.0, 0.0])
    delta, rstar, mstar = 1e-3, 0.25, 6e-4
    fn = kn * delta
    u = 4.7   # the cap falls between steps 22 and 23
    xi = np.zeros(3)
    caps = 0
    xs = []
    for k in range(1, 40):
        F, xi = o.history_law(n, delta, (0, u, 0), rstar, mstar, xi=xi)
        xs.append(xi[1])
        if -F[1] >= 0.3 * fn * (1 - 1e-12):
            caps += 1
            assert -F[1] == pytest.approx(0.3 * fn, rel=1e-6)   # float-stored kn, mu
            assert xi[1] == pytest.approx(0.3 * fn / kt, rel=1e-6)
        else:
            assert xi[1] == pytest.approx(k * u * DT, rel=1e-10)   # LSD: no tangential damping
            assert F[1] == pytest.approx(-kt * k * u * DT, rel=1e-6)
    # steps 1..floor(mu fn / (kt u dt)) stick, the rest of the 39 slide
    assert caps == 39 - int(math.floor(0.3 * fn / (kt * u * DT)))
    # unloading: from the cap, each reversed step removes u dt of spring
    x0 = xi[1]
    for k in range(1, 6):
        F, xi = o.history_law(n, delta, (0, -u, 0), rstar, mstar, xi=xi)
        assert xi[1] == pytest.approx(x0 - k * u * DT, rel=1e-10)
        assert F[1] == pytest.approx(-kt * xi[1], rel=1e-6)
    # normal force of the linear law with its damping -2 beta sqrt(kn m*)
    F, _ = o.history_law(n, delta, (-2.0, 0, 0), rstar, mstar)
    cn = -2.0 * beta_of(0.5) * math.sqrt(kn * mstar)
    assert F[0] == pytest.approx(fn + cn * 2.0, rel=1e-6)


def test_spring_rotation_keeps_length_and_tangency():
    o = _probe_oracle("lsd", mu=10.0)
    rng = np.random.default_rng(3)
    for _ in range(20):
        n = rng.normal(size=3)
        n /= np.linalg.norm(n)
        x0 = rng.normal(size=3) * 1e-4
        F, xi = o.history_law(n, 1e-3, (0, 0, 0), 0.25, 6e-4, xi=x0)
        assert np.linalg.norm(xi) == pytest.approx(np.linalg.norm(x0), rel=1e-12)
        assert abs(np.dot(xi, n)) < 1e-12 * np.linalg.norm(x0)
        # the tangential force is -kt times the rotated spring
        ft = F - np.dot(F, n) * n
        assert ft == pytest.approx(-1e5 * 2.0 / 7.0 * xi, rel=1e-6, abs=1e-12)


# ------------------------------------------------------ history-law walls at rest
def _settle_depth(deck, steps, axis, surface):
    o = Oracle(deck, bits=64)
    o.step(steps)
    st = o.state()
    return st, o


@pytest.mark.parametrize("law", ["hm", "lsd"])
@pytest.mark.parametrize("wall", ["plane", "cylinder", "cap"])
def test_history_law_walls_at_rest(law, wall):
    """Plane, inside of a cylinder and cylinder end cap under the history
    laws: a sphere at rest carries its weight with the Hertz overlap
    (3 m g / (4 E* sqrt r))^(2/3) (effective radius r against a wall) or the
    linear one m g / kn."""
    r, E, nu, kn = 0.5, 1e5, 0.3, 1e5
    if wall == "plane":
        walls = [S.infinite_plane_wobj((0.0, 2.0, 0.0), (0.0, 1.0, 0.0))]
        p, grav, top = (10.0, 2.0 + r, 10.0), (0, -G, 0), 2.0
    elif wall == "cylinder":
        # axis z through (R, R): the sphere rests at the bottom, y = 0
        R = 8.0
        walls = [S.cylinder_wobj(R, 10.0, (0.0, 0.0, 1.0), (0.0, 0.0, 0.0))]
        p, grav, top = (R, r, 5.0), (0, -G, 0), 0.0
    else:
        # axis y cylinder with its bottom cap at y = 1; gravity along -y
        walls = [S.cylinder_wobj(8.0, 10.0, (0.0, 1.0, 0.0), (0.0, 1.0, 0.0), bot=True)]
        p, grav, top = (8.0, 1.0 + r, 8.0), (0, -G, 0), 1.0
    deck = one_sphere(p, walls, law=law, rotation=False, E=E, nu=nu, kn=kn, gravity=grav)
    m = float(deck["pobj"][0].mass)
    o = Oracle(deck, bits=64)
    o.step(20000)
    st = o.state()
    depth = top + r - st["pos"][0, 1]
    if law == "hm":
        es, _ = estar_gstar(E, nu)
        want = (3.0 * m * G / (4.0 * es * math.sqrt(r))) ** (2.0 / 3.0)
    else:
        want = m * G / kn
    assert depth == pytest.approx(want, rel=2e-4)
    assert np.abs(st["vel"][0]).max() < 1e-4


@pytest.mark.parametrize("law", ["hm", "lsd"])
def test_sliding_block_decelerates_at_mu_g(law):
    """A sphere with rotation off sliding on a floor: once the spring reaches
    the Coulomb cap the friction is mu m g, so v(t) = u - mu g t."""
    r, mu, u = 0.5, 0.3, 20.0
    walls = [S.infinite_plane_wobj((0.0, 2.0, 0.0), (0.0, 1.0, 0.0))]
    deck = one_sphere((5.0, 2.0 + r, 10.0), walls, vel=(u, 0, 0), law=law, rotation=False, mu=mu, L=60.0)
    o = Oracle(deck, bits=64)
    o.step(3000)      # settle the normal contact and reach the cap
    v1 = o.state()["vel"][0, 0]
    o.step(2000)
    v2 = o.state()["vel"][0, 0]
    assert v2 > 0.0
    assert (v1 - v2) / (2000 * DT) == pytest.approx(mu * G, rel=2e-3)


# ---------------------------------------------------------- reference-law walls
def test_static_wall_friction_decelerates_by_005_mu_per_step():
    """Reference static surface (Collision_Response.cuh:63-79): a sphere
    (rotation off) sliding on a floor at its rest overlap loses 0.05 mu_s of
    speed per step, |Ft| = 0.05 (m/dt) mu_s |vt|/|v| with vt = v."""
    r, mu, u = 0.5, 0.4, 10.0
    deck = one_sphere((5.0, 2.0 + r, 10.0), [S.infinite_plane_wobj((0.0, 2.0, 0.0), (0.0, 1.0, 0.0))],
                      vel=(u, 0, 0), rotation=False, mu=mu, L=60.0)
    p = deck["pobj"][0]
    m = float(np.float32(p.mass))
    deck["pos"][0, 1] = np.float32(2.0 + r - m * G / p.surface_Kn)   # rest overlap
    o = Oracle(deck, bits=64)
    o.step(50)
    v = o.state()["vel"][0]
    assert v[0] == pytest.approx(u - 50 * 0.05 * mu, rel=1e-6)
    assert abs(v[1]) < 1e-3    # the float rest overlap is off by an ulp


def _drum(mill_rot_vel, R=10.0, Lz=8.0, caps=False):
    d = S.cylinder_wobj(R, Lz, (0.0, 0.0, 1.0), (0.0, 0.0, 0.0), top=caps, bot=caps)
    return d


def _drum_force(pos, vel, avel, rotating, W, mu=0.4, law="ref"):
    deck = one_sphere(pos, [_drum(W)], vel=vel, avel=avel, mu=mu, law=law)
    deck["sim"].MillRotVel = W if rotating else 0.0
    o = Oracle(deck, bits=64)
    o.set_drum_rotation(rotating)
    _, _, fw, tw, _ = o.forces()
    return fw[0], tw[0], deck


@pytest.mark.parametrize("law", ["ref", "hm"])
def test_moving_drum_wall_is_the_static_wall_in_the_wall_frame(law):
    """The drum (axis z through (R, R)) turns with angular velocity
    (0, 0, -MillRotVel); its surface velocity at the contact point
    cp = p - r n is w_d x (cp - c).  With mu_static = mu_kinetic a sphere
    with velocity v + v_wall against the moving drum feels the force a sphere
    with velocity v feels against the static drum (Galilean invariance of
    Collision_Response.cuh:116-233 against :4-109)."""
    R, r, W = 10.0, 0.5, 37.0
    rng = np.random.default_rng(7)
    for _ in range(6):
        ang = rng.uniform(0, 2 * np.pi)
        dist = R - r + 0.08                              # overlap 0.08 with the drum
        pos = np.array([R + dist * math.cos(ang), R + dist * math.sin(ang), 4.0])
        n = -(pos[:2] - R) / np.linalg.norm(pos[:2] - R)  # wall normal, towards the centre
        cp = pos - r * np.array([n[0], n[1], 0.0])
        vwall = np.cross([0.0, 0.0, -W], cp - [R, R, 0.0])
        v = rng.normal(size=3) * 5.0
        w = rng.normal(size=3) * 3.0
        Fs, Ts, _ = _drum_force(tuple(pos), tuple(v), tuple(w), False, W, law=law)
        Fm, Tm, _ = _drum_force(tuple(pos), tuple(v + vwall), tuple(w), True, W, law=law)
        scale = np.abs(Fs).max()
        assert np.allclose(Fm, Fs, rtol=0, atol=1e-6 * scale)
        assert np.allclose(Tm, Ts, rtol=0, atol=1e-6 * max(np.abs(Ts).max(), scale))


def test_co_moving_sphere_feels_only_the_spring():
    """The reference's friction is mu (m/dt) 0.05 |vt|/|v - v_wall|: any
    tangential slip, however small, gets the full magnitude.  So the values
    are chosen exact in binary (overlap 1/16, wall speed 37 * 10.0625): a
    sphere moving with the wall surface feels the spring alone."""
    R, r, W, dep = 10.0, 0.5, 37.0, 0.0625
    pos = np.array([R, r - dep, 4.0])                     # bottom of the drum
    cp = pos - r * np.array([0.0, 1.0, 0.0])
    vwall = np.cross([0.0, 0.0, -W], cp - [R, R, 0.0])
    assert np.all(vwall.astype(np.float32) == vwall)
    F, T_, deck = _drum_force(tuple(pos), tuple(vwall), (0, 0, 0), True, W)
    kn_s = deck["pobj"][0].surface_Kn
    assert F == pytest.approx([0.0, kn_s * dep, 0.0], rel=1e-12, abs=1e-12 * kn_s)
    assert np.abs(T_).max() < 1e-12 * kn_s
    # the wall moving the other way slips by 2 |v_wall|: full kinetic friction
    F2, _, _ = _drum_force(tuple(pos), tuple(-vwall), (0, 0, 0), True, W)
    m = float(deck["pobj"][0].mass)
    assert F2[0] == pytest.approx(0.05 * (m / DT) * 0.4 * np.sign(vwall[0]), rel=1e-6)


def test_end_plate_rest_frictionless_and_overwrite():
    """End plates (Collision_Response.cuh:504-526): spring-dashpot only, so a
    sphere settles on a cap with overlap m g / Kn_s and slides on it without
    losing speed or spinning up; a sphere touching the drum and a cap gets the
    cap force alone (the force is assigned, DEM_Kernels_Compute.cuh:539, 585)."""
    R, r, u = 8.0, 0.5, 7.0
    cyl = S.cylinder_wobj(R, 10.0, (0.0, 1.0, 0.0), (0.0, 1.0, 0.0), bot=True)
    deck = one_sphere((8.0, 1.0 + r, 8.0), [cyl], vel=(u, 0, 0), rotation=True)
    p = deck["pobj"][0]
    m = float(np.float32(p.mass))
    o = Oracle(deck, bits=64)
    o.step(300)
    st = o.state()
    assert st["vel"][0, 0] == u                          # no friction at all
    assert np.all(st["avel"][0] == 0.0)                  # no torque
    o.step(20000)
    st = o.state()
    depth = 1.0 + r - st["pos"][0, 1]
    assert depth == pytest.approx(m * G / p.surface_Kn, rel=1e-4)
    # corner: drum wall (x) and bottom cap (y) at once, sphere at rest
    dcap, ddrum = 0.03, 0.06
    pos = (8.0 + R - r + ddrum, 1.0 + r - dcap, 8.0)
    deck = one_sphere(pos, [cyl], rotation=True)
    o = Oracle(deck, bits=64)
    _, _, fw, _, _ = o.forces()
    assert fw[0] == pytest.approx([0.0, p.surface_Kn * dcap, 0.0], rel=1e-5, abs=1e-9)


def test_end_plate_restitution_is_the_oscillator_value():
    """A normal impact on an end plate: the unclamped linear spring-dashpot
    returns e' = e - 0.03808 (the value its damping is tuned for,
    Input/KSimulationData.cpp:993-996), up to the time discretisation."""
    R, r, v0, e = 8.0, 0.5, 30.0, 0.6
    cyl = S.cylinder_wobj(R, 10.0, (0.0, 1.0, 0.0), (0.0, 1.0, 0.0), bot=True)
    deck = one_sphere((8.0, 1.0 + r + 0.0001, 8.0), [cyl], vel=(0, -v0, 0), e=e, gravity=(0, 0, 0))
    o = Oracle(deck, bits=64)
    o.step(600)
    vy = o.state()["vel"][0, 1]
    assert vy / v0 == pytest.approx(e - 0.03808, abs=0.01)
