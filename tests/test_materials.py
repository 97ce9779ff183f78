"""Setup-time spring constants (bdem_derive_spring_constants, host only):
against the damped linear oscillator they are built from, and against the
reference's documented quirks (shifted restitution, same-type reduced mass
m/2 in KN mode, own mass in TC mode, row-order friction averaging)."""
import math

import numpy as np
import pytest

from bdem import lib as L
from bdem import scenarios as S
from bdem import types as T


def _sphere(radius, kn, mu):
    p = S.sphere_pobj(radius, 2500.0, 0.5, kn, mu)
    return p


def _zeta(cn, kn, m):
    return cn / (2.0 * math.sqrt(kn * m))


def _shift(e):
    return e - 0.03808 if 0.3 <= e <= 0.9 else e


@pytest.mark.parametrize("e", [0.2, 0.5, 0.9, 0.95])
def test_kn_mode_same_type_damping(e):
    p = _sphere(0.5, 1e5, 0.4)
    p.PP[0].Kn = 1e5
    p.surface_Kn = 2e5
    p.Dsurface_Kn = 3e5
    [q] = L.derive_spring_constants([p], L.SPRING_KN, [e], [e], [e])
    m = float(np.float32(q.mass))
    es = _shift(e)
    want = -math.log(es) / math.sqrt(math.log(es) ** 2 + math.pi ** 2)
    # same type: the reduced mass is m/2 (Input/KSimulationData.cpp:958)
    assert _zeta(q.PP[0].Cn, 1e5, m / 2) == pytest.approx(want, rel=1e-6)
    assert _zeta(q.surface_Cn, 2e5, m) == pytest.approx(want, rel=1e-6)
    assert _zeta(q.Dsurface_Cn, 3e5, m) == pytest.approx(want, rel=1e-6)
    assert q.PP[0].Cn == pytest.approx(S.cn_from_cor(float(np.float32(1e5)), m / 2, e), rel=1e-7)


def test_kn_mode_elastic_has_no_damping():
    p = _sphere(0.5, 1e5, 0.4)
    p.PP[0].Kn = 1e5
    [q] = L.derive_spring_constants([p], L.SPRING_KN, [1.0], [1.0], [1.0])
    assert q.PP[0].Cn == 0.0 and q.surface_Cn == 0.0 and q.Dsurface_Cn == 0.0


def test_kn_mode_cross_type_averages():
    a, b = _sphere(0.5, 1e5, 0.4), _sphere(0.3, 3e5, 0.2)
    a.PP[0].Kn, b.PP[1].Kn = 1e5, 3e5
    a.PP[0].Fric_static, a.PP[0].Fric_kinetic = 0.4, 0.3
    b.PP[1].Fric_static, b.PP[1].Fric_kinetic = 0.2, 0.1
    a, b = L.derive_spring_constants([a, b], L.SPRING_KN, [0.6, 0.8], [0.5, 0.5], [0.5, 0.5])
    assert a.PP[1].Kn == pytest.approx(2e5) and b.PP[0].Kn == pytest.approx(2e5)
    assert a.PP[1].Fric_static == pytest.approx(0.3) and a.PP[1].Fric_kinetic == pytest.approx(0.2)
    ma, mb = float(np.float32(a.mass)), float(np.float32(b.mass))
    mr = ma * mb / (ma + mb)
    es = _shift(0.7)
    want = -math.log(es) / math.sqrt(math.log(es) ** 2 + math.pi ** 2)
    assert _zeta(a.PP[1].Cn, 2e5, mr) == pytest.approx(want, rel=1e-6)
    assert a.PP[1].Cn == b.PP[0].Cn


@pytest.mark.parametrize("e", [0.25, 0.6])
def test_tc_mode_contact_time_and_restitution(e):
    """Kn = m (ln^2 e' + pi^2)/tc^2, Cn = -2 m ln e'/tc: the damped linear
    oscillator of mass m then lasts exactly tc (omega_d = pi/tc) and returns
    e' of the approach speed."""
    tc = 2e-4
    a, b = _sphere(0.5, 0.0, 0.4), _sphere(0.3, 0.0, 0.2)
    a, b = L.derive_spring_constants([a, b], L.SPRING_TC, [e, e], [e, e], [e, e], tc=(tc, tc, tc))
    ma, mb = float(np.float32(a.mass)), float(np.float32(b.mass))
    es = _shift(e)
    for kn, cn, m in ((a.PP[0].Kn, a.PP[0].Cn, ma),            # same type: the particle's own mass
                      (a.PP[1].Kn, a.PP[1].Cn, ma * mb / (ma + mb)),
                      (a.surface_Kn, a.surface_Cn, ma), (b.Dsurface_Kn, b.Dsurface_Cn, mb)):
        w0 = math.sqrt(kn / m)
        gam = cn / (2 * m)
        wd = math.sqrt(w0 * w0 - gam * gam)
        assert math.pi / wd == pytest.approx(tc, rel=1e-6)
        assert math.exp(-gam * tc) == pytest.approx(es, rel=1e-6)


def test_tc_mode_friction_row_order():
    """Set_SpringConstantsParticle_TC averages PP[0] in place: for type 1,
    PP[0] is overwritten with the mean before PP[1] reads it, so the
    same-type friction of type 1 is the two-type mean."""
    a, b = _sphere(0.5, 0.0, 0.4), _sphere(0.3, 0.0, 0.2)
    a.PP[0].Fric_static, b.PP[0].Fric_static = 0.4, 0.2
    a, b = L.derive_spring_constants([a, b], L.SPRING_TC, [0.5, 0.5], [0.5, 0.5], [0.5, 0.5],
                                     tc=(1e-4, 1e-4, 1e-4))
    assert a.PP[0].Fric_static == pytest.approx(0.4)
    assert a.PP[1].Fric_static == pytest.approx(0.3)
    assert b.PP[0].Fric_static == pytest.approx(0.3)
    assert b.PP[1].Fric_static == pytest.approx(0.3)


def test_rejects_bad_input():
    p = _sphere(0.5, 1e5, 0.4)
    with pytest.raises(L.BdemError):
        L.derive_spring_constants([p] * 7, L.SPRING_KN, [0.5] * 7, [0.5] * 7, [0.5] * 7)
    with pytest.raises(L.BdemError):
        L.derive_spring_constants([p], L.SPRING_TC, [0.5], [0.5], [0.5], tc=(0.0, 1e-4, 1e-4))
