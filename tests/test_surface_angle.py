"""The angle-of-repose fit (bdem_surface_angle, host only, no device):
a least-squares line through the non-empty bins of a surface profile.
Pinned by closed forms (tops on a line of known slope give its angle) and by
numpy's polyfit on random profiles."""
import math

import numpy as np
import pytest

import bdem
from bdem.lib import BdemError, surface_angle


@pytest.mark.parametrize("deg", [0.0, 12.5, 30.0, 45.0, 60.0])
@pytest.mark.parametrize("sign", [1.0, -1.0])
def test_tops_on_a_line_give_its_angle(deg, sign):
    lo, hi, bins = 2.0, 14.0, 24
    u = lo + (np.arange(bins) + 0.5) * (hi - lo) / bins
    top = (7.0 + sign * math.tan(math.radians(deg)) * u).astype(np.float32)
    slope, icpt, ang = surface_angle(top, lo, hi)
    assert ang == pytest.approx(deg, abs=1e-4)
    assert slope == pytest.approx(sign * math.tan(math.radians(deg)), abs=1e-5)
    assert icpt == pytest.approx(7.0, abs=1e-4)


def test_empty_bins_are_skipped():
    lo, hi, bins = 0.0, 20.0, 40
    u = lo + (np.arange(bins) + 0.5) * (hi - lo) / bins
    top = (3.0 + 0.4 * u).astype(np.float32)
    top[::3] = np.nan
    top[-5:] = np.nan
    slope, _, _ = surface_angle(top, lo, hi)
    assert slope == pytest.approx(0.4, abs=1e-6)


def test_random_profile_matches_polyfit():
    rng = np.random.default_rng(3)
    lo, hi, bins = -3.0, 17.0, 57
    top = rng.uniform(0, 10, bins).astype(np.float32)
    top[rng.random(bins) < 0.2] = np.nan
    u = lo + (np.arange(bins) + 0.5) * (hi - lo) / bins
    ok = ~np.isnan(top)
    b, a = np.polyfit(u[ok], top[ok].astype(np.float64), 1)
    slope, icpt, ang = surface_angle(top, lo, hi)
    assert slope == pytest.approx(b, rel=1e-9)
    assert icpt == pytest.approx(a, rel=1e-9)
    assert ang == pytest.approx(math.degrees(math.atan(abs(b))), rel=1e-9)


def test_fewer_than_two_bins_is_an_error():
    top = np.full(10, np.nan, np.float32)
    top[4] = 1.0
    with pytest.raises(BdemError):
        surface_angle(top, 0.0, 10.0)
    assert "bdem_get_surface_profile" in bdem.lib.EXPORTED
