"""The allocation-free surface profile (bdem_get_surface_profile_noalloc)
against bdem_get_surface_profile: bit-identical heights (NaN where empty)
on two calls in a row, the second on a smaller pack with fewer bins, so the
shared scratch must be cleared between calls."""
import numpy as np
import pytest

from bdem import scenarios as S
from helpers import make_device

pytestmark = pytest.mark.gpu


def _same(a, b):
    assert a.shape == b.shape
    assert np.array_equal(np.isnan(a), np.isnan(b))
    ok = ~np.isnan(a)
    assert np.array_equal(a[ok].view(np.uint32), b[ok].view(np.uint32))


def test_noalloc_profile_equals_profile_twice():
    big = S.random_deck(6000, seed=61, L=14.0)
    small = S.random_deck(700, seed=62, L=9.0)
    print("[case] noalloc profile: 6000 then 700 spheres", flush=True)
    a = make_device(big)
    a.step(3)
    ref1 = a.surface_profile(0.0, 14.0, 512, axis=0)
    got1 = a.surface_profile_noalloc(0.0, 14.0, 512, axis=0)
    _same(got1, ref1)
    assert np.count_nonzero(~np.isnan(got1)) > 400
    b = make_device(small)
    b.step(2)
    ref2 = b.surface_profile(-1.0, 10.0, 96, axis=2)
    got2 = b.surface_profile_noalloc(-1.0, 10.0, 96, axis=2)
    _same(got2, ref2)
    # the first context again after the second: nothing left over
    _same(a.surface_profile_noalloc(0.0, 14.0, 512, axis=0), ref1)
    a.close()
    b.close()
