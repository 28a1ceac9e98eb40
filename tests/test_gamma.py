"""Known answers of the gamma-index checker (dpm-g4cpp_amd/python/gamma.py),
which every 1 %/1 mm acceptance in the tests and the benchmark goes through.

The global gamma of Low et al. (Med. Phys. 25 (1998) 656) at a reference
point r is min over r' of sqrt((D_e(r') - D_r(r))^2 / (dd Dmax)^2 + |r' - r|^2
/ dta^2). Closed forms used here:

* identical distributions: gamma = 0 everywhere;
* a flat field offset by c: no spatial shift helps, gamma = c / (dd Dmax);
* a linear ramp D_r(x) = a x against D_e(x) = a (x - s) + c: gamma is the
  distance from a point to a line in the (dose / dd Dmax, position / dta)
  plane, |c - a s| / (dd Dmax) / sqrt(1 + (a dta / (dd Dmax))^2);
* points below the cutoff (10 % of the reference maximum) are not counted.
"""
import numpy as np
import pytest

import gamma


def test_identical_distributions_pass_everywhere():
    rng = np.random.default_rng(1)
    d = rng.uniform(0.0, 1.0, size=(12, 10, 8))
    g, mask = gamma.gamma_map(d, d, spacing_mm=2.0)
    assert np.all(g[mask] == 0.0)
    assert gamma.gamma_pass_rate_3d(d, d, 2.0) == 1.0
    assert gamma.gamma_pass_rate_1d(d[:, 0, 0], d[:, 0, 0], 1.0) == 1.0


@pytest.mark.parametrize("offset,expect", [(0.02, 0.0), (0.005, 1.0), (0.0099, 1.0), (0.0101, 0.0)])
def test_flat_field_offset(offset, expect):
    ref = np.full(40, 2.0)
    ev = ref * (1.0 + offset)
    g, mask = gamma.gamma_map(ref, ev, spacing_mm=1.0, sub=8)
    np.testing.assert_allclose(g[mask], offset / 0.01, rtol=1e-9)
    assert gamma.gamma_pass_rate_1d(ref, ev, 1.0) == expect


def _ramp(n=80, spacing=1.0, a=0.02, s=0.0, c=0.0):
    x = np.arange(n) * spacing
    ref = 1.0 + a * x
    ev = 1.0 + a * (x - s) + c
    return x, ref, ev


@pytest.mark.parametrize("a,s,c", [(0.02, 0.5, 0.0), (0.05, 0.3, 0.004), (0.01, -0.7, 0.01), (0.08, 1.5, 0.0)])
def test_ramp_gamma_is_the_point_to_line_distance(a, s, c):
    x, ref, ev = _ramp(a=a, s=s, c=c)
    dmax = ref.max()
    dd, dta = 0.01 * dmax, 1.0
    g, mask = gamma.gamma_map(ref, ev, spacing_mm=1.0, dd=0.01, dta_mm=1.0, sub=8)
    exact = abs(c - a * s) / dd / np.sqrt(1.0 + (a * dta / dd) ** 2)
    # interior points (the evaluated grid is clamped at its ends); the search
    # runs over offsets of spacing / 8 within dta, so it over-estimates by at
    # most the gamma change of half an offset step
    inner = mask & (x > 3) & (x < x[-1] - 3)
    bound = exact <= 1.0
    if bound:
        step = 1.0 / 8
        slack = 0.5 * step / dta + 1e-9
        assert np.all(g[inner] >= exact - 1e-9) and np.all(g[inner] <= exact + slack), (g[inner].min(), exact)
    else:
        # outside dta the search stops at |r' - r| = dta: gamma exceeds 1
        assert np.all(g[inner] > 1.0)


def test_ramp_shift_below_and_beyond_the_distance_to_agreement():
    # a steep ramp (10 % of the maximum per mm): a 0.6 mm shift passes by
    # distance to agreement although the dose differs by 6 %; 2 mm fails
    _, ref, ev = _ramp(a=0.1, s=0.6)
    assert gamma.gamma_pass_rate_1d(ref, ev, 1.0) > 0.95
    _, ref, ev = _ramp(a=0.1, s=2.0)
    assert gamma.gamma_pass_rate_1d(ref, ev, 1.0) < 0.1


def test_points_below_the_cutoff_are_not_counted():
    ref = np.concatenate([np.full(20, 0.05), np.full(20, 1.0)])
    ev = ref.copy()
    ev[:20] = 0.09      # 80 % off, but below 10 % of the maximum
    assert gamma.gamma_pass_rate_1d(ref, ev, 1.0) == 1.0
    # a 3 % plateau over bins 22..28: its edge bins find the 1.0 level in
    # reach (at offset o toward the neighbour, gamma^2 = 9 (1 - o)^2 + o^2,
    # minimum 0.9 at o = 0.9 mm); bins 23..27 do not (5 of the 20 counted)
    ev[22:29] = 1.03
    assert gamma.gamma_pass_rate_1d(ref, ev, 1.0) == pytest.approx(15 / 20)


def test_ramp_in_3d_matches_the_closed_form():
    # a ramp along x on 2 mm voxels: a 0.5 mm shift (inside dta, below the
    # voxel size) is found by the 3-D search's 0.5 mm offsets
    a, s = 0.04, 0.5
    z, y, x = np.meshgrid(*(np.arange(16) * 2.0 for _ in range(3)), indexing="ij")
    ref = 1.0 + a * x
    ev = 1.0 + a * (x - s)
    sub = max(2, int(np.ceil(2.0 * 2.0 / 1.0)))
    g, mask = gamma.gamma_map(ref, ev, spacing_mm=2.0, sub=sub)
    inner = mask & (x > 4) & (x < 26)
    dd = 0.01 * ref.max()
    exact = a * s / dd / np.sqrt(1.0 + (a / dd) ** 2)
    step = 2.0 / sub
    assert np.all(g[inner] >= exact - 1e-9) and np.all(g[inner] <= exact + 0.5 * step + 1e-9)
    assert gamma.gamma_pass_rate_3d(ref, ev, 2.0) == 1.0
    ev = 1.0 + a * (x - 3.0)
    assert gamma.gamma_pass_rate_3d(ref, ev, 2.0) < 0.2


def test_shifted_gaussian_in_3d():
    z, y, x = np.meshgrid(*(np.arange(24) * 2.0 for _ in range(3)), indexing="ij")
    def blob(dx):
        return np.exp(-((x - 24.0 - dx) ** 2 + (y - 24.0) ** 2 + (z - 24.0) ** 2) / (2 * 8.0 ** 2))
    ref = blob(0.0)
    assert gamma.gamma_pass_rate_3d(ref, blob(0.5), 2.0) == 1.0
    assert gamma.gamma_pass_rate_3d(ref, blob(3.0), 2.0) < 0.9


def test_empty_reference_is_an_error():
    with pytest.raises(ValueError):
        gamma.gamma_pass_rate_1d(np.zeros(5), np.ones(5), 1.0)
