"""Pins of the oracle's extensions beyond the reference that the benchmark
configurations depend on: the history-batch sum / sum-of-squares estimator,
the rectangular field source and the Woodcock majorant scale.

Each check is something the mathematics fixes, and each is shown to fail on a
planted mistake in the oracle (oracle_lib.fault):

* batches: a run with nbatch batches must equal, batch by batch, independent
  runs over each batch's history range (batch b holds the histories
  ceil(b N / nb) .. ceil((b+1) N / nb) - 1): sum = sum of the runs' tallies,
  sum2 = sum of their squares, voxel by voxel;
* field source: entry points uniform on [x0 - fx/2, x0 + fx/2) x [y0 - fy/2,
  y0 + fy/2) (mean, variance fx^2/12, range, Kolmogorov-Smirnov), on the
  source plane, along +z;
* Woodcock tracking: scaling the majorant by any s >= max(rho/rho_nominal)
  leaves the dose distribution unchanged (only null events are added) while
  the number of photon steps grows in proportion to s (an affine function of
  s: each track's path makes s K steps in the phantom and one out of it).
"""
import numpy as np
import pytest
from scipy import stats

import oracle_lib

RHO3 = np.array([1.0, 4.54, 1.85])   # nominal densities of the configuration-3 tables (water, Ti, bone)


def batch_ranges(n, nb):
    """History ranges of the batches: batch b = [ceil(b n / nb), ceil((b+1) n / nb))."""
    starts = [-(-b * n // nb) for b in range(nb + 1)]
    return [(starts[b], starts[b + 1]) for b in range(nb)]


def test_batch_ranges_partition_by_the_floor_rule():
    # the estimator's rule: history i (0-based within the run) is in batch floor(i nb / n)
    for n, nb in ((4000, 8), (1001, 7), (10, 10), (97, 3)):
        rows = np.array([(i * nb) // n for i in range(n)])
        for b, (lo, hi) in enumerate(batch_ranges(n, nb)):
            assert np.all(rows[lo:hi] == b) and np.sum(rows == b) == hi - lo


def _phantom():
    rng = np.random.default_rng(4)
    mat = np.zeros((16, 14, 14), np.uint8)
    mat[5:9, :, :] = 1
    mat[10:13, 3:10, 2:9] = 2
    rho = (RHO3[mat] * rng.lognormal(0.0, 0.05, mat.shape)).astype(np.float32)
    scale = float(np.max(rho / RHO3[mat]))
    return mat, rho, scale


def _batch_expectation(tables3, n, nb, kw):
    mat, rho, scale = _phantom()
    s = np.zeros(mat.size)
    s2 = np.zeros(mat.size)
    for lo, hi in batch_ranges(n, nb):
        r = oracle_lib.run_voxel(tables3, mat, rho, 2.0, hi - lo, hist_begin=kw["hist_begin"] + lo,
                                 majorant_scale=scale, **{k: v for k, v in kw.items() if k != "hist_begin"})
        d = r["edep"].ravel()
        s += d
        s2 += d * d
    return s, s2


def _batch_run(tables3, n, nb, kw):
    mat, rho, scale = _phantom()
    r = oracle_lib.run_voxel(tables3, mat, rho, 2.0, n, nbatch=nb, majorant_scale=scale, **kw)
    return r["batch_sum"], r["batch_sum2"], r["edep"].ravel()


BATCH_KW = dict(ptype=0, emin=0.5, emax=12.0, field=18.0, seed=31, hist_begin=5000)


def test_batch_sum_and_sum2_equal_per_batch_runs(tables3):
    n, nb = 4000, 8
    s_ref, s2_ref = _batch_expectation(tables3, n, nb, BATCH_KW)
    s, s2, edep = _batch_run(tables3, n, nb, BATCH_KW)
    np.testing.assert_allclose(s, s_ref, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(s2, s2_ref, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(edep, s_ref, rtol=1e-12, atol=1e-12)
    assert np.count_nonzero(s2) > 500


def test_batch_estimator_catches_a_shifted_row_mapping(tables3):
    n, nb = 4000, 8
    s_ref, s2_ref = _batch_expectation(tables3, n, nb, BATCH_KW)
    with oracle_lib.fault(oracle_lib.FAULT_BATCH_ROW):
        s, s2, _ = _batch_run(tables3, n, nb, BATCH_KW)
    # the sum is blind to the mapping; the sum of squares is not
    np.testing.assert_allclose(s, s_ref, rtol=1e-12, atol=1e-12)
    rel = np.abs(s2 - s2_ref) / np.maximum(s2_ref, 1e-300)
    assert rel.max() > 1e-6, rel.max()


def _field_check(p, x0, y0, fx, fy, z0):
    """Failures (a list of text) of the field-entry-point distribution."""
    bad = []
    n = p.shape[0]
    for col, c, w in ((0, x0, fx), (1, y0, fy)):
        v = p[:, col]
        if not (v.min() >= c - 0.5 * w and v.max() < c + 0.5 * w):
            bad.append("range %d" % col)
        if v.max() - v.min() < w * (1 - 1e-3):
            bad.append("span %d" % col)
        if abs(v.mean() - c) > 4 * w / np.sqrt(12 * n):
            bad.append("mean %d" % col)
        # variance of a uniform width w: w^2 / 12 (relative standard error 0.894 / sqrt(n))
        if abs(v.var() / (w * w / 12.0) - 1.0) > 5 * 0.894 / np.sqrt(n):
            bad.append("variance %d" % col)
        if stats.kstest((v - (c - 0.5 * w)) / w, "uniform").pvalue < 1e-4:
            bad.append("ks %d" % col)
    if not np.all(p[:, 2] == z0):
        bad.append("plane")
    if not (np.all(p[:, 3] == 0.0) and np.all(p[:, 4] == 0.0) and np.all(p[:, 5] == 1.0)):
        bad.append("direction")
    return bad


def test_field_source_is_uniform_over_the_field(tables0):
    # a 2 mm phantom 4 voxels wide: nearly every primary starts outside it
    # laterally and leaves at once, so the run costs only the source
    mat = np.zeros((4, 4, 4), np.uint8)
    rho = np.ones((4, 4, 4), np.float32)
    n, x0, y0, fx, fy = 200000, 5.0, -3.0, 30.0, 20.0
    r = oracle_lib.run_voxel(tables0, mat, rho, 2.0, n, ptype=0, e0=2.0, field=fx, field_y=fy, x0=x0, y0=y0,
                             seed=3, primaries=True)
    assert _field_check(r["prim"], x0, y0, fx, fy, -4.0) == []
    np.testing.assert_array_equal(r["prim"][:, 6], 2.0)
    with oracle_lib.fault(oracle_lib.FAULT_FIELD_WIDTH):
        bad = oracle_lib.run_voxel(tables0, mat, rho, 2.0, n, ptype=0, e0=2.0, field=fx, field_y=fy, x0=x0, y0=y0,
                                   seed=3, primaries=True)
    fails = _field_check(bad["prim"], x0, y0, fx, fy, -4.0)
    assert "variance 0" in fails and "variance 1" in fails and "span 0" in fails, fails


def _majorant_runs(tables3, scale_factor, n=3000, runs=10):
    """Depth-dose profiles of `runs` independent runs (seeds 41, 42, ...) of a
    photon field in a phantom with a Ti slab and a bone block, with the
    majorant scaled by scale_factor x max(rho / rho_nominal); returns the
    profiles [runs, nz] and the summed photon-step and Compton counts."""
    rng = np.random.default_rng(12)
    nz, ny, nx = 30, 16, 16
    mat = np.zeros((nz, ny, nx), np.uint8)
    mat[6:10] = 1                # a Ti slab
    mat[14:22, 4:12, 4:12] = 2   # a bone block
    rho = (RHO3[mat] * rng.lognormal(0.0, 0.05, mat.shape)).astype(np.float32)
    s0 = float(np.max(rho / RHO3[mat]))
    prof, g, kn = [], 0, 0
    for k in range(runs):
        r = oracle_lib.run_voxel(tables3, mat, rho, 2.0, n, ptype=0, emin=0.2, emax=3.0, field=16.0, seed=41 + k,
                                 majorant_scale=s0 * scale_factor)
        prof.append(r["edep"].sum(axis=(1, 2)))
        g += r["counters"]["g_steps"]
        kn += r["counters"]["kn"]
    return np.array(prof), g, kn


def _profile_z(a, b):
    """z-scores per depth bin of two sets of independent-run profiles, and of their totals."""
    def mean_se(p):
        return p.mean(axis=0), p.std(axis=0, ddof=1) / np.sqrt(p.shape[0])
    (ma, sa), (mb, sb) = mean_se(a), mean_se(b)
    z = (ma - mb) / np.sqrt(sa ** 2 + sb ** 2)
    ta, tb = a.sum(axis=1), b.sum(axis=1)
    zt = (ta.mean() - tb.mean()) / np.sqrt(ta.var(ddof=1) / ta.size + tb.var(ddof=1) / tb.size)
    return z, zt


def test_majorant_scale_leaves_the_dose_unchanged(tables3):
    a, g1, kn1 = _majorant_runs(tables3, 1.0)
    b, g2, kn2 = _majorant_runs(tables3, 2.0)
    z, zt = _profile_z(a, b)
    # 30 depth bins of two independent estimates (t-like with 9 degrees of freedom each)
    assert np.mean(z * z) < 2.0 and np.max(np.abs(z)) < 5.0 and abs(zt) < 4.0, (np.mean(z * z), z, zt)
    # null events: a photon track of path integral K of the majorant makes on
    # average s K steps that end inside the phantom, plus one that leaves it,
    # so the step count is affine in s: equal increments from s = 1, 2, 3
    _, g3, kn3 = _majorant_runs(tables3, 3.0)
    assert 1.0 < g2 / g1 < 2.0, g2 / g1
    assert abs((g3 - g2) / (g2 - g1) - 1.0) < 0.03, (g1, g2, g3)
    # real interactions do not change
    np.testing.assert_allclose([kn2, kn3], [kn1, kn1], rtol=0.03)


def test_majorant_check_catches_unscaled_interaction_probabilities(tables3):
    a, _, _ = _majorant_runs(tables3, 1.0)
    with oracle_lib.fault(oracle_lib.FAULT_MAJORANT):
        b, _, _ = _majorant_runs(tables3, 2.0)
    z, zt = _profile_z(a, b)
    assert np.mean(z * z) > 10.0 and abs(zt) > 10.0, (np.mean(z * z), zt)
