"""Phase-space source on the GPU (dpm_simulate_phsp) against the FP64 oracle
and against the sources it generalises.

* FP64 transport of mixed records (e-, e+, photons; positions inside and
  outside the phantom, random directions) equals the oracle history by
  history (a few per cent of the long histories may take another branch on
  the last bit of a transcendental; outside their voxels the doses agree to
  1e-9);
* records restating the pencil beam give the pencil source's histories;
* batch statistics with mixed particle kinds: the batch sums add up to the
  tally and the sum of squares equals that of per-batch runs;
* energy balance of FP32 runs, no termination-guard exits.
"""
import numpy as np
import pytest

import dpm_amd
import oracle_lib
from test_gpu_estimators import _divergent

EMC2 = 0.510991
RHO3 = np.array([1.0, 4.54, 1.85])


def _phantom():
    rng = np.random.default_rng(8)
    mat = np.zeros((20, 16, 14), np.uint8)
    mat[6:9] = 1
    mat[12:17, 3:12, 2:10] = 2
    rho = (RHO3[mat] * rng.lognormal(0.0, 0.05, mat.shape)).astype(np.float32)
    return mat, rho, float(np.max(rho / RHO3[mat]))


def _records(n, seed):
    """Mixed records, photons first (the order of the history ids)."""
    rng = np.random.default_rng(seed)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    kind = rng.choice([-1, 0, 1], n, p=[0.4, 0.4, 0.2])
    pos = np.column_stack([rng.uniform(-20, 20, n), rng.uniform(-22, 22, n), rng.uniform(-28, 28, n)])
    rows = np.column_stack([kind, rng.uniform(1.0, 15.0, n), pos, d])
    return rows[np.argsort(rows[:, 0] != 0, kind="stable")]


@pytest.mark.gpu
def test_phsp_fp64_matches_the_oracle(gpu_tables3, tables3):
    mat, rho, scale = _phantom()
    ph = dpm_amd.Phantom.voxel(gpu_tables3, mat, rho, 2.0)
    n = 3000
    rows = _records(n, 5)
    d, _, st, he = dpm_amd.simulate_phsp(gpu_tables3, ph, rows, seed=9, precision=dpm_amd.DPM_FP64,
                                         per_history=True)
    assert st["guard_exits"] == 0
    ref = oracle_lib.run_voxel(tables3, mat, rho, 2.0, n, seed=9, majorant_scale=scale, per_history=True, phsp=rows)
    np.testing.assert_allclose(st["e_primary"], rows[:, 1].sum(), rtol=1e-12)
    rel = np.abs(he - ref["hist_edep"]) / np.maximum(np.abs(ref["hist_edep"]), 1e-300)
    assert np.mean((rel < 1e-9) | ((he == 0) & (ref["hist_edep"] == 0))) >= 0.97
    bad, foot = _divergent(
        lambda a, k: dpm_amd.simulate_phsp(gpu_tables3, ph, rows[a:a + k], seed=9, hist_begin=a,
                                           precision=dpm_amd.DPM_FP64)[0].ravel(),
        lambda a, k: oracle_lib.run_voxel(tables3, mat, rho, 2.0, k, seed=9, majorant_scale=scale, hist_begin=a,
                                          phsp=rows[a:a + k])["edep"].ravel(),
        0, n)
    assert len(bad) <= 0.03 * n, len(bad)
    m = ~foot
    print("phase space: %d of %d histories differ, %d voxels compared" % (len(bad), n, m.sum()))
    np.testing.assert_allclose(d.ravel()[m], ref["edep"].ravel()[m], rtol=1e-9, atol=1e-12)
    ph.close()


@pytest.mark.gpu
def test_pencil_records_equal_the_pencil_source(gpu_tables0):
    mat, rho = np.zeros((30, 20, 20), np.uint8), np.ones((30, 20, 20), np.float32)
    ph = dpm_amd.Phantom.voxel(gpu_tables0, mat, rho, 2.0)
    n = 20000
    for ptype, e0 in ((-1, 12.0), (0, 8.0)):
        rows = np.zeros((n, 8))
        rows[:, 0], rows[:, 1], rows[:, 2], rows[:, 3], rows[:, 4], rows[:, 7] = ptype, e0, 2.5, -3.0, -30.0, 1.0
        d1, _, s1, h1 = dpm_amd.simulate_phsp(gpu_tables0, ph, rows, seed=3, per_history=True)
        d0, _, s0, h0 = dpm_amd.simulate(gpu_tables0, ph, n, particle=ptype, energy=e0, x0=2.5, y0=-3.0, seed=3,
                                         per_history=True)
        np.testing.assert_allclose(h1, h0, rtol=1e-12, atol=1e-12)
        np.testing.assert_allclose(d1, d0, rtol=1e-9, atol=1e-12)
        assert s1["counters"] == s0["counters"]
    ph.close()


@pytest.mark.gpu
def test_phsp_batches_with_mixed_particles(gpu_tables3):
    mat, rho, _ = _phantom()
    ph = dpm_amd.Phantom.voxel(gpu_tables3, mat, rho, 2.0)
    n, nb = 40000, 4
    rows = _records(n, 11)
    s, s2, st, _ = dpm_amd.simulate_phsp(gpu_tables3, ph, rows, seed=2, nbatch=nb)
    plain, _, st1, _ = dpm_amd.simulate_phsp(gpu_tables3, ph, rows, seed=2)
    np.testing.assert_allclose(s, plain, rtol=1e-9, atol=1e-12)
    es2 = np.zeros(plain.shape)
    for b in range(nb):
        lo, hi = -(-b * n // nb), -(-(b + 1) * n // nb)
        d, _, _, _ = dpm_amd.simulate_phsp(gpu_tables3, ph, rows[lo:hi], seed=2, hist_begin=lo)
        es2 += d ** 2
    np.testing.assert_allclose(s2, es2, rtol=1e-9, atol=1e-12)
    bal = (st1["e_scored"] + st1["e_escaped"]) - (st1["e_primary"] + 2 * EMC2 * (st1["n_annih"] - st1["n_pair"]))
    assert abs(bal) < 1e-5 * st1["e_primary"], bal
    assert st1["guard_exits"] == 0
    ph.close()
