"""Configurations 3 and 5 at their full sizes, checked through properties that
do not depend on the size (the FP64 oracle cannot run them in seconds; the
reduced-size recipes of the same phantoms are compared with it history by
history in test_gpu_transport.py):

* energy conservation of the whole run: scored + escaped energy equals the
  primaries' kinetic energy plus 2 mc^2 per annihilation minus 2 mc^2 per
  pair production;
* batch statistics: the batch sum of squares bounds the squared sum
  (nbatch * sum2 >= sum^2 per voxel, Cauchy-Schwarz) and the batch sums add
  up to the tally;
* FP32 transport against FP64 transport of the same histories: 1 %/1 mm
  gamma pass rate >= 99 % over voxels above 10 % of the maximum dose;
* configuration 5: the mean sampled primary energy equals the mean of the
  triangular (0.1, 2, 18) MeV spectrum within 5 standard errors.
"""
import numpy as np
import pytest

import dpm_amd
import gamma
import phantoms

EMC2 = 0.510991


def _balance(st, e_primary):
    return (st["e_scored"] + st["e_escaped"]) - (e_primary + 2 * EMC2 * (st["n_annih"] - st["n_pair"]))


def _batch_checks(dose, sum2, nbatch):
    # dose is the sum over batches; sum2 the sum of squared batch tallies
    assert np.all(sum2 >= -1e-300)
    assert np.all(nbatch * sum2 * (1 + 1e-9) + 1e-300 >= dose * dose)


@pytest.mark.gpu
def test_config3_full_size(gpu_tables3):
    # 256x256x128 voxels of 2 mm: water with 40 bone and 4 Ti ellipsoids and
    # lognormal(0.05) density jitter; 12 MeV e- 10x10 cm2 field, 1e8 histories
    mat, rho = phantoms.heterogeneous(256, 256, 128, 2.0, seed=1234)
    ph = dpm_amd.Phantom.voxel(gpu_tables3, mat, rho, 2.0)
    n, e0, nbatch = 100_000_000, 12.0, 100
    kw = dict(particle=-1, energy=e0, field=100.0, seed=3, nbatch=nbatch)
    print("config 3 full size: FP32 run of %d histories" % n, flush=True)
    d32, s32, st32, _ = dpm_amd.simulate(gpu_tables3, ph, n, **kw)
    bal = _balance(st32, n * e0)
    assert abs(bal) < 1e-5 * n * e0, bal
    assert abs(d32.sum() - st32["e_scored"]) < 1e-9 * st32["e_scored"]
    _batch_checks(d32, s32, nbatch)
    print("config 3 full size: FP64 run", flush=True)
    d64, _, st64, _ = dpm_amd.simulate(gpu_tables3, ph, n, precision=dpm_amd.DPM_FP64, **kw)
    bal64 = _balance(st64, n * e0)
    assert abs(bal64) < 1e-9 * n * e0, bal64
    pr = gamma.gamma_pass_rate_3d(d64 / rho, d32 / rho, spacing_mm=2.0, dd=0.01, dta_mm=1.0, cutoff=0.1)
    print("config 3 full size: gamma FP32 vs FP64 %.5f" % pr, flush=True)
    assert pr >= 0.99, pr
    ph.close()


@pytest.mark.gpu
def test_config5_full_size(gpu_tables3):
    # 512x512x256 voxels of 1 mm (the heterogeneous recipe, ellipsoids x2);
    # 36 gantry fields every 10 degrees, 10x10 cm2, triangular (0.1, 2, 18)
    # MeV photons, uneven history counts in [1e7, 5e7], 1e9 in total
    mat, rho = phantoms.heterogeneous(512, 512, 256, 1.0, seed=2024, axis_scale=2.0)
    ph = dpm_amd.Phantom.voxel(gpu_tables3, mat, rho, 1.0)
    plan = phantoms.config5_fields(1e9)
    n = sum(f[3] for f in plan)
    assert n == 1_000_000_000 and min(f[3] for f in plan) >= 1e7 and max(f[3] for f in plan) <= 5e7
    kw = dict(particle=0, emin=0.1, emode=2.0, emax=18.0, field=100.0, seed=5)
    print("config 5 full size: FP32 run of %d histories in %d fields" % (n, len(plan)), flush=True)
    d32, s32, st32 = dpm_amd.simulate_fields(gpu_tables3, ph, plan, nbatch=100, **kw)
    bal = _balance(st32, st32["e_primary"])
    assert abs(bal) < 1e-5 * st32["e_primary"], bal
    _batch_checks(d32, s32, 100)
    mean_e = st32["e_primary"] / n
    sd = np.sqrt((0.1 ** 2 + 2.0 ** 2 + 18.0 ** 2 - 0.1 * 2.0 - 0.1 * 18.0 - 2.0 * 18.0) / 18.0)
    assert abs(mean_e - (0.1 + 2.0 + 18.0) / 3) < 5 * sd / np.sqrt(n), mean_e
    # FP32 vs FP64 on a tenth of the plan (same angles and proportions),
    # gamma over the central 128^3 mm^3 where the fields cross
    sub = phantoms.config5_fields(1e8)
    print("config 5 full size: FP32 and FP64 runs of 1e8 histories", flush=True)
    a32, _, _ = dpm_amd.simulate_fields(gpu_tables3, ph, sub, **kw)
    a64, _, st64 = dpm_amd.simulate_fields(gpu_tables3, ph, sub, precision=dpm_amd.DPM_FP64, **kw)
    bal64 = _balance(st64, st64["e_primary"])
    assert abs(bal64) < 1e-9 * st64["e_primary"], bal64
    c = (slice(64, 192), slice(192, 320), slice(192, 320))
    pr = gamma.gamma_pass_rate_3d(a64[c] / rho[c], a32[c] / rho[c], spacing_mm=1.0, dd=0.01, dta_mm=1.0, cutoff=0.1)
    print("config 5 full size: gamma FP32 vs FP64 (central 128^3) %.5f" % pr, flush=True)
    assert pr >= 0.99, pr
    ph.close()


@pytest.mark.gpu
def test_config4_full_size_beamlets(gpu_tables0):
    # 512 beamlets (e- 6-18 MeV, photons U(1, 20) MeV, entry points on the
    # 2 mm grid over 160x160 mm2, lognormal history counts) on the 150^3 water
    # phantom in one batched FP32 run; properties of the CSR dose-influence
    # matrix that hold at any size, and four rows against plain runs
    mat, rho = phantoms.water_cube(150)
    ph = dpm_amd.Phantom.voxel(gpu_tables0, mat, rho, 2.0)
    bl = phantoms.beamlets(512, seed=4321)
    thr = 1e-3
    total = sum(b[4] for b in bl)
    print("config 4 full size: %d beamlets, %d histories" % (len(bl), total), flush=True)
    row_ptr, col, val, st = dpm_amd.simulate_beamlets(gpu_tables0, ph, bl, rel_threshold=thr, seed=8)
    assert st["guard_exits"] == 0
    # energy balance of the whole run
    bal = _balance(st, st["e_primary"])
    assert abs(bal) < 1e-5 * st["e_primary"], bal
    np.testing.assert_allclose(st["e_primary"], sum(b[3] * b[4] for b in bl), rtol=1e-6)
    # CSR structure: rows in order, voxel indices valid and unique per row,
    # every kept value above the threshold of its row's maximum
    assert row_ptr[0] == 0 and np.all(np.diff(row_ptr) >= 1) and row_ptr[-1] == col.size
    assert col.min() >= 0 and col.max() < ph.nbins and np.all(val > 0)
    csr_total = 0.0
    for b in range(len(bl)):
        c, v = col[row_ptr[b]:row_ptr[b + 1]], val[row_ptr[b]:row_ptr[b + 1]]
        assert np.unique(c).size == c.size
        assert v.min() >= thr * v.max() * (1 - 1e-12)
        csr_total += v.sum()
    # the thresholded rows hold part of the scored energy, never more
    assert csr_total <= st["e_scored"] * (1 + 1e-9)
    assert csr_total > 0.5 * st["e_scored"]
    # the same beamlets in 4 groups of 128 dense rows: the same CSR
    g4 = dpm_amd.simulate_beamlets(gpu_tables0, ph, bl, rel_threshold=thr, seed=8, tally_budget=128 * 8 * ph.nbins)
    assert g4[3]["beamlet_groups"] == 4
    np.testing.assert_array_equal(g4[0], row_ptr)
    np.testing.assert_array_equal(g4[1], col)
    np.testing.assert_allclose(g4[2], val, rtol=1e-12)
    del g4
    # four rows (the largest, the smallest, an electron and a photon beamlet)
    # against plain per-beamlet runs of the same history ids
    off = np.concatenate([[0], np.cumsum([b[4] for b in bl])])
    counts = np.array([b[4] for b in bl])
    picks = {int(np.argmax(counts)), int(np.argmin(counts)),
             next(i for i, b in enumerate(bl) if b[2] == -1), next(i for i, b in enumerate(bl) if b[2] == 0)}
    for b in sorted(picks):
        x, y, p, e, n = bl[b]
        d, _, _, _ = dpm_amd.simulate(gpu_tables0, ph, n, particle=p, energy=e, x0=x, y0=y, seed=8,
                                      hist_begin=int(off[b]))
        d = d.ravel()
        keep = np.nonzero((d > thr * d.max()) & (d > 0))[0]
        c, v = col[row_ptr[b]:row_ptr[b + 1]], val[row_ptr[b]:row_ptr[b + 1]]
        o = np.argsort(c)
        np.testing.assert_array_equal(c[o], keep)
        np.testing.assert_allclose(v[o], d[keep], rtol=1e-9)
        print("config 4: beamlet %d (%s %.1f MeV, %d histories) equals its plain run: %d voxels"
              % (b, "e-" if p == -1 else "gamma", e, n, keep.size), flush=True)
    ph.close()


@pytest.mark.gpu
def test_config4_beamlets_on_the_config5_phantom(gpu_tables3):
    # the 512 beamlets of configuration 4 on the 512x512x256 phantom at 1 mm
    # (67 M voxels, 537 MB per dense FP64 row): the default 16 GiB budget runs
    # them in 16 groups of 32 rows; energy balance, CSR structure and two rows
    # against plain runs
    mat, rho = phantoms.heterogeneous(512, 512, 256, 1.0, seed=2024, axis_scale=2.0)
    ph = dpm_amd.Phantom.voxel(gpu_tables3, mat, rho, 1.0)
    del mat
    bl = phantoms.beamlets(512, seed=4321)
    thr = 1e-3
    print("config 4 beamlets on the config-5 phantom: %d histories" % sum(b[4] for b in bl), flush=True)
    row_ptr, col, val, st = dpm_amd.simulate_beamlets(gpu_tables3, ph, bl, rel_threshold=thr, seed=8)
    per_group = (16 << 30) // (8 * ph.nbins)
    assert st["beamlet_groups"] == -(-512 // per_group), st["beamlet_groups"]
    assert st["guard_exits"] == 0
    bal = _balance(st, st["e_primary"])
    assert abs(bal) < 1e-5 * st["e_primary"], bal
    assert row_ptr[0] == 0 and np.all(np.diff(row_ptr) >= 1) and row_ptr[-1] == col.size
    assert col.min() >= 0 and col.max() < ph.nbins and np.all(val > 0)
    assert val.sum() <= st["e_scored"] * (1 + 1e-9)
    off = np.concatenate([[0], np.cumsum([b[4] for b in bl])])
    for b in (0, per_group + 1):   # in the first and in the second group
        x, y, p, e, n = bl[b]
        d, _, _, _ = dpm_amd.simulate(gpu_tables3, ph, n, particle=p, energy=e, x0=x, y0=y, seed=8,
                                      hist_begin=int(off[b]))
        d = d.ravel()
        keep = np.nonzero((d > thr * d.max()) & (d > 0))[0]
        c, v = col[row_ptr[b]:row_ptr[b + 1]], val[row_ptr[b]:row_ptr[b + 1]]
        o = np.argsort(c)
        np.testing.assert_array_equal(c[o], keep)
        np.testing.assert_allclose(v[o], d[keep], rtol=1e-9)
    print("config 4 on the config-5 phantom: %d groups, %d nonzeros, %.2f s device" %
          (st["beamlet_groups"], col.size, st["seconds"]), flush=True)
    ph.close()
