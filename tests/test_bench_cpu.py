"""Host-side helpers of bench.py (no GPU): the per-voxel FP32 error report and
the hist.sim reader used by the configuration-1 leg."""
import importlib.util
import os

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def bench():
    spec = importlib.util.spec_from_file_location("bench_module", os.path.join(ROOT, "bench.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


def test_voxel_error_report_statistics(bench):
    # 100 batch values per voxel with mean 1 and sigma 0.1: the total has
    # sigma 0.1 * sqrt(100) = 1, i.e. 1 % of the total 100. An "FP64" dose
    # that differs from the FP32 total by N(0, 1) per voxel must come out at
    # the half-normal's median 0.674 sigma, with 68.3 % of voxels within 1 sigma.
    rng = np.random.default_rng(11)
    nb, nv = 100, 20000
    x = rng.normal(1.0, 0.1, size=(nb, nv))
    t = x.sum(0)
    s2 = (x * x).sum(0)
    d64 = t + rng.normal(0.0, 1.0, nv)
    r = bench.voxel_error_report(t, s2, d64, nb, cutoff=0.0)
    assert r["voxels"] == nv
    assert abs(r["stat_rel_sigma_median"] - 0.01) < 5e-4
    assert abs(r["err_over_sigma_median"] - 0.6745) < 0.03
    assert abs(r["frac_err_below_sigma"] - 0.6827) < 0.02
    assert abs(r["rel_err_median"] - 0.6745 * 0.01) < 5e-4
    # identical doses: zero error
    z = bench.voxel_error_report(t, s2, t.copy(), nb, cutoff=0.0)
    assert z["rel_err_max"] == 0.0 and z["frac_err_below_sigma"] == 1.0
    # the cutoff keeps only voxels above the fraction of the FP64 maximum
    d = t.copy()
    d[: nv // 2] *= 0.01
    assert bench.voxel_error_report(d, s2, d, nb, cutoff=0.1)["voxels"] == nv - nv // 2


def test_read_hist_sim_reference_fixture(bench):
    # the compiled reference's own output (10^4 histories, configuration 3 slab)
    mean, rows = bench._read_hist_sim(os.path.join(ROOT, "tests", "golden", "ref_hist_cfg3_e15_n10000.sim"))
    assert abs(mean - 13.624) < 1e-9
    assert rows.shape[1] == 3 and rows.shape[0] >= 100
    np.testing.assert_allclose(rows[:, 0], 0.05 + 0.1 * np.arange(rows.shape[0]), rtol=1e-4)
    assert rows[:, 1].max() > 0 and np.all(rows[:, 2] >= 0)


def test_host_cpus_records_the_cores_used(bench):
    h = bench.host_cpus()
    assert h["affinity_cpus"] == len(os.sched_getaffinity(0))
    assert 1 <= h["workers"] <= min(bench.CPU_SHARE, h["affinity_cpus"])
    assert "physical_cores" in h or "lscpu_error" in h


def test_equal_history_report(bench):
    import gamma
    rng = np.random.default_rng(3)
    d = rng.uniform(0.0, 1.0, size=(10, 12, 14))
    rho = np.full(d.shape, 2.0)
    r = bench.equal_history_report(gamma, d, d.copy(), rho, 2.0)
    assert r["gamma_pass_rate_gpu_fp32_vs_cpu"] == 1.0 and r["max_abs_diff_above_cutoff_rel_to_max"] == 0.0
    assert r["voxels_compared"] == int((d >= 0.1 * d.max()).sum())
    # a flat dose 3 % high fails every voxel at 1 %/1 mm (gamma = 3)
    flat = np.ones(d.shape)
    bad = bench.equal_history_report(gamma, flat, flat * 1.03, rho, 2.0)
    assert bad["gamma_pass_rate_gpu_fp32_vs_cpu"] == 0.0
    assert abs(bad["total_energy_rel_diff"] - 0.03) < 1e-12


def test_more_than_one_gpu_is_refused():
    # one GPU only: the command exits non-zero with a message before touching a device
    import subprocess
    import sys
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode != 0
    assert "--gpus 2" in r.stderr and r.stdout.strip() == ""
