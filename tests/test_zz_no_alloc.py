"""dpm_simulate_noalloc (the entry point that allocates no device memory)
against dpm_simulate: the same histories give the same tally, to the order of
FP64 atomic additions (1e-9 relative per bin), and the same event counters;
two calls in a row, the second with a smaller input, so that the second
finds the first one's contents in the shared device scratch. Device arrays
for the results come from torch (the entry point writes device memory)."""
import numpy as np
import pytest

import dpm_amd

torch = pytest.importorskip("torch")


def _dev(n):
    return torch.zeros(n, dtype=torch.float64, device="cuda")


def _check_against_plain(t, ph, n, kw, precision=dpm_amd.DPM_FP32):
    ref, ref_steps, st_ref, _ = dpm_amd.simulate(t, ph, n, precision=precision, **kw)
    d_tally, d_steps = _dev(ph.nbins), _dev(ph.nbins)
    torch.cuda.synchronize()
    slab = ph.shape is None
    st = dpm_amd.simulate_noalloc(t, ph, n, d_tally.data_ptr(), d_steps.data_ptr() if slab else None,
                                  precision=precision, **kw)
    torch.cuda.synchronize()
    got = d_tally.cpu().numpy()
    np.testing.assert_allclose(got, np.asarray(ref).ravel(), rtol=1e-9, atol=1e-12)
    if slab:
        np.testing.assert_array_equal(d_steps.cpu().numpy(), np.asarray(ref_steps).ravel())
    for k in ("sb", "moller", "kn", "gs", "secondaries", "deposits", "e_steps", "g_steps"):
        assert st["counters"][k] == st_ref["counters"][k], k
    np.testing.assert_allclose(st["e_scored"], st_ref["e_scored"], rtol=1e-9)
    np.testing.assert_allclose(st["e_escaped"], st_ref["e_escaped"], rtol=1e-9, atol=1e-9)
    assert st["guard_exits"] == 0
    return st


@pytest.mark.gpu
def test_noalloc_equals_simulate_twice_second_smaller(gpu_tables0, gpu_tables3):
    # voxel phantom, photon field: several scratch-sized chunks, then a smaller run
    mat, rho = np.zeros((40, 40, 40), np.uint8), np.ones((40, 40, 40), np.float32)
    ph = dpm_amd.Phantom.voxel(gpu_tables0, mat, rho, 2.0)
    kw = dict(particle=0, emin=0.5, emax=20.0, field=60.0, seed=5)
    st = _check_against_plain(gpu_tables0, ph, 200000, kw)
    assert st["rounds"] >= 8
    _check_against_plain(gpu_tables0, ph, 7000, dict(kw, seed=6))
    # FP64 transport through the same scratch
    _check_against_plain(gpu_tables0, ph, 30000, dict(kw, seed=7), precision=dpm_amd.DPM_FP64)
    ph.close()
    # the reference's slab and pencil beam (depth tally and step counts)
    slab = dpm_amd.Phantom.slab(gpu_tables3, 1.0, 3)
    _check_against_plain(gpu_tables3, slab, 50000, dict(particle=-1, energy=15.0, seed=9))
    _check_against_plain(gpu_tables3, slab, 3000, dict(particle=-1, energy=15.0, seed=10))
    slab.close()


@pytest.mark.gpu
def test_noalloc_rejects_what_needs_allocation(gpu_tables0):
    mat, rho = np.zeros((8, 8, 8), np.uint8), np.ones((8, 8, 8), np.float32)
    ph = dpm_amd.Phantom.voxel(gpu_tables0, mat, rho, 2.0)
    d = _dev(ph.nbins)
    torch.cuda.synchronize()
    L = dpm_amd.lib()
    src = dpm_amd._source(0, 5.0, 0.0, 0.0, 0.0, 0.0, 0.0, None, None)
    for opt in (dpm_amd.RunOptions(seed=1, nbatch=10), dpm_amd.RunOptions(seed=1, pipes=2),
                dpm_amd.RunOptions(seed=1, graph=1), dpm_amd.RunOptions(seed=1, chunk=1 << 30),
                dpm_amd.RunOptions(seed=1, waves_per_cu=64)):
        rc = L.dpm_simulate_noalloc(gpu_tables0.h, ph.h, src, 1000, opt, dpm_amd.ctypes.c_void_p(d.data_ptr()),
                                    None, None)
        assert rc == -1, rc
    assert L.dpm_simulate_noalloc(gpu_tables0.h, ph.h, src, 1000, None, None, None, None) == -1
    ph.close()
