"""GPU parity of the collimated gantry field (point source at sad, jaws with
unscattered transmission exp(-n tau / cos theta)) through the C ABI: the
FP64 engine against the oracle history by history, the FP32 dose against
the FP64 dose, open jaws against the plain field, and the argument checks.
Every run reports zero termination-guard exits."""
import numpy as np
import pytest

import dpm_amd
import gamma
import oracle_lib
from test_gpu_transport import small_heterogeneous

EMC2 = 0.510991
FP64 = dpm_amd.DPM_FP64
COLL = dict(sad=150.0, jaws=(-12.0, 9.0, -7.0, 14.0), jaw_tau=0.8)


def _balance(st):
    return (st["e_scored"] + st["e_escaped"]) - (st["e_primary"] + 2 * EMC2 * (st["n_annih"] - st["n_pair"]))


@pytest.mark.gpu
@pytest.mark.parametrize("particle,energy", [(0, None), (-1, 12.0)])
def test_collimated_field_fp64_history_by_history(gpu_tables3, tables3, particle, energy):
    mat, rho = small_heterogeneous((20, 22, 24), seed=5)       # 40 x 44 x 48 mm at 2 mm
    ph = dpm_amd.Phantom.voxel(gpu_tables3, mat, rho, 2.0)
    scale = float(np.max(rho / np.array([1.0, 4.54, 1.85])[mat]))
    n = 3000
    kw = dict(field=50.0, x0=3.0, y0=-2.0, seed=61, **COLL)
    if energy is None:
        kw.update(emin=0.1, emax=18.0)
        g_kw, o_kw = dict(particle=0, emode=2.0), dict(ptype=0, emode=2.0)
    else:
        g_kw, o_kw = dict(particle=particle, energy=energy), dict(ptype=particle, e0=energy)
    _, _, st, he = dpm_amd.simulate(gpu_tables3, ph, n, angle=35.0, precision=FP64, per_history=True, **g_kw, **kw)
    ref = oracle_lib.run_voxel(tables3, mat, rho, 2.0, n, angle=35.0, per_history=True, majorant_scale=scale,
                               **o_kw, **kw)
    rel = np.abs(he - ref["hist_edep"]) / np.maximum(np.abs(ref["hist_edep"]), 1e-300)
    ok = (rel < 1e-9) | ((np.abs(he) < 1e-300) & (np.abs(ref["hist_edep"]) < 1e-300))
    print("collimated %s: %.4f of histories agree, %d absorbed in the jaws (oracle)"
          % ("photons" if particle == 0 else "e-", ok.mean(), ref["n_jaw"]), flush=True)
    assert ref["n_jaw"] > 0.2 * n
    assert np.mean(ok) >= 0.99, np.mean(ok)
    # the absorbed histories are the same on both sides
    assert np.sum((np.abs(he) < 1e-300) & (np.abs(ref["hist_edep"]) < 1e-300)) >= ref["n_jaw"]
    np.testing.assert_allclose(st["e_primary"], ref["e_primary"], rtol=1e-12)
    assert abs(_balance(st)) < 1e-9 * st["e_primary"]
    assert st["guard_exits"] == 0


@pytest.mark.gpu
def test_collimated_field_fp32_dose_matches_fp64(gpu_tables0):
    mat = np.zeros((48, 48, 48), np.uint8)
    rho = np.ones(mat.shape, np.float32)
    ph = dpm_amd.Phantom.voxel(gpu_tables0, mat, rho, 2.0)
    n = 400000
    kw = dict(particle=0, emin=0.5, emax=6.0, field=60.0, angle=20.0, seed=7, sad=200.0,
              jaws=(-20.0, 15.0, -10.0, 25.0), jaw_tau=2.0)
    d32, _, st32, _ = dpm_amd.simulate(gpu_tables0, ph, n, **kw)
    d64, _, st64, _ = dpm_amd.simulate(gpu_tables0, ph, n, precision=FP64, **kw)
    pr = gamma.gamma_pass_rate_3d(d64, d32, spacing_mm=2.0, dd=0.01, dta_mm=1.0, cutoff=0.1)
    print("collimated field FP32 vs FP64 gamma %.5f" % pr, flush=True)
    assert pr >= 0.99, pr
    assert abs(_balance(st32)) < 1e-5 * st32["e_primary"]
    assert st32["guard_exits"] == 0 and st64["guard_exits"] == 0


@pytest.mark.gpu
def test_open_jaws_equal_the_plain_field(gpu_tables0):
    mat = np.zeros((20, 20, 20), np.uint8)
    rho = np.ones(mat.shape, np.float32)
    ph = dpm_amd.Phantom.voxel(gpu_tables0, mat, rho, 2.0)
    kw = dict(particle=0, energy=6.0, field=30.0, angle=60.0, seed=3, precision=FP64, per_history=True)
    d0, _, _, h0 = dpm_amd.simulate(gpu_tables0, ph, 4000, **kw)
    d1, _, _, h1 = dpm_amd.simulate(gpu_tables0, ph, 4000, jaws=(-16.0, 16.0, -16.0, 16.0), jaw_tau=5.0, **kw)
    np.testing.assert_allclose(h1, h0, rtol=1e-12, atol=0.0)   # a history's secondaries add in any order
    np.testing.assert_allclose(d1, d0, rtol=1e-12, atol=0.0)   # FP64 atomics add in any order


@pytest.mark.gpu
def test_collimator_arguments_are_checked(gpu_tables0):
    mat = np.zeros((20, 20, 20), np.uint8)
    rho = np.ones(mat.shape, np.float32)
    ph = dpm_amd.Phantom.voxel(gpu_tables0, mat, rho, 2.0)      # 40 mm cube, half-diagonal 34.6 mm
    kw = dict(particle=0, energy=6.0, field=30.0, angle=0.0)
    for bad in (dict(sad=30.0), dict(sad=-1.0), dict(jaws=(5.0, -5.0, -5.0, 5.0)),
                dict(jaws=(-5.0, 5.0, -5.0, 5.0), jaw_tau=-1.0), dict(jaws=(-5.0, float("nan"), -5.0, 5.0))):
        with pytest.raises(dpm_amd.DpmError):
            dpm_amd.simulate(gpu_tables0, ph, 10, **kw, **bad)
    d, _, st, _ = dpm_amd.simulate(gpu_tables0, ph, 1000, sad=35.0, **kw)
    assert st["guard_exits"] == 0 and d.sum() > 0.0
    # collimation belongs to gantry fields: a parallel field or pencil with it is refused
    for kind_kw in (dict(field=30.0), dict()):
        for coll in (dict(sad=100.0), dict(jaws=(-5.0, 5.0, -5.0, 5.0))):
            with pytest.raises(dpm_amd.DpmError):
                dpm_amd.simulate(gpu_tables0, ph, 10, particle=0, energy=6.0, **kind_kw, **coll)
