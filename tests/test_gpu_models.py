"""GPU model-level tests: the device form of the reference's BremTest,
MollerTest and MSCAngularDeflectionTest (testAPI.cc:92-199), plus the Compton
final state. dpm_perform runs N single interactions of a primary along +z
(each with its own Philox stream); their histograms must agree with the
stored FP64-oracle histograms of tests/golden/model_hists_cfg3.npz (two-sample
chi-square, independent random numbers), and each final state must satisfy
its kinematic relation sample by sample.
"""
import os

import numpy as np
import pytest
from scipy import stats

import dpm_amd

EMC2 = 0.510991
HERE = os.path.dirname(os.path.abspath(__file__))
CASES = {  # name: (kind, energy, material)
    "brem_e10_water": ("brem", 10.0, 0),
    "brem_e5_ti": ("brem", 5.0, 1),
    "moller_e10": ("moller", 10.0, 0),
    "msc_e5": ("msc", 5.0, 0),
    "msc_e0.5": ("msc", 0.5, 0),
    "compton_e2": ("compton", 2.0, 0),
}


@pytest.fixture(scope="module")
def hists():
    return np.load(os.path.join(HERE, "golden", "model_hists_cfg3.npz"))


def variable(kind, e, v):
    if kind in ("brem", "moller"):
        return np.log10(v / e)
    if kind == "msc":
        return np.log10(np.maximum(v, 1e-300))
    return v / e


def two_sample_p(a, b):
    na, nb = a.sum(), b.sum()
    m = (a + b) >= 50
    ka, kb = np.sqrt(nb / na), np.sqrt(na / nb)
    chi2 = np.sum((ka * a[m] - kb * b[m]) ** 2 / (a[m] + b[m]))
    return float(stats.chi2.sf(chi2, int(m.sum()) - 1))


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [dpm_amd.DPM_FP32, dpm_amd.DPM_FP64])
@pytest.mark.parametrize("name", list(CASES))
def test_final_state_histograms_match_reference_model(gpu_tables3, hists, name, precision):
    kind, e, imat = CASES[name]
    n = 2_000_000
    v, c = gpu_tables3.perform(kind, n, e, imat=imat, seed=77 + list(CASES).index(name), precision=precision)
    x = variable(kind, e, v)
    ed = hists[name + "_edges"]
    mine = np.histogram(x, bins=ed)[0]
    # outside the range: only the exact zero deflection of GS (k2 = 0, 6e-8 per sample)
    assert n - mine.sum() <= 5, (name, n - mine.sum())
    p = two_sample_p(mine, hists[name + "_counts"])
    assert p > 1e-4, (name, p)
    # kinematics of each final state (SimDPMLike.cc)
    if kind == "brem":
        dum0 = EMC2 * 0.7071067812 / (e + EMC2)
        np.testing.assert_allclose(c, 1.0 - dum0 * dum0, rtol=1e-6)
        assert np.all((v > 0.05 * (1 - 1e-6)) & (v <= e * (1 + 1e-6)))
    elif kind == "moller":
        np.testing.assert_allclose(c, np.sqrt(v * (e + 2 * EMC2) / (e * (v + 2 * EMC2))), rtol=1e-6)
        assert np.all((v >= 0.2 * (1 - 1e-6)) & (v <= 0.5 * e * (1 + 1e-6)))
    elif kind == "msc":
        np.testing.assert_allclose(c, 1.0 - v, atol=2e-7)
    else:
        eps = v / e
        np.testing.assert_allclose(c, 1.0 - (1.0 - eps) / (eps * e / EMC2), atol=2e-6)


@pytest.mark.gpu
def test_perform_rejects_impossible_interactions(gpu_tables3):
    with pytest.raises(dpm_amd.DpmError):
        gpu_tables3.perform("moller", 10, 0.3)      # needs E > 2 ecut
    with pytest.raises(dpm_amd.DpmError):
        gpu_tables3.perform("brem", 10, 5.0, imat=9)


@pytest.mark.gpu
@pytest.mark.parametrize("e", [0.06, 0.5, 4.0, 20.0])
def test_compton_directions_stay_real_at_the_kinematic_edge(gpu_tables3, e):
    # FP32 rounding can put cos(theta) of the smallest photon energy
    # (eps = 1/(1 + 2 kappa), backscatter) a few ulp beyond -1; the clamped
    # sine keeps every direction a real unit vector
    n = 2_000_000
    v, c = gpu_tables3.perform("compton", n, e, seed=91)
    eps = v / e
    assert np.all(np.isfinite(v)) and np.all(np.isfinite(c))
    assert np.all(np.abs(c) <= 1.0 + 1e-6)
    edge = eps < (1.0 / (1.0 + 2.0 * e / EMC2)) * (1.0 + 1e-4)
    assert edge.sum() > 0, "no sample near the backscatter edge"
    assert np.all(c[edge] < -0.99)
