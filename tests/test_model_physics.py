"""Model-level validation of the reference samplers on the generated tables
(no GPU): the stored histograms of tests/golden/model_hists_cfg3.npz (FP64
oracle samplers, bit-exact to the reference's, 4e6 independent samples per
case) against the analytic cross sections they tabulate:

* Moller: d sigma / d eps = y^2 + (1/eps)(1/eps - g) + (1/(1-eps))(1/(1-eps) - g)
  (MollerTableBuilder.cc:152-168) for eps = T/E in [ecut/E, 1/2];
* Klein-Nishina: d sigma / d eps = (1/eps + eps)(1 - eps sin^2/(1 + eps^2))
  (KNTableBuilder.cc:182-190) for eps in [1/(1+2 kappa), 1];
* Seltzer-Berger: k d sigma / dk proportional to sum_i n_i Z_i^2 chi_i(T, k/T)
  of the NIST scaled DCS the reference ships (SBTables/data/NIST_BREM),
  interpolated as the reference's builder does (log-log in T, natural spline
  in ln(1 - kappa + 1e-12)), for k in [gcut, T], in water and titanium.

A DCS with the Moller exchange term dropped, a Klein-Nishina shape without
its energy dependence and a Bethe-Heitler brem spectrum must fail the same
checks.
"""
import os

import numpy as np
import pytest

import dpm_tablegen as tg

EMC2 = 0.510991
ECUT = 0.2
HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def hists():
    return np.load(os.path.join(HERE, "golden", "model_hists_cfg3.npz"))


def moller_dcs(eps, t, exchange=True):
    gam = t / EMC2 + 1.0
    g = (2.0 * gam - 1.0) / (gam * gam)
    y = (gam - 1.0) / gam
    d = y * y + (1.0 / eps) * (1.0 / eps - g)
    if exchange:
        d = d + (1.0 / (1.0 - eps)) * (1.0 / (1.0 - eps) - g)
    return d


def kn_dcs(eps, e):
    kappa = e / EMC2
    omc = (1.0 / eps - 1.0) / kappa
    sin2 = omc * (2.0 - omc)
    return (1.0 / eps + eps) * (1.0 - eps * sin2 / (1.0 + eps * eps))


def bin_probabilities(edges, pdf_x, lo, hi):
    p = np.zeros(len(edges) - 1)
    for i, (a, b) in enumerate(zip(edges[:-1], edges[1:])):
        a, b = max(a, lo), min(b, hi)
        if b > a:
            x = np.linspace(a, b, 2001)
            p[i] = np.trapezoid(pdf_x(x), x)
    return p / p.sum()


def chi2_per_dof(counts, p, n):
    m = p * n > 1000
    exp = p[m] * n
    return float(np.sum((counts[m] - exp) ** 2 / exp) / m.sum())


def test_moller_sampler_follows_the_moller_dcs(hists):
    e = 10.0
    n = int(hists["n"])
    ed, c = hists["moller_e10_edges"], hists["moller_e10_counts"]
    lo, hi = np.log10(ECUT / e), np.log10(0.5)
    # histogram variable x = log10(eps): pdf_x ~ eps * dsigma/deps
    p = bin_probabilities(ed, lambda x: 10 ** x * moller_dcs(10 ** x, e), lo, hi)
    assert chi2_per_dof(c, p, n) < 2.0
    p_bad = bin_probabilities(ed, lambda x: 10 ** x * moller_dcs(10 ** x, e, exchange=False), lo, hi)
    assert chi2_per_dof(c, p_bad, n) > 20.0


def test_kn_sampler_follows_klein_nishina(hists):
    e = 2.0
    n = int(hists["n"])
    ed, c = hists["compton_e2_edges"], hists["compton_e2_counts"]
    p = bin_probabilities(ed, lambda x: kn_dcs(x, e), 1.0 / (1.0 + 2.0 * e / EMC2), 1.0)
    assert chi2_per_dof(c, p, n) < 2.0
    # Thomson-like mistake (no energy dependence of the shape) fails
    p_bad = bin_probabilities(ed, lambda x: 1.0 + 0.0 * x, 1.0 / (1.0 + 2.0 * e / EMC2), 1.0)
    assert chi2_per_dof(c, p_bad, n) > 20.0


@pytest.mark.parametrize("case,material,e", [("brem_e10_water", "G4_WATER", 10.0), ("brem_e5_ti", "G4_Ti", 5.0)])
def test_sb_sampler_follows_the_nist_scaled_dcs(hists, case, material, e):
    n = int(hists["n"])
    ed, c = hists[case + "_edges"], hists[case + "_counts"]
    m = tg.Material(material)
    _, kappa, _ = tg.sb_data()
    sp = tg.NaturalSpline(tg.sb_phi(kappa), tg.sb_material_dcs(m, e))
    lo, hi = np.log10(0.05 / e), 0.0          # k from gcut = 50 keV to T
    # histogram variable x = log10(k/T): pdf_x ~ k dsigma/dk
    p = bin_probabilities(ed, lambda x: sp(tg.sb_phi(10 ** x)), lo, hi)
    assert chi2_per_dof(c, p, n) < 2.0
    p_bad = bin_probabilities(ed, lambda x: tg.brem_kdsdk(m, e, e * 10 ** x), lo, hi)
    assert chi2_per_dof(c, p_bad, n) > 20.0
