"""Checks of the Geant4-free table generator: the alias tables sample the
tabulated linear pdf exactly, the GS RITA tables invert their distribution,
and key physics values sit near published data (NIST XCOM / ESTAR)."""
import numpy as np

import dpm_tablegen as tg


def _alias_bin_probabilities(w, alias):
    """Exact probability of each bin under the run-time rule
    (SimLinAliasData.cc:27-30): bin i with 1/(n-1), kept with prob w_i."""
    nb = len(w) - 1
    p = np.zeros(nb)
    for i in range(nb):
        keep = min(max(w[i], 0.0), 1.0)
        p[i] += keep / nb
        p[alias[i]] += (1.0 - keep) / nb
    return p


def test_alias_tables_reproduce_trapezoid_masses():
    rng = np.random.default_rng(1)
    for n in (3, 4, 8, 64):
        x = np.sort(rng.uniform(0, 1, n))
        x[0], x[-1] = 0.0, 1.0
        y = rng.uniform(0.0, 5.0, n)
        w, al = tg.linear_alias_table(x, y)
        mass = 0.5 * (x[1:] - x[:-1]) * (y[1:] + y[:-1])
        np.testing.assert_allclose(_alias_bin_probabilities(w, al), mass / mass.sum(), atol=1e-12)


def test_alias_sampling_matches_pdf_by_brute_force():
    # sample the full run-time algorithm on a lattice of uniforms and compare
    # the empirical cdf with the exact one of the linear pdf
    x = np.array([0.0, 0.2, 0.5, 1.0])
    y = np.array([1.0, 3.0, 0.5, 2.0])
    w, al = tg.linear_alias_table(x, y)
    m = 400
    r1 = (np.arange(m) + 0.5) / m
    r2 = (np.arange(m) + 0.5) / m
    vals = []
    for a in r1:
        rest = a * 3
        i = int(rest)
        if w[i] < rest - i:
            i = al[i]
        for b in r2:
            xv, xd = x[i], x[i + 1] - x[i]
            dum = (y[i + 1] - y[i]) / y[i]
            if abs(dum) > 0.1:
                vals.append(xv - xd / dum * (1.0 - np.sqrt(1.0 + b * dum * (dum + 2.0))))
            else:
                vals.append(xv + b * xd * (1.0 - 0.5 * dum * (b - 1.0) * (1.0 + dum * b)))
    vals = np.sort(vals)
    grid = np.linspace(0, 1, 101)
    # exact cdf of the piecewise-linear pdf
    cdf = np.concatenate([[0.0], np.cumsum(0.5 * (x[1:] - x[:-1]) * (y[1:] + y[:-1]))])
    cdf /= cdf[-1]

    def exact(t):
        j = min(np.searchsorted(x, t, side="right") - 1, 2)
        h = t - x[j]
        s = (y[j + 1] - y[j]) / (x[j + 1] - x[j])
        area = cdf[j] * 1.0
        tot = np.sum(0.5 * (x[1:] - x[:-1]) * (y[1:] + y[:-1]))
        return area + (y[j] * h + 0.5 * s * h * h) / tot

    emp = np.searchsorted(vals, grid, side="right") / vals.size
    ex = np.array([exact(t) for t in grid])
    assert np.max(np.abs(emp - ex)) < 5e-3


def test_screened_rutherford_moments():
    # Legendre moments of 1/(1-mu+2A)^2 from the Q_l recurrence equal direct quadrature
    a = np.array([1e-3, 1e-2, 0.3])
    omg = tg.sr_legendre_moments(a, 6)
    mu, wq = np.polynomial.legendre.leggauss(4000)
    for j, aa in enumerate(a):
        f = 1.0 / (1.0 - mu + 2 * aa) ** 2
        norm = np.sum(wq * f)
        for l in range(1, 7):
            pl = np.polynomial.legendre.legval(mu, [0] * l + [1])
            gl = np.sum(wq * f * pl) / norm
            assert abs((1 - gl) - omg[l, j]) < 2e-6 * max(1.0, omg[l, j])


def test_gs_rita_tables_invert_the_distribution():
    nel = np.array([50.0, 400.0])
    scr = np.array([1e-3, 1e-4])
    a_tr, tabs = tg.gs_tables(nel, scr, ncum=64, nu=1024)
    for i in range(2):
        u = tabs[i, :, 0]
        assert u[0] == 0.0 and u[-1] == 1.0 and np.all(np.diff(u) >= 0)
        # the rational interpolation is monotone within each cumulative bin
        d = 1.0 / 63
        for j in range(63):
            a, b = tabs[i, j, 1], tabs[i, j, 2]
            r = np.linspace(0, d, 50)
            uu = u[j] + (1 + a + b) * d * r / (d * d + a * d * r + b * r * r) * (u[j + 1] - u[j])
            assert np.all(np.diff(uu) >= -1e-12)


def test_physics_values_near_published_data():
    w = tg.Material("G4_WATER")
    c, p, q = tg.photon_mxsec(w, np.array([1.0, 10.0]))
    mu = (c + p + q) * 10.0  # cm2/g
    assert abs(mu[0] - 0.0707) / 0.0707 < 0.02      # XCOM, 1 MeV (no coherent)
    assert abs(mu[1] - 0.0222) / 0.0222 < 0.03      # XCOM, 10 MeV
    s = tg.restricted_collision_dedx(w, np.array([1.0, 10.0]), np.array([1.0, 10.0]))
    assert abs(s[0] - 1.849) / 1.849 < 0.02         # ESTAR collision stopping power
    assert abs(s[1] - 1.968) / 1.968 < 0.02
