#!/usr/bin/env python3
"""Geant4-free physics-table generator for the DPM transport engine.

The reference produces its run-time tables with ``dpm_GenerateData`` on top of
Geant4 (``dpm_GenerateData.cc:52-128``, ``DataInit/``); Geant4 and its data sets
are not available here, so this module computes the same *quantities* from
published closed-form physics and writes them in the reference's on-disk
formats, byte-layout compatible with the run-time readers in
``Simulation/src/Sim*.cc``:

================================  =========================================  ==============================
file                              format of (reference writer)               read by (reference reader)
================================  =========================================  ==============================
mat.dat                           DataInit/inc/ElectronData.hh:444-481       SimMaterialData.cc:6-53
el_iemfp.dat (never read)         DataInit/inc/ElectronData.hh:483-498       --
el_itr1mfp.dat                    DataInit/inc/ElectronData.hh:500-517       SimITr1MFPElastic.cc:13-56
el_screeningPar.dat (never read)  DataInit/inc/ElectronData.hh:519-534       --
el_scatStrength.dat               DataInit/inc/ElectronData.hh:536-551       SimMaxScatStrength.cc:6-43
el_GSDtrData.dat                  DataInit/inc/ElectronData.hh:553-586       SimGSTables.cc:19-80
eloss_rdedx.dat                   DataInit/inc/ElectronData.hh:588-603       SimStoppingPower.cc:18-61
eloss_rRange.dat (never read)     DataInit/inc/ElectronData.hh:605-619       --
imfp_brem.dat                     DataInit/inc/ElectronData.hh:621-637       SimIMFPBrem.cc:18-61
imfp_moller.dat                   DataInit/inc/ElectronData.hh:639-654       SimIMFPMoller.cc:6-43
ioni_MollerDtrData.dat            DataInit/src/MollerTableBuilder.cc:120-149 SimMollerTables.cc:20-67
brem_SBDtrData.dat                DataInit/src/SBTableBuilder.cc:296-344     SimSBTables.cc:21-90
imfp_{compton,pairp,photoe,total} DataInit/inc/PhotonData.hh:174-243         SimIMFPPhoton.cc:19-63
imfp_globalMax.dat                DataInit/inc/PhotonData.hh:246-257         SimIMFPMaxPhoton.cc:6-38
compton_KNDtrData.dat             DataInit/src/KNTableBuilder.cc:149-178     SimKNTables.cc:20-67
================================  =========================================  ==============================

Energy grids follow the reference generator: 128 log-spaced e- energies on
[0.99 ecut, Emax] (``dpm_GenerateData.cc:91-96``, ``ElectronData.hh:161-180``),
brem IMFP on [max(ecut,gcut), Emax] and Moller IMFP on [2 ecut, Emax]
(``InitElectronData.cc:254-285``), SB/Moller alias tables 128 x 64, KN 32 x 64,
GS 256 x 64 on [ecut, Emax], photon data 1024 *linear* points on
[0.99 gcut, Emax] (``PhotonData.hh:60-71``).

Physics models (written here from the cited formulae and data, not from
Geant4):

* elastic: screened-Rutherford (Wentzel) DCS per element with Moliere's
  screening parameter; exact transport coefficients of that DCS.
* collision stopping power: Berger-Seltzer restricted formula (ICRU 37) with
  Sternheimer's density-effect parameters (tabulated x0, x1, m, delta0 of
  Sternheimer, Berger & Seltzer 1984; general Sternheimer-Peierls rules for
  other materials).
* bremsstrahlung: the Seltzer-Berger scaled DCS (the reference's
  SBTables/data/NIST_BREM values of the elements used, kept in
  sb_nist_brem.npz): sampling tables by the reference's own construction
  (SBTableBuilder.cc:69-241), and the brem cross section and restricted
  radiative stopping power as integrals of the same spline DCS. The
  Bethe-Heitler closed form (Koch & Motz 1959, 3BN limits) is kept as a
  cross-check.
* Moller: Moller DCS and its closed-form restricted integrals.
* photons: Klein-Nishina (Compton), Bethe-Heitler pair with a threshold
  factor, and a Z^4.6 E^-3.5 photo-absorption fit normalised to water.
* MSC: Goudsmit-Saunderson distribution of the screened-Rutherford DCS with the
  Legendre moments from Q_l(1+2A) (Miller backward recurrence), the q2+
  (two-or-more scattering) part, transformed with mu(u)=1-2au/(1-u+a) and
  tabulated for rational-interpolation (RITA) inversion.

The tables are therefore close to, but not, Geant4's (the reference test
case's depth dose agrees with the Geant4-table run at 1 %/1 mm); GPU-vs-CPU
parity is checked on identical files written by this module.
"""
import argparse
import math
import os
import sys

import numpy as np

# --------------------------------------------------------------------------
# constants (CODATA 2018 rounded; MeV, cm)
# --------------------------------------------------------------------------
EMC2 = 0.51099895          # electron rest energy [MeV]
R_E = 2.8179403262e-13     # classical electron radius [cm]
ALPHA = 1.0 / 137.035999084
N_A = 6.02214076e23
HBARC = 197.3269804e-13    # [MeV cm]
BOHR = 0.529177210903e-8   # [cm]

# Z, atomic weight [g/mol], mean excitation energy [eV]
ELEMENTS = {
    "H": (1, 1.00794, 19.2), "C": (6, 12.0107, 81.0), "N": (7, 14.0067, 82.0),
    "O": (8, 15.9994, 95.0), "Na": (11, 22.98977, 149.0), "Mg": (12, 24.305, 156.0),
    "P": (15, 30.973761, 173.0), "S": (16, 32.065, 180.0), "Ar": (18, 39.948, 188.0),
    "Ca": (20, 40.078, 191.0), "Ti": (22, 47.867, 233.0), "W": (74, 183.84, 727.0),
}

# name -> (density g/cm3, material I [eV], {element: mass fraction})
MATERIALS = {
    "G4_WATER": (1.0, 78.0, {"H": 0.111894, "O": 0.888106}),
    "G4_AIR": (0.00120479, 85.7, {"C": 0.000124, "N": 0.755268, "O": 0.231781, "Ar": 0.012827}),
    "G4_BONE_CORTICAL_ICRP": (1.92, 110.0, {"H": 0.034, "C": 0.155, "N": 0.042, "O": 0.435, "Na": 0.001,
                                            "Mg": 0.002, "P": 0.103, "S": 0.003, "Ca": 0.225}),
    "G4_Ti": (4.54, 233.0, {"Ti": 1.0}),
    "G4_BONE_COMPACT_ICRU": (1.85, 91.9, {"H": 0.064, "C": 0.278, "N": 0.027, "O": 0.41, "Mg": 0.002,
                                          "P": 0.07, "S": 0.002, "Ca": 0.147}),
    "G4_W": (19.3, 727.0, {"W": 1.0}),
}


class Config:
    """Mirror of the reference ``Configuration`` defaults and pre-defined
    indices (``Configuration.hh:50-113``)."""

    def __init__(self, index=0):
        self.n_ekin_el = 128
        self.n_ekin_ph = 1024
        self.emin = 0.001
        self.emax = 21.0
        self.ecut = 0.20
        self.gcut = 0.05
        self.s_low = 5.0
        self.s_high = 10.0
        self.e_cross = 12.0
        self.geom_index = 0
        self.materials = ["G4_WATER"]
        if index == 1:
            self.geom_index = 1
            self.materials = ["G4_WATER", "G4_AIR"]
        elif index == 2:
            self.geom_index = 2
            self.materials = ["G4_WATER", "G4_BONE_CORTICAL_ICRP"]
        elif index == 3:
            self.geom_index = 3
            self.materials = ["G4_WATER", "G4_Ti", "G4_BONE_COMPACT_ICRU"]
        elif index == 4:
            self.s_low, self.s_high = 2.0, 5.0
            self.materials = ["G4_Ti"]
        elif index == 5:
            self.s_low, self.s_high = 1.0, 1.0
            self.materials = ["G4_W"]


# --------------------------------------------------------------------------
# material description
# --------------------------------------------------------------------------
class Material:
    def __init__(self, name):
        rho, iev, comp = MATERIALS[name]
        self.name = name
        self.rho = rho
        self.I = iev * 1e-6
        z = np.array([ELEMENTS[e][0] for e in comp], dtype=float)
        a = np.array([ELEMENTS[e][1] for e in comp], dtype=float)
        w = np.array([comp[e] for e in comp], dtype=float)
        w = w / w.sum()
        atoms = w / a                       # mol atoms per gram per element
        self.Z = z
        self.A = a
        self.n = atoms / atoms.sum()        # fractional number of atoms
        self.atoms_per_gram = atoms * N_A   # per element
        self.Zbar = float(np.sum(self.n * z))
        self.Z2bar = float(np.sum(self.n * z * z))
        self.Zsbar = float(np.sum(self.n * z * (z + 1.0)))
        self.Abar = float(np.sum(self.n * a))
        self.ZAbar = float(np.sum(self.n * a * z))
        self.Imean_el = float(np.exp(np.sum(self.n * np.log([ELEMENTS[e][2] * 1e-6 for e in comp]))))
        self.inel_corr = self.Zsbar / self.Z2bar
        self.electrons_per_gram = float(np.sum(self.atoms_per_gram * z))


def log_grid(e0, e1, n):
    g = np.exp(np.log(e0) + np.arange(n) * (np.log(e1 / e0) / (n - 1)))
    g[0], g[-1] = e0, e1
    return g


def kinematics(t):
    pc = np.sqrt(t * (t + 2.0 * EMC2))
    etot = t + EMC2
    beta = pc / etot
    return pc, beta, etot


# --------------------------------------------------------------------------
# elastic scattering: screened Rutherford with Moliere screening
# --------------------------------------------------------------------------
def moliere_screening(z, t):
    pc, beta, _ = kinematics(t)
    a_tf = 0.88534 * BOHR * z ** (-1.0 / 3.0)
    return (HBARC / (2.0 * pc * a_tf)) ** 2 * (1.13 + 3.76 * (ALPHA * z / beta) ** 2)


def sr_transport_coef(a):
    """1-G_1 of the screened Rutherford DCS with screening parameter a."""
    return 2.0 * a * ((1.0 + a) * np.log1p(1.0 / a) - 1.0)


def elastic_data(mat, t):
    """Per-gram elastic and first-transport inverse mfps [1/cm per g/cm3]
    (nuclear Z^2 only; the inelastic correction is applied by the caller)."""
    pc, beta, _ = kinematics(t)
    pbc = pc * beta
    sig_el = np.zeros_like(t)
    sig_tr = np.zeros_like(t)
    for z, npg in zip(mat.Z, mat.atoms_per_gram):
        a = moliere_screening(z, t)
        s = math.pi * (z * R_E * EMC2 / pbc) ** 2 / (a * (1.0 + a))
        sig_el += npg * s
        sig_tr += npg * s * sr_transport_coef(a)
    return sig_el, sig_tr


def find_screening(g1):
    """Screening parameter A with 1-G_1(A) = g1 (bisection in log A)."""
    g1 = np.atleast_1d(np.asarray(g1, dtype=float))
    lo = np.full_like(g1, -40.0)
    hi = np.full_like(g1, 12.0)
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        val = sr_transport_coef(np.exp(mid))
        up = val < g1
        lo = np.where(up, mid, lo)
        hi = np.where(up, hi, mid)
    return np.exp(0.5 * (lo + hi))


# --------------------------------------------------------------------------
# stopping power
# --------------------------------------------------------------------------
# Sternheimer density-effect parameters (Sternheimer, Berger & Seltzer, At.
# Data Nucl. Data Tables 30 (1984) 261): the table's I [eV] and Cbar (for the
# table's own I and density; density_effect recomputes Cbar for the
# material's), x0, x1, m, delta0
STERNHEIMER = {
    "G4_WATER": (75.0, 3.5017, 0.2400, 2.8004, 3.4773, 0.0),
    "G4_AIR": (85.7, 10.5961, 1.7418, 4.2759, 3.3994, 0.0),
    "G4_BONE_CORTICAL_ICRP": (106.4, 3.6488, 0.1161, 3.0919, 3.5919, 0.0),
    "G4_Ti": (233.0, 4.4450, 0.0957, 3.0386, 3.0679, 0.12),
    "G4_BONE_COMPACT_ICRU": (91.9, 3.3390, 0.0944, 3.0201, 3.6419, 0.0),
    "G4_W": (727.0, 5.4059, 0.2167, 3.4960, 2.8447, 0.14),
}


def plasma_energy(mat):
    """hbar omega_p = 28.816 eV sqrt(rho Z/A) [MeV]."""
    return 28.816e-6 * math.sqrt(mat.rho * mat.Zbar / mat.Abar)


def density_effect(mat, t):
    """delta(beta gamma) of the Sternheimer parametrisation: 0 (insulators) or
    delta0 10^(2(x - x0)) (conductors) below x0 = log10(p/mc), 4.6052 x - Cbar
    + a (x1 - x)^m up to x1, 4.6052 x - Cbar above, with Cbar = 1 + 2
    ln(I/hbar omega_p) of the material's own I and density and a from
    continuity at x0. Tabulated materials use their published x0, x1, m and
    delta0; the others the general Sternheimer-Peierls rules."""
    pc, _, _ = kinematics(t)
    x = np.log10(pc / EMC2)
    cbar = 1.0 + 2.0 * math.log(mat.I / plasma_energy(mat))
    if mat.name in STERNHEIMER:
        _, _, x0, x1, m, d0 = STERNHEIMER[mat.name]
    else:
        d0, m = 0.0, 3.0
        if mat.rho < 0.01:  # gases
            x0, x1 = (1.6, 4.0) if cbar < 10.0 else (0.326 * cbar - 1.5, 4.0)
        elif mat.I < 100e-6:
            x0, x1 = (0.2, 2.0) if cbar < 3.681 else (0.326 * cbar - 1.0, 2.0)
        else:
            x0, x1 = (0.2, 3.0) if cbar < 5.215 else (0.326 * cbar - 1.5, 3.0)
    a = (cbar - 4.6052 * x0 + d0) / (x1 - x0) ** m
    d = np.where(x < x0, d0 * 10.0 ** (2.0 * (x - x0)), 4.6052 * x - cbar + a * np.clip(x1 - x, 0.0, None) ** m)
    return np.where(x >= x1, 4.6052 * x - cbar, d)


def restricted_collision_dedx(mat, t, ecut):
    """Berger-Seltzer restricted collision stopping power [MeV cm2/g]."""
    tau = t / EMC2
    gam = tau + 1.0
    beta2 = 1.0 - 1.0 / (gam * gam)
    dmax = np.minimum(ecut, 0.5 * t) / EMC2
    i_r = mat.I / EMC2
    f = (-1.0 - beta2 + np.log((tau - dmax) * dmax) + tau / (tau - dmax)
         + (0.5 * dmax * dmax + (2.0 * tau + 1.0) * np.log(1.0 - dmax / tau)) / (gam * gam))
    dedx = (2.0 * math.pi * R_E ** 2 * EMC2 * mat.electrons_per_gram / beta2
            * (np.log(2.0 * (tau + 2.0) / (i_r * i_r)) + f - density_effect(mat, t)))
    return np.maximum(dedx, 0.0)


# --------------------------------------------------------------------------
# bremsstrahlung
# --------------------------------------------------------------------------
def brem_kdsdk(mat, t, k):
    """k dsigma/dk per gram [cm2/g] (Bethe-Heitler, screening by the smaller
    logarithm). t, k broadcast together, 0 < k <= t. A cross-check of the
    Seltzer-Berger values the tables are built from."""
    e0 = t + EMC2
    e = e0 - k
    v = e / e0
    bracket = 1.0 + v * v - (2.0 / 3.0) * v
    total = 0.0
    for z, npg in zip(mat.Z, mat.atoms_per_gram):
        l_unscr = np.log(2.0 * e0 * e / (EMC2 * k)) - 0.5
        l_scr = math.log(183.0 * z ** (-1.0 / 3.0))
        lg = np.clip(np.minimum(l_unscr, l_scr), 0.0, None)
        total = total + npg * 4.0 * ALPHA * R_E ** 2 * z * (z + 1.0) * (bracket * lg + np.where(l_unscr > l_scr, v / 9.0, 0.0))
    return total


def _gl(n):
    x, w = np.polynomial.legendre.leggauss(n)
    return 0.5 * (x + 1.0), 0.5 * w


def brem_integrals(mat, t, gcut):
    """Restricted radiative stopping power (k<gcut) [MeV cm2/g] and brem
    macroscopic cross section (k>gcut) [1/cm per g/cm3]."""
    x, w = _gl(48)
    t = np.asarray(t, dtype=float)
    kc = np.minimum(gcut, t)
    # int_0^kc k dsigma/dk dk  (integrand smooth in k)
    kk = kc[:, None] * x[None, :]
    srad = np.sum(w[None, :] * brem_kdsdk(mat, t[:, None], kk), axis=1) * kc
    # int_gcut^t dsigma/dk dk = int (k dsigma/dk) dln k
    lo = np.log(gcut)
    hi = np.log(np.maximum(t, gcut * (1.0 + 1e-12)))
    lk = lo + (hi - lo)[:, None] * x[None, :]
    sig = np.sum(w[None, :] * brem_kdsdk(mat, t[:, None], np.exp(lk)), axis=1) * (hi - lo)
    sig = np.where(t > gcut, sig, 0.0)
    return srad, sig


# Seltzer-Berger scaled DCS chi(Z, T, kappa) = (beta^2/Z^2) k dsigma/dk [mb]
# of the ELEMENTS, on the NIST grid of 113 kinetic energies x 32 reduced
# photon energies kappa = k/T (written by make_sb_data.py from the reference's
# SBTables/data/NIST_BREM files)
SB_DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "sb_nist_brem.npz")
MBARN_CM2 = 1.0e-27
_SB = None


def sb_data():
    """(energy[113] MeV, kappa[32], {Z: chi[113, 32] in mb})."""
    global _SB
    if _SB is None:
        with np.load(SB_DATA, allow_pickle=False) as d:
            _SB = (d["energy"].copy(), d["kappa"].copy(),
                   {int(z): d["chi"][i].copy() for i, z in enumerate(d["Z"])})
    return _SB


class NaturalSpline:
    """Natural cubic spline (zero second derivative at both ends) through
    (x_i, y_i) with strictly monotone x, increasing or decreasing.

    Evaluation follows Utils/src/Spline.cc:73-108: the segment is found by
    bisection and clamped to the end segments outside [x_0, x_{n-1}], and a
    segment whose two node values sum to zero gives zero. The coefficients are
    the textbook ones (second derivatives from the tridiagonal continuity
    system of the first derivative, solved by forward elimination)."""

    def __init__(self, x, y):
        x = np.asarray(x, dtype=float)
        y = np.asarray(y, dtype=float)
        n = len(x)
        if n < 2 or len(y) != n:
            raise ValueError("spline needs >= 2 matching nodes")
        h = np.diff(x)
        if not (np.all(h > 0) or np.all(h < 0)):
            raise ValueError("spline nodes must be strictly monotone")
        s = np.diff(y) / h
        m2 = np.zeros(n)
        if n > 2:
            diag = 2.0 * (h[:-1] + h[1:])
            rhs = 6.0 * (s[1:] - s[:-1])
            for i in range(1, n - 2):
                f = h[i] / diag[i - 1]
                diag[i] -= f * h[i]
                rhs[i] -= f * rhs[i - 1]
            m2[n - 2] = rhs[-1] / diag[-1]
            for i in range(n - 4, -1, -1):
                m2[i + 1] = (rhs[i] - h[i + 1] * m2[i + 2]) / diag[i]
        self.x, self.y, self.m2 = x, y, m2
        self.b = s - h * (2.0 * m2[:-1] + m2[1:]) / 6.0
        self.c = 0.5 * m2[:-1]
        self.d = (m2[1:] - m2[:-1]) / (6.0 * h)
        self.increasing = h[0] > 0

    def segment(self, v):
        """Index m of the segment that holds v (clamped to 0 .. n-2)."""
        x, n = self.x, len(self.x)
        v = np.asarray(v, dtype=float)
        if self.increasing:
            m = np.searchsorted(x[:-1], v, side="left")
            m = np.where(m > 0, m - 1, 0)
        else:
            m = n - np.searchsorted(x[::-1], v, side="right") - 1
        return np.clip(m, 0, n - 2)

    def __call__(self, v):
        m = self.segment(v)
        t = np.asarray(v, dtype=float) - self.x[m]
        val = self.y[m] + t * (self.b[m] + t * (self.c[m] + t * self.d[m]))
        return np.where(self.y[m] + self.y[m + 1] == 0.0, 0.0, val)


def sb_phi(kappa):
    """phi(kappa) = ln(1 - kappa + 1e-12), the spline variable of the scaled
    DCS (SBTableBuilder.cc:146-148): it spreads the tip kappa -> 1."""
    return np.log(1.0 - np.asarray(kappa, dtype=float) + 1.0e-12)


def sb_material_dcs(mat, t, sb=None):
    """X_j = sum_i n_i Z_i^2 chi_i(T, kappa_j) [cm^2/g] at the 32 kappa nodes,
    n_i in atoms per gram. chi_i is interpolated linearly in (ln T, ln chi)
    between the grid energies E_e <= T < E_e+1, per kappa node
    (SBTableBuilder.cc:131-160)."""
    energy, kappa, chi = sb or sb_data()
    e = int(np.searchsorted(energy, t, side="right")) - 1
    if not 0 <= e < len(energy) - 1:
        raise ValueError("T = %g MeV outside the scaled-DCS grid" % t)
    dum0 = math.log(t / energy[e])
    dum1 = math.log(energy[e + 1] / energy[e])
    x = np.zeros(len(kappa))
    for z, npg in zip(mat.Z, mat.atoms_per_gram):
        lo = chi[int(round(z))][e] * MBARN_CM2
        hi = chi[int(round(z))][e + 1] * MBARN_CM2
        x += np.exp(np.log(hi / lo) / dum1 * dum0 + np.log(lo)) * npg * z * z
    return x


def sb_alias_points(mat, gcut, t, ns=64, sb=None):
    """The ns (xi, p(xi)) points of one SB sampling table at kinetic energy T,
    restating SBTableBuilder::BuildOneLinAlias (SBTableBuilder.cc:120-241):
    xi(kappa) = 1 - ln(kappa)/ln(kappa_c), kappa_c = gcut/T, spans the photon
    energies gcut..T; the points are xi = 0 (spline value at kappa_c), the
    loaded nodes kappa_c < kappa_k < 1 with their node values, xi = 1 with the
    kappa = 1 value, then, until ns points, the midpoint of the interval whose
    linear interpolation is furthest from the spline, with its spline value.
    p is in the reference's per-volume unit (atoms/mm^3 x mm^2)."""
    sb = sb or sb_data()
    kappa = sb[1]
    nk = len(kappa)
    dcs = sb_material_dcs(mat, t, sb) * 0.1 * mat.rho
    sp = NaturalSpline(sb_phi(kappa), dcs)
    kc = gcut / t
    lkc = math.log(kc)
    ki = 0
    while ki < nk - 1 and kc >= kappa[ki]:
        ki += 1
    ki -= 1
    xs = [0.0]
    ys = [float(sp(sb_phi(kc)))]
    for k in range(ki + 1, nk - 1):
        xs.append(1.0 - math.log(kappa[k]) / lkc)
        ys.append(float(dcs[k]))
    xs.append(1.0)
    ys.append(float(dcs[nk - 1]))
    while len(xs) < ns:
        xa, ya = np.array(xs), np.array(ys)
        xm = 0.5 * (xa[:-1] + xa[1:])
        lin = 0.5 * (ya[:-1] + ya[1:])
        spv = sp(sb_phi(kc * np.exp(-xm * lkc)))
        j = int(np.argmax(np.abs(lin - spv)))
        xs.insert(j + 1, float(xm[j]))
        ys.insert(j + 1, float(spv[j]))
    return np.array(xs), np.array(ys)


def sb_brem_integrals(mat, t, gcut, sb=None, ngl=16):
    """Restricted radiative stopping power S_rad(T) = int_0^min(gcut,T) k
    dsigma/dk dk [MeV cm^2/g] and the brem cross section sigma(T) =
    int_gcut^T dsigma/dk dk [cm^2/g] from the scaled DCS: per atom
    k dsigma/dk = Z^2 chi(kappa)/beta^2, so with X the spline of
    sb_material_dcs in phi(kappa) and kappa_c = min(gcut/T, 1),
    S_rad = T/beta^2 int_0^kappa_c X dkappa, sigma = 1/beta^2
    int_kappa_c^1 X dln(kappa). Gauss-Legendre (ngl points) on every interval
    between the kappa nodes."""
    sb = sb or sb_data()
    kappa = sb[1]
    x, w = _gl(ngl)
    t = np.atleast_1d(np.asarray(t, dtype=float))
    srad = np.zeros(len(t))
    sig = np.zeros(len(t))
    for i, ti in enumerate(t):
        sp = NaturalSpline(sb_phi(kappa), sb_material_dcs(mat, ti, sb))
        beta2 = ti * (ti + 2.0 * EMC2) / (ti + EMC2) ** 2
        kc = min(gcut / ti, 1.0)
        nodes = np.concatenate([[0.0], kappa[(kappa > 0.0) & (kappa < kc)], [kc]])
        a, b = nodes[:-1, None], nodes[1:, None]
        kk = a + (b - a) * x[None, :]
        srad[i] = np.sum((b - a) * w[None, :] * sp(sb_phi(kk))) * ti / beta2
        if kc < 1.0:
            nodes = np.log(np.concatenate([[kc], kappa[(kappa > kc) & (kappa < 1.0)], [1.0]]))
            a, b = nodes[:-1, None], nodes[1:, None]
            uu = a + (b - a) * x[None, :]
            sig[i] = np.sum((b - a) * w[None, :] * sp(sb_phi(np.exp(uu)))) / beta2
    return srad, sig


# --------------------------------------------------------------------------
# Moller
# --------------------------------------------------------------------------
def moller_xsec(mat, t, ecut):
    """Restricted Moller macroscopic cross section [1/cm per g/cm3] for
    secondary energies in [ecut, t/2]."""
    tau = t / EMC2
    gam = tau + 1.0
    beta2 = 1.0 - 1.0 / (gam * gam)
    xc = ecut / t
    g = (2.0 * gam - 1.0) / (gam * gam)
    y = (gam - 1.0) / gam
    val = (y * y * (0.5 - xc) + 1.0 / xc - 1.0 / (1.0 - xc) - g * np.log((1.0 - xc) / xc))
    s = 2.0 * math.pi * R_E ** 2 * EMC2 / (beta2 * t) * val
    return np.where(t > 2.0 * ecut, mat.electrons_per_gram * s, 0.0)


def moller_pdf_xi(xi, ecut, t):
    """eps dsigma/deps of the Moller DCS in the variable
    xi = ln(T/ecut)/ln(t/(2 ecut)) used by SimMollerTables.cc:83-88."""
    tsec = ecut * np.exp(xi * np.log(t / (2.0 * ecut)))
    eps = tsec / t
    gam = t / EMC2 + 1.0
    g = (2.0 * gam - 1.0) / (gam * gam)
    y = (gam - 1.0) / gam
    dcs = y * y + (1.0 / eps) * (1.0 / eps - g) + (1.0 / (1.0 - eps)) * (1.0 / (1.0 - eps) - g)
    return eps * dcs


def kn_pdf_xi(xi, e):
    """eps dsigma/deps of the Klein-Nishina DCS in the variable
    xi with eps = exp(ln(1+2kappa)(xi-1)) (SimKNTables.cc:85-89)."""
    kappa = e / EMC2
    eps = np.exp(np.log1p(2.0 * kappa) * (xi - 1.0))
    one_minus_cos = (1.0 / eps - 1.0) / kappa
    sin2 = one_minus_cos * (2.0 - one_minus_cos)
    return eps * (1.0 / eps + eps) * (1.0 - eps * sin2 / (1.0 + eps * eps))


# --------------------------------------------------------------------------
# photons
# --------------------------------------------------------------------------
def kn_total(e):
    k = e / EMC2
    l2k = np.log1p(2.0 * k)
    return 2.0 * math.pi * R_E ** 2 * (((1.0 + k) / (k * k)) * (2.0 * (1.0 + k) / (1.0 + 2.0 * k) - l2k / k)
                                       + l2k / (2.0 * k) - (1.0 + 3.0 * k) / (1.0 + 2.0 * k) ** 2)


def pair_xsec_atom(z, e):
    k = e / EMC2
    with np.errstate(divide="ignore", invalid="ignore"):
        unscr = (28.0 / 9.0) * np.log(2.0 * k) - 218.0 / 27.0
        scr = (28.0 / 9.0) * math.log(183.0 * z ** (-1.0 / 3.0)) - 2.0 / 27.0
        thr = (2.0 * math.pi / 3.0) * ((k - 2.0) / k) ** 3
        g = np.minimum(np.maximum(unscr, thr), scr)
    g = np.where(k > 2.0, g, 0.0)
    return ALPHA * R_E ** 2 * z * (z + 1.0) * g


_PE_C = None


def photoel_xsec_atom(z, e):
    """Z^4.6 E^-3.5 photo-absorption per atom, normalised so that water has
    mu_pe/rho = 0.00276 cm2/g at 100 keV."""
    global _PE_C
    if _PE_C is None:
        water = Material("G4_WATER")
        s = np.sum(water.atoms_per_gram * water.Z ** 4.6)
        _PE_C = 0.00276 / s
    return _PE_C * z ** 4.6 * (e / 0.1) ** (-3.5)


def photon_mxsec(mat, e):
    """(compton, pair, photoelectric) macroscopic cross sections [1/mm] at the
    nominal density."""
    comp = mat.electrons_per_gram * kn_total(e)
    pair = np.zeros_like(e)
    pe = np.zeros_like(e)
    for z, npg in zip(mat.Z, mat.atoms_per_gram):
        pair += npg * pair_xsec_atom(z, e)
        pe += npg * photoel_xsec_atom(z, e)
    f = mat.rho * 0.1
    return comp * f, pair * f, pe * f


# --------------------------------------------------------------------------
# alias tables (Walker) for the linear-pdf samplers
# --------------------------------------------------------------------------
def linear_alias_table(x, y):
    """Walker alias table over the len(x)-1 trapezoid bins of the linearly
    interpolated pdf (x_i, y_i); the run-time rule is: take bin i with
    probability w_i, else bin alias_i (SimLinAliasData.cc:25-30)."""
    nb = len(x) - 1
    mass = 0.5 * (x[1:] - x[:-1]) * (y[1:] + y[:-1])
    mass = np.maximum(mass, 1e-300)
    q = mass * nb / mass.sum()
    w = np.ones(len(x))
    alias = np.arange(len(x), dtype=np.int64)
    small = [i for i in range(nb) if q[i] < 1.0]
    large = [i for i in range(nb) if q[i] >= 1.0]
    while small and large:
        s = small.pop()
        l = large[-1]
        w[s] = q[s]
        alias[s] = l
        q[l] -= 1.0 - q[s]
        if q[l] < 1.0:
            large.pop()
            small.append(l)
    for i in small + large:
        w[i] = 1.0
        alias[i] = i
    return w, alias


# --------------------------------------------------------------------------
# Goudsmit-Saunderson angular distributions
# --------------------------------------------------------------------------
def sr_legendre_moments(a_scr, lmax):
    """1-G_l (l=0..lmax) of the screened Rutherford DCS 1/(1-mu+2A)^2 for an
    array of screening parameters, from G_l = l (Q_{l-1}(x) - x Q_l(x)) with
    x = 1+2A and Q_l by Miller's backward recurrence normalised to Q_0."""
    a_scr = np.asarray(a_scr, dtype=float)
    x = 1.0 + 2.0 * a_scr
    eta = np.arccosh(x)
    lstart = int(lmax + 40.0 / float(np.min(eta)) + 50)
    q_next = np.zeros_like(x)
    q_cur = np.full_like(x, 1e-280)
    qs = np.zeros((lmax + 1, x.size))
    for l in range(lstart, 0, -1):
        q_prev = ((2 * l + 1) * x * q_cur - (l + 1) * q_next) / l
        q_next, q_cur = q_cur, q_prev
        if l - 1 <= lmax:
            qs[l - 1] = q_cur
        big = np.abs(q_cur) > 1e200
        if np.any(big):
            sc = np.where(big, 1e-200, 1.0)
            q_cur *= sc
            q_next *= sc
            qs *= sc[None, :]
    q0 = 0.5 * np.log((x + 1.0) / (x - 1.0))
    qs *= (q0 / qs[0])[None, :]
    one_minus_g = np.zeros((lmax + 1, x.size))
    ls = np.arange(1, lmax + 1)[:, None]
    g = ls * (qs[:-1] - x[None, :] * qs[1:])
    one_minus_g[1:] = 1.0 - g
    one_minus_g[1] = sr_transport_coef(a_scr)   # closed form for l=1
    return np.maximum(one_minus_g, 0.0)


def gs_q2plus_coefs(nel, one_minus_g):
    """Legendre coefficients xi_l of the two-or-more-scattering part of the GS
    distribution."""
    en = np.exp(-nel)
    xi = np.exp(-nel * one_minus_g) - en - en * nel * (1.0 - one_minus_g)
    return xi / (1.0 - en - nel * en)


def optimal_transform(mean_one_minus_mu):
    """Transformation parameter a such that the screened-Rutherford pdf with
    screening a has the same <1-mu> as the MSC distribution (then
    mu(u)=1-2au/(1-u+a) maps that pdf to the uniform one)."""
    target = np.clip(mean_one_minus_mu, 1e-14, 0.999999)
    return find_screening(target)


def gs_tables(nel, a_scr, ncum=64, nu=1024):
    """Build the (u, a, b) RITA inversion tables for the transformed GS q2+
    pdfs at each (N_el, screening) pair. Returns (transform_param[n],
    tables[n, ncum, 3])."""
    nel = np.asarray(nel, dtype=float)
    a_scr = np.asarray(a_scr, dtype=float)
    n = nel.size
    # truncation of the Legendre sums: xi_l < 1e-13 relative
    lmax = 64
    while True:
        omg = sr_legendre_moments(a_scr, lmax)
        xi = gs_q2plus_coefs(nel[None, :], omg)
        if np.all(np.abs(xi[-1]) < 1e-13) or lmax >= 20000:
            break
        lmax *= 2
    xi = gs_q2plus_coefs(nel[None, :], omg)
    mean_1mmu = 1.0 - xi[1]
    a_tr = optimal_transform(mean_1mmu)
    u = np.linspace(0.0, 1.0, nu)
    kd = 1.0 / (1.0 - u[None, :] + a_tr[:, None])
    mu = 1.0 - 2.0 * a_tr[:, None] * u[None, :] * kd
    # Legendre sum F(mu) = sum (l+1/2) xi_l P_l(mu)
    p0 = np.ones_like(mu)
    p1 = mu.copy()
    f = 0.5 * xi[0][:, None] * p0 + 1.5 * xi[1][:, None] * p1
    for l in range(1, lmax):
        p2 = ((2 * l + 1) * mu * p1 - l * p0) / (l + 1)
        f += (l + 1.5) * xi[l + 1][:, None] * p2
        p0, p1 = p1, p2
    q = f * 2.0 * a_tr[:, None] * (1.0 + a_tr[:, None]) * kd * kd
    q = np.maximum(q, 1e-12 * np.max(q, axis=1, keepdims=True))
    tables = np.zeros((n, ncum, 3))
    dc = 1.0 / (ncum - 1)
    for i in range(n):
        qi = q[i]
        cum = np.concatenate([[0.0], np.cumsum(0.5 * (qi[1:] + qi[:-1]) * np.diff(u))])
        qi = qi / cum[-1]
        cum = cum / cum[-1]
        targets = np.arange(ncum) * dc
        idx = np.clip(np.searchsorted(cum, targets, side="right") - 1, 0, nu - 2)
        # invert the linear pdf within the dense bin
        ul = u[idx]
        h = u[idx + 1] - ul
        pa = qi[idx]
        pb = qi[idx + 1]
        r = targets - cum[idx]
        slope = (pb - pa) / h
        with np.errstate(divide="ignore", invalid="ignore"):
            disc = np.sqrt(np.maximum(pa * pa + 2.0 * slope * r, 0.0))
            t_lin = np.where(np.abs(slope) > 1e-12 * np.maximum(pa, 1e-300), (disc - pa) / slope, r / np.maximum(pa, 1e-300))
        uu = np.clip(ul + t_lin, 0.0, 1.0)
        uu[0], uu[-1] = 0.0, 1.0
        uu = np.maximum.accumulate(uu)
        pu = np.interp(uu, u, qi)
        du = np.diff(uu)
        with np.errstate(divide="ignore", invalid="ignore"):
            d1 = dc / du
            b = 1.0 - d1 * d1 / (pu[:-1] * pu[1:])
            a = d1 / pu[:-1] - 1.0 - b
        bad = ~np.isfinite(a) | ~np.isfinite(b) | (b > 0.99) | (np.abs(a) > 50.0) | (np.abs(b) > 50.0)
        a = np.where(bad, 0.0, a)
        b = np.where(bad, 0.0, b)
        tables[i, :, 0] = uu
        tables[i, :-1, 1] = a
        tables[i, :-1, 2] = b
    return a_tr, tables


# --------------------------------------------------------------------------
# driver
# --------------------------------------------------------------------------
def generate(outdir, index, verbose=True):
    cfg = Config(index)
    mats = [Material(m) for m in cfg.materials]
    nmat = len(mats)
    os.makedirs(outdir, exist_ok=True)
    ecut, gcut, emax = cfg.ecut, cfg.gcut, cfg.emax
    egrid = log_grid(0.99 * ecut, emax, cfg.n_ekin_el)
    ref = mats[0]

    def log(msg):
        if verbose:
            print(msg, flush=True)

    log(f" === dpm_tablegen: configuration {index} -> {outdir} ({', '.join(cfg.materials)})")
    # ---- per material electron quantities on the common grid
    per = []
    for m in mats:
        sig_el, sig_tr = elastic_data(m, egrid)         # 1/cm per g/cm3
        lam_el = 1.0 / (sig_el * m.rho) * 10.0          # mm
        lam_tr = 1.0 / (sig_tr * m.rho) * 10.0          # mm
        scr = find_screening(sig_tr / sig_el)
        srad, _ = sb_brem_integrals(m, egrid, gcut)
        dedx = restricted_collision_dedx(m, egrid, ecut) + srad   # MeV cm2/g
        fx = np.arctan(5.0 * np.log(egrid / cfg.e_cross)) / math.pi + 0.5
        smax = cfg.s_low * (cfg.s_high / cfg.s_low) ** fx
        k1 = m.inel_corr * smax / lam_tr
        per.append(dict(lam_el=lam_el, lam_tr=lam_tr, scr=scr, dedx=dedx, k1=k1))

    with open(os.path.join(outdir, "mat.dat"), "w") as f:
        f.write("# Secondary electron and gamma production thresholds:\n %.8E %.8E\n" % (ecut, gcut))
        f.write("# MSC maximum step length parameters:\n %.8E %.8E %.8E\n" % (cfg.s_low, cfg.s_high, cfg.e_cross))
        f.write("# Number of materials:\n %d\n" % nmat)
        for im, m in enumerate(mats):
            f.write("# \n")
            f.write("# Material index, name and density [g/cm3] data:\n")
            f.write(" %d %16s %.8E\n" % (im, m.name, m.rho))
            f.write("# Z, Z^2, Z(Z+1), w [amu], wZ [amu], I [MeV] (all are weighted for molecules)\n")
            f.write(" %.8E  %.8E  %.8E  %.8E  %.8E  %.8E\n" % (m.Zbar, m.Z2bar, m.Zsbar, m.Abar, m.ZAbar, m.I))
            f.write("# [Z/w]/[Z/w]_ref, [Z^2/w]/[Z^2/w]_ref (all are weighted for molecules) ref: im=0\n")
            f.write(" %.8E  %.8E\n" % (m.Zbar / m.Abar * ref.Abar / ref.Zbar, m.Z2bar / m.Abar * ref.Abar / ref.Z2bar))

    with open(os.path.join(outdir, "el_iemfp.dat"), "w") as f:
        f.write("# Inverse elastic mean free path, i.e. macroscopic cross section, data\n")
        f.write("# The `_scpc` are the scattering power corrected, i.e. inelastic scattering correected, data) \n")
        f.write("# E [MeV]   1/Lambda_elastic [1/mm] 1/Lambda_elastic_scpc [1/mm]  \n")
        f.write("# number of data per materials\n %d\n" % len(egrid))
        f.write("# -----------------------------\n")
        for im, m in enumerate(mats):
            f.write("#\n# Material name: %s\n#\n" % m.name)
            for ie, e in enumerate(egrid):
                emfp = per[im]["lam_el"][ie]
                f.write("%.8E  %.8E  %.8E\n" % (e, 1.0 / emfp, m.inel_corr / emfp))

    with open(os.path.join(outdir, "el_itr1mfp.dat"), "w") as f:
        f.write("# Inverse 1-rst transport mean free path, i.e. macroscopic cross section, data\n")
        f.write("# The `scpc` are the scattering power corrected, i.e. inelastic scattering correected, data) \n")
        f.write("# E [MeV]   1/Lambda_1 [1/mm][cm3/g] 1/Lambda_1^{scpc} [1/mm][cm3/g]  \n")
        f.write("# number of data per materials and #material\n %d %d\n" % (len(egrid), nmat))
        f.write("# -----------------------------\n")
        for im, m in enumerate(mats):
            f.write("#\n# Material name: %s\n#\n" % m.name)
            for ie, e in enumerate(egrid):
                tr = per[im]["lam_tr"][ie]
                f.write("%.8E  %.8E  %.8E\n" % (e, 1.0 / tr / m.rho, m.inel_corr / tr / m.rho))

    with open(os.path.join(outdir, "el_screeningPar.dat"), "w") as f:
        f.write("# Screening parameter of the screened Rutherford DCS\n")
        f.write("# (i.e. \\eta(E_i), such that the accurate G_1(\\eta) = mfp_el/mfp_tr1) \n")
        f.write("# E [MeV]   Screening_parameter [] \n")
        f.write("# number of data per materials\n %d\n" % len(egrid))
        f.write("# -----------------------------\n")
        for im, m in enumerate(mats):
            f.write("#\n# Material name: %s\n#\n" % m.name)
            for ie, e in enumerate(egrid):
                f.write("%.8E  %.8E\n" % (e, per[im]["scr"][ie]))

    with open(os.path.join(outdir, "el_scatStrength.dat"), "w") as f:
        f.write("# The max scattering strength `K_1(E)` allowed in a given MSC step.\n")
        f.write("# Max K_1(E) is approximated as max-MSC-step-length(E)/tr1-mfp(E).\n")
        f.write("# E [MeV]   Max-scattering-strength [] \n")
        f.write("# number of data per materials\n %d\n" % len(egrid))
        f.write("# -----------------------------\n")
        for im, m in enumerate(mats):
            f.write("#\n# Material name: %s\n#\n" % m.name)
            for ie, e in enumerate(egrid):
                f.write("%.8E  %.8E\n" % (e, per[im]["k1"][ie]))

    # ---- GS tables
    log("     GS angular tables ...")
    gs_e = log_grid(ecut, emax, 256)
    with open(os.path.join(outdir, "el_GSDtrData.dat"), "w") as f:
        f.write("# Data related to the transformed GS angular distributions:\n")
        f.write("# (optimal transformation parameters `a` and tables for fast,\n")
        f.write("#  rejection free sampling from the q(u;s/emfp,scrPar,a) pdf-s)\n")
        f.write("# #ekin, emin [MeV], emax [MeV] and #cumulative:\n")
        f.write("%d  %.14E  %.14E  %d\n" % (len(gs_e), gs_e[0], gs_e[-1], 64))
        f.write("# For each materials: #ekin tables and 3x#cumulative data in each table,\n")
        f.write("# i, E_i [MeV] and transformation parameter are given before each table.\n")
        f.write("# -----------------------------\n")
        for im, m in enumerate(mats):
            lg = np.log(egrid)
            k1 = np.interp(np.log(gs_e), lg, per[im]["k1"])
            inv_tr = np.interp(np.log(gs_e), lg, 1.0 / per[im]["lam_tr"])
            inv_el = np.interp(np.log(gs_e), lg, 1.0 / per[im]["lam_el"])
            g1 = inv_tr / inv_el
            nel = k1 / g1
            scr = find_screening(g1)
            a_tr, tabs = gs_tables(nel, scr)
            f.write("#\n# Material name: %s\n#\n" % m.name)
            for ie in range(len(gs_e)):
                f.write("%d  %.14E  %.14E\n" % (ie, gs_e[ie], a_tr[ie]))
                for iu in range(64):
                    f.write("%22.14E  %22.14E  %22.14E " % (tabs[ie, iu, 0], tabs[ie, iu, 1], tabs[ie, iu, 2]))
                    if (iu + 1) % 2 == 0 or iu == 63:
                        f.write("\n")

    # ---- energy loss
    with open(os.path.join(outdir, "eloss_rdedx.dat"), "w") as f:
        # the reference writer fuses its first two comment lines (no newline)
        f.write("# The restricted, total (collisonal and radiative) stopping power")
        f.write("# E [MeV]   dE/dx [MeV/mm][cm3/g] \n")
        f.write("# number of data per materials and #material\n %d %d\n" % (len(egrid), nmat))
        f.write("# -----------------------------\n")
        for im, m in enumerate(mats):
            f.write("#\n# Material name: %s\n#\n" % m.name)
            for ie, e in enumerate(egrid):
                f.write("%.8E  %.8E\n" % (e, per[im]["dedx"][ie] * 0.1))

    with open(os.path.join(outdir, "eloss_rRange.dat"), "w") as f:
        f.write("# The restricted, total (collisonal and radiative) range")
        f.write("# E [MeV]   Range [mm] \n")
        f.write("# number of data per materials\n %d\n" % len(egrid))
        f.write("# -----------------------------\n")
        for im, m in enumerate(mats):
            f.write("#\n# Material name: %s\n#\n" % m.name)
            dedx_mm = per[im]["dedx"] * 0.1 * m.rho
            rng = np.concatenate([[0.0], np.cumsum(np.diff(egrid) * 0.5 * (1.0 / dedx_mm[1:] + 1.0 / dedx_mm[:-1]))])
            for ie, e in enumerate(egrid):
                f.write("%.8E  %.8E\n" % (e, rng[ie]))

    # ---- brem IMFP
    bgrid = log_grid(max(ecut, gcut), emax, cfg.n_ekin_el)
    with open(os.path.join(outdir, "imfp_brem.dat"), "w") as f:
        f.write("# The restricted imfp for bremsstrahlung interaction\n")
        f.write("# E [MeV]   Brem_IMFP [1/mm][cm3/g] \n")
        f.write("# number of data per materials and #materials\n %d %d\n" % (len(bgrid), nmat))
        f.write("# -----------------------------\n")
        for im, m in enumerate(mats):
            _, sig = sb_brem_integrals(m, bgrid, gcut)
            sig = np.maximum(sig * 0.1, 1e-20 / m.rho)
            f.write("#\n# Material name: %s\n#\n" % m.name)
            for ie, e in enumerate(bgrid):
                f.write("%.8E  %.8E\n" % (e, sig[ie]))

    # ---- Moller IMFP
    mgrid = log_grid(2.0 * ecut, emax, cfg.n_ekin_el)
    with open(os.path.join(outdir, "imfp_moller.dat"), "w") as f:
        f.write("# The restricted imfp for Moller interaction\n")
        f.write("# E [MeV]   Moller_IMFP [1/mm][cm3/g] \n")
        f.write("# number of data per materials\n %d\n" % len(mgrid))
        f.write("# -----------------------------\n")
        for im, m in enumerate(mats):
            f.write("#\n# Material name: %s\n#\n" % m.name)
            sig = np.maximum(moller_xsec(m, mgrid, ecut) * 0.1, 1e-20 / m.rho)
            for ie, e in enumerate(mgrid):
                f.write("%.8E  %.8E\n" % (e, sig[ie]))

    xi64 = np.linspace(0.0, 1.0, 64)
    xi64[0], xi64[-1] = 0.0, 1.0

    def write_alias(f, x, y):
        w, al = linear_alias_table(x, y)
        for i in range(len(x)):
            f.write("%22.14E  %22.14E  %22.14E  %d " % (x[i], y[i], w[i], al[i]))
            if (i + 1) % 2 == 0 or i == len(x) - 1:
                f.write("\n")

    # ---- Moller sampling tables
    with open(os.path.join(outdir, "ioni_MollerDtrData.dat"), "w") as f:
        f.write("# Tables for rejection free sampling of energy transferred\n")
        f.write("# to the secondary electron in Moller interaction.\n")
        f.write("# The size of the discrete primary (np) and secondary (ns) energies first \n")
        f.write("# then the primary energy and the corresponding (4 x ns) sampling table data \n")
        f.write("# each of the (np) discrete primary energies. \n")
        f.write("%d %d\n" % (len(mgrid), 64))
        for ie, e in enumerate(mgrid):
            f.write("%22.14E\n" % e)
            ep = e + (1e-6 if ie == 0 else (-1e-6 if ie == len(mgrid) - 1 else 0.0))
            write_alias(f, xi64, moller_pdf_xi(xi64, ecut, ep))

    # ---- SB sampling tables
    sgrid = log_grid(gcut, emax, 128)
    with open(os.path.join(outdir, "brem_SBDtrData.dat"), "w") as f:
        f.write("# Tables for rejection free sampling of the emitted photon energy\n")
        f.write("# in electron bremsstrahlung interaction according to the Seltzer-Berger DCS.\n")
        f.write("# The size of the discrete primary (np), secondary (ns) energies and #materials first \n")
        f.write("# followed by the discrete primary energy grid and for each material (np) tables with \n")
        f.write("# (4 x ns) sampling table data each. \n")
        f.write("%d %d %d\n" % (len(sgrid), 64, nmat))
        for ie, e in enumerate(sgrid):
            f.write("%22.14E " % e)
            if (ie + 1) % 6 == 0 or ie == len(sgrid) - 1:
                f.write("\n")
        for im, m in enumerate(mats):
            f.write("#\n# Material name: %s\n#\n" % m.name)
            for ie, e in enumerate(sgrid):
                ep = e + (1e-6 if ie == 0 else (-1e-6 if ie == len(sgrid) - 1 else 0.0))
                write_alias(f, *sb_alias_points(m, gcut, ep))

    # ---- photons
    pgrid = np.empty(cfg.n_ekin_ph)
    pe0, pe1 = 0.99 * gcut, emax
    inv_d = 1.0 / ((pe1 - pe0) / (cfg.n_ekin_ph - 1))
    pgrid[0], pgrid[-1] = pe0, pe1
    pgrid[1:-1] = pe0 + np.arange(1, cfg.n_ekin_ph - 1) / inv_d
    comp, pair, phot, tot = [], [], [], []
    for m in mats:
        c, p, q = photon_mxsec(m, pgrid)
        c, p, q = np.maximum(c, 1e-20), np.maximum(p, 1e-20), np.maximum(q, 1e-20)
        comp.append(c), pair.append(p), phot.append(q), tot.append(c + p + q)
    gmax = np.max(np.array(tot), axis=0)
    for fname, title, data in (("imfp_compton.dat", "# The imfp for Compton scattering\n", comp),
                               ("imfp_pairp.dat", "# The imfp for e- e+ pair production.\n", pair),
                               ("imfp_photoe.dat", "# The imfp for photoelectric effect.\n", phot),
                               ("imfp_total.dat", "# The total imfp = compton + pair + photoe  .\n", tot)):
        col = {"imfp_compton.dat": "Compton_IMFP", "imfp_pairp.dat": "Pair_IMFP",
               "imfp_photoe.dat": "PhotoEl_IMFP", "imfp_total.dat": "Total_IMFP"}[fname]
        with open(os.path.join(outdir, fname), "w") as f:
            f.write(title)
            f.write("# E [MeV]   %s [1/mm] \n" % col)
            f.write("# number of data per materials and #material\n %d %d\n" % (len(pgrid), nmat))
            f.write("# -----------------------------\n")
            for im, m in enumerate(mats):
                f.write("#\n# Material name: %s\n#\n" % m.name)
                for ie, e in enumerate(pgrid):
                    f.write("%.8E  %.8E\n" % (e, data[im][ie] / m.rho))
    with open(os.path.join(outdir, "imfp_globalMax.dat"), "w") as f:
        f.write("# The global max tota-imfp.\n")
        f.write("# E [MeV]   Max-Total_IMFP [1/mm] \n")
        f.write("# number of data per materials\n %d\n" % len(pgrid))
        f.write("# -----------------------------\n")
        for ie, e in enumerate(pgrid):
            f.write("%.8E  %.8E\n" % (e, gmax[ie]))

    kgrid = log_grid(gcut, emax, 32)
    with open(os.path.join(outdir, "compton_KNDtrData.dat"), "w") as f:
        f.write("# Tables for rejection free sampling of reduced energy transfer\n")
        f.write("# in Compton scattering according to the Klein-Nishina DCS.\n")
        f.write("# The size of the discrete primary (np) and secondary (ns) energies first \n")
        f.write("# then the primary energy and the corresponding (4 x ns) sampling table data \n")
        f.write("# each of the (np) discrete primary energies. \n")
        f.write("%d %d\n" % (len(kgrid), 64))
        for ie, e in enumerate(kgrid):
            f.write("%22.14E\n" % e)
            ep = e + (1e-6 if ie == 0 else (-1e-6 if ie == len(kgrid) - 1 else 0.0))
            write_alias(f, xi64, kn_pdf_xi(xi64, ep))
    log("     done.")
    return outdir


TABLE_FILES = ("mat.dat", "el_itr1mfp.dat", "el_scatStrength.dat", "el_GSDtrData.dat", "eloss_rdedx.dat",
               "imfp_brem.dat", "imfp_moller.dat", "ioni_MollerDtrData.dat", "brem_SBDtrData.dat",
               "imfp_globalMax.dat", "imfp_total.dat", "imfp_compton.dat", "imfp_pairp.dat",
               "compton_KNDtrData.dat")


def tables_digest(d):
    """sha256 over the 14 run-time table files (names and bytes)."""
    import hashlib
    h = hashlib.sha256()
    for f in TABLE_FILES:
        with open(os.path.join(d, f), "rb") as fh:
            h.update(f.encode())
            h.update(fh.read())
    return h.hexdigest()


def ensure_tables(outdir, index):
    """Generate the tables into ``outdir`` unless a complete set is there."""
    stamp = os.path.join(outdir, ".complete")
    if os.path.exists(stamp):
        return outdir
    generate(outdir, index, verbose=False)
    with open(stamp, "w") as f:
        f.write("%d\n" % index)
    return outdir


def main(argv=None):
    ap = argparse.ArgumentParser(description="Geant4-free DPM table generator (dpm_GenerateData flags)")
    ap.add_argument("-d", "--output-data-dir", default="./data")
    ap.add_argument("-c", "--configuration-index", type=int, default=0)
    args = ap.parse_args(argv)
    generate(args.output_data_dir, args.configuration_index)


if __name__ == "__main__":
    sys.exit(main())
