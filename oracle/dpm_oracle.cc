// dpm_oracle.cc -- TEST INFRASTRUCTURE ONLY (the parity oracle).
//
// FP64 CPU restatement of the reference DPM transport of bl406/dpm-g4cpp.
// Every function names the reference file:line it follows. Arithmetic is
// written in the same operation order as the reference so that, in the
// mt19937 mode, the depth-dose output is bit-identical to the reference's
// dpm_Simulate run on the same table files (checked by tests/ against the
// reference compiled from its own sources into oracle/_ref).
//
// Extensions beyond the reference (used by the GPU engine and by the parity
// tests, never by the mt19937 reference-compatible path), each pinned by the
// tests named beside it:
//  - a Philox4x32-10 per-particle random stream (Salmon et al., SC'11;
//    Random123 known-answer vectors),
//  - a 3-D voxel phantom (per-voxel material and density) with a 3-D tally
//    (reduces to the reference slab; closed-form ray-box distances),
//  - parallel rectangular and gantry field sources (uniformity and width
//    checks with a planted half-width field), flat, triangular and tabulated
//    energy spectra (inverse-CDF checks over the whole uniform lattice),
//  - phase-space particle lists (reduce to the pencil beam bit for bit;
//    hand ray-box projection),
//  - collimated gantry fields: point source and jaws (exact expectation of
//    the jaw transmission, a planted missing obliquity; open jaws reduce to
//    the parallel field bit for bit),
//  - Woodcock majorant scaling (dose invariance and affine step counts, a
//    planted unscaled interaction probability),
//  - per-batch sum / sum-of-squares tallies (equal to per-batch runs, a
//    planted row shift) and byte-model counters.
#include "dpm_oracle.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>
#include <algorithm>

namespace {

std::string g_err;
// test hook: planted mistakes for the tests (0 = none):
//  1 drop the Compton electron,
//  2 batch rows shifted by one history (history ih counted in the batch of ih + 1),
//  3 field source half as wide (x0 + (u - 0.5) * fx / 2),
//  4 Woodcock interaction probabilities from the unscaled majorant (only the
//    step length scaled),
//  5 jaw transmission without the obliquity (exp(-n tau) instead of
//    exp(-n tau / cos theta))
int g_fault = 0;

// ---------------------------------------------------------------------------
// Spline with log-spaced nodes: Simulation/inc/SimDataSpline.hh:38-50,
// Simulation/src/SimDataSpline.cc:9-67 (natural cubic spline coefficients).
// ---------------------------------------------------------------------------
struct LogSpline {
  int n = 0;
  double logMinX = 0, invLogDelta = 0;
  std::vector<double> X, Y, B, C, D;

  void resize(int m) { n = m; X.assign(m, 0); Y.assign(m, 0); B.assign(m, 0); C.assign(m, 0); D.assign(m, 0); }
  void set(int i, double x, double y) {
    X[i] = x; Y[i] = y;
    if (i == 0) logMinX = std::log(x);
    if (i == 1) invLogDelta = 1. / (std::log(x) - logMinX);
    if (i == n - 1) setup();
  }
  // same recurrences and operation order as SimDataSpline::SetUp (SimDataSpline.cc:33-67)
  void setup() {
    const int last = n - 1;
    double s = 0.0, r = 0.0;
    for (int m = 0; m < last; ++m) {
      D[m] = X[m + 1] - X[m];
      r = (Y[m + 1] - Y[m]) / D[m];
      C[m] = r - s;
      s = r;
    }
    r = 0.0; s = 0.0;
    C[0] = 0.0; C[last] = 0.0;
    for (int m = 1; m < last; ++m) {
      C[m] = C[m] + r * C[m - 1];
      B[m] = 2.0 * (X[m - 1] - X[m + 1]) - r * s;
      s = D[m];
      r = s / B[m];
    }
    for (int k = 1, m = last - 1; k < last; ++k, --m) {
      C[m] = (D[m] * C[m + 1] - C[m]) / B[m];
    }
    for (int m = 0; m < last; ++m) {
      s = D[m];
      r = C[m + 1] - C[m];
      D[m] = r / s;
      C[m] = 3.0 * C[m];
      B[m] = (Y[m + 1] - Y[m]) / s - (C[m] + r) * s;
    }
  }
  // SimDataSpline::GetValue (SimDataSpline.hh:47-50, 38-41)
  double value(double x) const {
    const double lx = std::log(x);
    int i = (int)((lx - logMinX) * invLogDelta);
    i = std::max(0, std::min(n - 2, i));
    const double dx = x - X[i];
    return Y[i] + dx * (B[i] + dx * (C[i] + dx * D[i]));
  }
};

// Linear-index linear interpolation: Simulation/inc/SimDataLinear.hh:38-45
struct LinTable {
  int n = 0;
  double minX = 0, invDelta = 0;
  std::vector<double> X, Y;
  void resize(int m) { n = m; X.assign(m, 0); Y.assign(m, 0); }
  void set(int i, double x, double y) {
    X[i] = x; Y[i] = y;
    if (i == 0) minX = x;
    if (i == 1) invDelta = 1. / (x - minX);
  }
  double value(double x) const {
    int i = (int)((x - minX) * invDelta);
    i = std::max(0, std::min(n - 2, i));
    return (Y[i + 1] - Y[i]) * (x - X[i]) / (X[i + 1] - X[i]) + Y[i];
  }
};

// Clamped table: the thin wrappers of SimITr1MFPElastic.hh:35-39,
// SimStoppingPower.hh:36-42, SimIMFPBrem.hh:35-41, SimIMFPMoller.hh:35-39,
// SimMaxScatStrength.hh:39-43 (energy clamp [Emin, Emax-1e-6], floor 1e-20).
struct ClampedSpline {
  LogSpline s;
  double emin = 0, emax = 0;
  double value(double e) const {
    const double ee = std::min(emax - 1.0E-6, std::max(emin, e));
    return std::max(1.0E-20, s.value(ee));
  }
};
struct ClampedLinear {
  LinTable t;
  double emin = 0, emax = 0;
  double value(double e) const {
    const double ee = std::min(emax - 1.0E-6, std::max(emin, e));
    return std::max(1.0E-20, t.value(ee));
  }
};

// Walker alias + linear pdf: Simulation/src/SimLinAliasData.cc:25-42
struct AliasTable {
  int n = 0;
  std::vector<double> x, y, w;
  std::vector<int> al;
  double sample(double r1, double r2, int* bin) const {
    const double rest = r1 * (n - 1);
    int i = (int)(rest);
    if (w[i] < rest - i) i = al[i];
    if (bin) *bin = i;
    const double xv = x[i];
    const double xd = x[i + 1] - xv;
    if (y[i] > 0.0) {
      const double dum = (y[i + 1] - y[i]) / y[i];
      if (std::abs(dum) > 0.1)
        return xv - xd / dum * (1.0 - std::sqrt(1.0 + r2 * dum * (dum + 2.0)));
      return xv + r2 * xd * (1.0 - 0.5 * dum * (r2 - 1.0) * (1.0 + dum * r2));
    }
    return xv + xd * std::sqrt(r2);
  }
};

struct EnergyAliasSet {   // common part of SimSBTables / SimMollerTables / SimKNTables
  int ne = 0, ns = 0;
  double minE = 0, logMinE = 0, invLogDelta = 0;
  std::vector<AliasTable> t;   // [mat*ne + ie]
};

struct GSTable {
  int ne = 0, ns = 0;
  double minE = 0, logMinE = 0, invLogDelta = 0, deltaCum = 0;
  std::vector<double> egrid, transf;
  std::vector<double> u, a, b;   // [ie*ns + is]
};

struct Data {
  bool loaded = false;
  double ecut = 0, gcut = 0;
  int nmat = 0;
  std::vector<double> rho, mollerScale;
  std::vector<ClampedSpline> itr1, dedx, brem;
  ClampedSpline moller, k1;
  ClampedLinear gmax;
  std::vector<ClampedLinear> ptot, pcomp, ppair;
  EnergyAliasSet sb, mo, kn;
  GSTable gs;
};
Data g;

// ---------------------------------------------------------------------------
// readers: same files, comment-line counts and fields as the reference
// ---------------------------------------------------------------------------
struct File {
  FILE* f = nullptr;
  char line[1024];
  explicit File(const std::string& p) { f = std::fopen(p.c_str(), "r"); }
  ~File() { if (f) std::fclose(f); }
  void skip(int k) { for (int i = 0; i < k; ++i) if (!std::fgets(line, sizeof(line), f)) return; }
};

bool open_or_fail(File& fl, const std::string& p) {
  if (!fl.f) { g_err = "file = " + p + " not found!"; return false; }
  return true;
}

// SimMaterialData.cc:6-53
bool load_material(const std::string& dir) {
  const std::string p = dir + "/mat.dat";
  File fl(p);
  if (!open_or_fail(fl, p)) return false;
  double d1, d2, d3;
  fl.skip(1);
  if (std::fscanf(fl.f, "%lg %lg\n", &g.ecut, &g.gcut) != 2) { g_err = "bad mat.dat"; return false; }
  fl.skip(1);
  if (std::fscanf(fl.f, "%lg %lg %lg\n", &d1, &d2, &d3) != 3) { g_err = "bad mat.dat"; return false; }
  fl.skip(1);
  if (std::fscanf(fl.f, "%d\n", &g.nmat) != 1 || g.nmat <= 0 || g.nmat > 255) { g_err = "bad mat.dat"; return false; }
  g.rho.assign(g.nmat, 0); g.mollerScale.assign(g.nmat, 0);
  for (int im = 0; im < g.nmat; ++im) {
    int idum; char name[512];
    fl.skip(2);
    if (std::fscanf(fl.f, "%d %s %lg\n", &idum, name, &g.rho[im]) != 3) { g_err = "bad mat.dat"; return false; }
    fl.skip(3);
    if (std::fscanf(fl.f, "%lg %lg\n", &g.mollerScale[im], &d1) != 2) { g_err = "bad mat.dat"; return false; }
  }
  return true;
}

// per-material spline files: SimITr1MFPElastic.cc:13-56 (3 columns, col 3 used),
// SimStoppingPower.cc:18-61, SimIMFPBrem.cc:18-61
bool load_mat_splines(const std::string& p, int head, int cols, std::vector<ClampedSpline>& out) {
  File fl(p);
  if (!open_or_fail(fl, p)) return false;
  fl.skip(head);
  int nd, nm;
  if (std::fscanf(fl.f, "%d %d\n", &nd, &nm) != 2 || nd < 2) { g_err = "bad " + p; return false; }
  fl.skip(1);
  out.assign(nm, ClampedSpline());
  for (int im = 0; im < nm; ++im) {
    fl.skip(3);
    out[im].s.resize(nd);
    for (int i = 0; i < nd; ++i) {
      double e, v, dd;
      int ok = (cols == 3) ? std::fscanf(fl.f, "%lg %lg %lg\n", &e, &dd, &v) : std::fscanf(fl.f, "%lg %lg\n", &e, &v);
      if (ok != cols) { g_err = "bad " + p; return false; }
      out[im].s.set(i, e, v);
      if (i == 0) out[im].emin = e;
      if (i == nd - 1) out[im].emax = e;
    }
  }
  // the reference keeps Emin/Emax of material 0 for all materials
  for (int im = 1; im < nm; ++im) { out[im].emin = out[0].emin; out[im].emax = out[0].emax; }
  return true;
}

// single (reference-material) spline files: SimIMFPMoller.cc:6-43, SimMaxScatStrength.cc:6-43
bool load_ref_spline(const std::string& p, int head, ClampedSpline& out) {
  File fl(p);
  if (!open_or_fail(fl, p)) return false;
  fl.skip(head);
  int nd;
  if (std::fscanf(fl.f, "%d\n", &nd) != 1 || nd < 2) { g_err = "bad " + p; return false; }
  fl.skip(4);
  out.s.resize(nd);
  for (int i = 0; i < nd; ++i) {
    double e, v;
    if (std::fscanf(fl.f, "%lg %lg", &e, &v) != 2) { g_err = "bad " + p; return false; }
    out.s.set(i, e, v);
    if (i == 0) out.emin = e;
    if (i == nd - 1) out.emax = e;
  }
  return true;
}

// SimIMFPMaxPhoton.cc:6-38
bool load_gmax(const std::string& dir) {
  const std::string p = dir + "/imfp_globalMax.dat";
  File fl(p);
  if (!open_or_fail(fl, p)) return false;
  fl.skip(3);
  int nd;
  if (std::fscanf(fl.f, "%d\n", &nd) != 1 || nd < 2) { g_err = "bad " + p; return false; }
  fl.skip(1);
  g.gmax.t.resize(nd);
  for (int i = 0; i < nd; ++i) {
    double e, v;
    if (std::fscanf(fl.f, "%lg %lg", &e, &v) != 2) { g_err = "bad " + p; return false; }
    g.gmax.t.set(i, e, v);
    if (i == 0) g.gmax.emin = e;
    if (i == nd - 1) g.gmax.emax = e;
  }
  return true;
}

// SimIMFPPhoton.cc:19-63
bool load_photon(const std::string& p, std::vector<ClampedLinear>& out) {
  File fl(p);
  if (!open_or_fail(fl, p)) return false;
  fl.skip(3);
  int nd, nm;
  if (std::fscanf(fl.f, "%d %d\n", &nd, &nm) != 2 || nd < 2) { g_err = "bad " + p; return false; }
  fl.skip(1);
  out.assign(nm, ClampedLinear());
  for (int im = 0; im < nm; ++im) {
    fl.skip(3);
    out[im].t.resize(nd);
    for (int i = 0; i < nd; ++i) {
      double e, v;
      if (std::fscanf(fl.f, "%lg %lg\n", &e, &v) != 2) { g_err = "bad " + p; return false; }
      out[im].t.set(i, e, v);
      if (i == 0) out[im].emin = e;
      if (i == nd - 1) out[im].emax = e;
    }
  }
  for (int im = 1; im < nm; ++im) { out[im].emin = out[0].emin; out[im].emax = out[0].emax; }
  return true;
}

void set_egrid(EnergyAliasSet& s, int ie, double e) {
  if (ie == 0) { s.minE = e; s.logMinE = std::log(e); }
  if (ie == 1) s.invLogDelta = 1. / (std::log(e) - s.logMinE);
}

// one sampling table; the SB reader's format ends in "\n" (SimSBTables.cc:83),
// which also swallows the line ends before the next material's comment lines
bool read_alias(FILE* f, AliasTable& t, int ns, bool eatws) {
  t.n = ns; t.x.resize(ns); t.y.resize(ns); t.w.resize(ns); t.al.resize(ns);
  const char* fmt = eatws ? "%lg %lg %lg %d\n" : "%lg %lg %lg %d";
  for (int is = 0; is < ns; ++is)
    if (std::fscanf(f, fmt, &t.x[is], &t.y[is], &t.w[is], &t.al[is]) != 4) return false;
  return true;
}

// SimSBTables.cc:21-90
bool load_sb(const std::string& dir) {
  const std::string p = dir + "/brem_SBDtrData.dat";
  File fl(p);
  if (!open_or_fail(fl, p)) return false;
  fl.skip(5);
  int nm;
  if (std::fscanf(fl.f, "%d %d %d", &g.sb.ne, &g.sb.ns, &nm) != 3) { g_err = "bad " + p; return false; }
  for (int ie = 0; ie < g.sb.ne; ++ie) {
    double e;
    if (std::fscanf(fl.f, "%lg\n", &e) != 1) { g_err = "bad " + p; return false; }
    set_egrid(g.sb, ie, e);
  }
  g.sb.t.assign((size_t)nm * g.sb.ne, AliasTable());
  for (int im = 0; im < nm; ++im) {
    fl.skip(3);
    for (int ie = 0; ie < g.sb.ne; ++ie)
      if (!read_alias(fl.f, g.sb.t[(size_t)im * g.sb.ne + ie], g.sb.ns, true)) { g_err = "bad " + p; return false; }
  }
  return true;
}

// SimMollerTables.cc:20-67 and SimKNTables.cc:20-67 (identical layout)
bool load_energy_alias(const std::string& p, EnergyAliasSet& s) {
  File fl(p);
  if (!open_or_fail(fl, p)) return false;
  fl.skip(5);
  if (std::fscanf(fl.f, "%d %d", &s.ne, &s.ns) != 2) { g_err = "bad " + p; return false; }
  s.t.assign(s.ne, AliasTable());
  for (int ie = 0; ie < s.ne; ++ie) {
    double e;
    if (std::fscanf(fl.f, "%lg", &e) != 1) { g_err = "bad " + p; return false; }
    set_egrid(s, ie, e);
    if (!read_alias(fl.f, s.t[ie], s.ns, false)) { g_err = "bad " + p; return false; }
  }
  return true;
}

// SimGSTables.cc:19-80 (material 0 only)
bool load_gs(const std::string& dir) {
  const std::string p = dir + "/el_GSDtrData.dat";
  File fl(p);
  if (!open_or_fail(fl, p)) return false;
  fl.skip(4);
  double d;
  GSTable& t = g.gs;
  if (std::fscanf(fl.f, "%d %lg %lg %d\n", &t.ne, &d, &d, &t.ns) != 4) { g_err = "bad " + p; return false; }
  t.deltaCum = 1.0 / (t.ns - 1);
  fl.skip(6);
  t.egrid.resize(t.ne); t.transf.resize(t.ne);
  t.u.resize((size_t)t.ne * t.ns); t.a.resize(t.u.size()); t.b.resize(t.u.size());
  for (int ie = 0; ie < t.ne; ++ie) {
    int idum; double e, tp;
    if (std::fscanf(fl.f, "%d %lg %lg", &idum, &e, &tp) != 3) { g_err = "bad " + p; return false; }
    t.egrid[ie] = e; t.transf[ie] = tp;
    if (ie == 0) { t.minE = e; t.logMinE = std::log(e); }
    if (ie == 1) t.invLogDelta = 1. / (std::log(e) - t.logMinE);
    for (int is = 0; is < t.ns; ++is) {
      const size_t k = (size_t)ie * t.ns + is;
      if (std::fscanf(fl.f, "%lg %lg %lg", &t.u[k], &t.a[k], &t.b[k]) != 3) { g_err = "bad " + p; return false; }
    }
  }
  return true;
}

// ---------------------------------------------------------------------------
// samplers
// ---------------------------------------------------------------------------
// SimSBTables::SampleEnergyTransfer (SimSBTables.cc:93-114)
double sample_sb(double eprim, int imat, double r1, double r2, double r3, int* ebin, int* abin) {
  if (imat < 0) return 0.0;
  const double lpe = std::log(eprim);
  double ph = (lpe - g.sb.logMinE) * g.sb.invLogDelta;
  int ie = (int)ph;
  ph -= ie;
  if (r1 < ph) ++ie;
  if (ebin) *ebin = ie;
  const double xi = g.sb.t[(size_t)imat * g.sb.ne + ie].sample(r2, r3, abin);
  const double kappac = g.sb.minE / eprim;
  const double kappa = kappac * std::exp(-xi * (g.sb.logMinE - lpe));
  return kappa * eprim;
}

// SimMollerTables::SampleEnergyTransfer (SimMollerTables.cc:69-89)
double sample_moller(double eprim, double r1, double r2, double r3, int* ebin, int* abin) {
  const double lpe = std::log(eprim);
  double ph = (lpe - g.mo.logMinE) * g.mo.invLogDelta;
  int ie = (int)ph;
  ph -= ie;
  if (r1 < ph) ++ie;
  if (ebin) *ebin = ie;
  const double xi = g.mo.t[ie].sample(r2, r3, abin);
  const double dum1 = lpe - g.mo.logMinE;
  return std::exp(xi * dum1) * 0.5 * g.mo.minE;
}

// SimKNTables::SampleEnergyTransfer (SimKNTables.cc:69-90)
double sample_kn(double egamma, double r1, double r2, double r3, int* ebin, int* abin) {
  const double kEMC2 = 0.510991;
  const double lpe = std::log(egamma);
  double ph = (lpe - g.kn.logMinE) * g.kn.invLogDelta;
  int ie = (int)ph;
  ph -= ie;
  if (r1 < ph) ++ie;
  if (ebin) *ebin = ie;
  const double xi = g.kn.t[ie].sample(r2, r3, abin);
  const double kappa = egamma / kEMC2;
  return std::exp(std::log(1. + 2. * kappa) * (xi - 1.));
}

// SimGSTables::SampleAngularDeflection (SimGSTables.cc:84-119)
// 1 - cos(theta) of SimGSTables::SampleAngularDeflection, i.e. the
// reference's 2 a U / (1 - U + a) before it is subtracted from 1
double sample_gs_omu(double eprim, double r1, double r2, int* ebin, int* cbin) {
  const GSTable& t = g.gs;
  const double lpe = std::log(eprim);
  double ph = (lpe - t.logMinE) * t.invLogDelta;
  int ie = (int)ph;
  const int ielow = ie;
  ph -= ie;
  if (r1 < ph) ++ie;
  if (ebin) *ebin = ie;
  const int il = (int)(r2 / t.deltaCum);
  if (cbin) *cbin = il;
  const double resid = r2 - il * t.deltaCum;
  const size_t k = (size_t)ie * t.ns + il;
  const double parA = t.a[k], parB = t.b[k], u0 = t.u[k], u1 = t.u[k + 1];
  const double dum1 = (1.0 + parA + parB) * t.deltaCum * resid;
  const double dum2 = t.deltaCum * t.deltaCum + parA * t.deltaCum * resid + parB * resid * resid;
  const double theU = u0 + dum1 / dum2 * (u1 - u0);
  const double a0 = t.transf[ielow], a1 = t.transf[ielow + 1];
  const double e0 = t.egrid[ielow], e1 = t.egrid[ielow + 1];
  const double parTransf = (a1 - a0) / (e1 - e0) * (eprim - e0) + a0;
  return 2. * parTransf * theU / (1. - theU + parTransf);
}

// SimGSTables::SampleAngularDeflection (SimGSTables.cc:84-119): cos(theta)
double sample_gs(double eprim, double r1, double r2, int* ebin, int* cbin) {
  return 1. - sample_gs_omu(eprim, r1, r2, ebin, cbin);
}

// ---------------------------------------------------------------------------
// random numbers
// ---------------------------------------------------------------------------
inline void philox_round(uint32_t c[4], const uint32_t k[2]) {
  const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
  const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
  const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
  const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
  const uint32_t n0 = hi1 ^ c[1] ^ k[0];
  const uint32_t n2 = hi0 ^ c[3] ^ k[1];
  c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
}

void philox10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t c[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
  uint32_t k[2] = {key[0], key[1]};
  for (int r = 0; r < 10; ++r) {
    if (r > 0) { k[0] += 0x9E3779B9u; k[1] += 0xBB67AE85u; }
    philox_round(c, k);
  }
  out[0] = c[0]; out[1] = c[1]; out[2] = c[2]; out[3] = c[3];
}

inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// serial of the n-th secondary created by a particle with serial `parent`
inline uint64_t child_serial(uint64_t parent, uint32_t n) {
  return mix64(parent ^ (0x9E3779B97F4A7C15ull * (uint64_t)(n + 1)));
}

// Track record: Simulation/inc/Track.hh:10-22 plus the Philox stream state
struct Track {
  double pos[3] = {0, 0, 0};
  double dir[3] = {0, 0, 1};
  int box[3] = {0, 0, 0};
  int type = -3;
  int mat = -1;
  double ekin = 0, edep = 0;
  uint32_t hist = 0;
  uint64_t serial = 0;
  uint32_t ndraw = 0;
  uint32_t nchild = 0;
};

struct Rng {
  int mode = ORACLE_RNG_MT19937;
  bool rtl = true;
  std::mt19937 mt;
  std::uniform_real_distribution<double> dis{0.0, 1.0};
  uint32_t key[2] = {0, 0};
  Track* cur = nullptr;     // Philox: stream of the particle under tracking
  uint32_t blk[4];
  uint32_t blkid = 0xFFFFFFFFu;

  uint32_t draw24() {
    Track& t = *cur;
    const uint32_t b = t.ndraw >> 2;
    if (b != blkid || cached_hist != t.hist || cached_serial != t.serial) {
      const uint32_t ctr[4] = {b, (uint32_t)t.serial, (uint32_t)(t.serial >> 32), t.hist};
      philox10(ctr, key, blk);
      blkid = b; cached_hist = t.hist; cached_serial = t.serial;
    }
    const uint32_t w = blk[t.ndraw & 3];
    ++t.ndraw;
    return w >> 8;
  }
  uint32_t cached_hist = 0xFFFFFFFFu;
  uint64_t cached_serial = ~0ull;

  double u() {
    if (mode == ORACLE_RNG_MT19937) return dis(mt);
    return (double)draw24() * (1.0 / 16777216.0);
  }
  // ln(R) of the exponential draws (SimDPMLike.cc:138,146,192,211,476); the
  // Philox layout takes R=(k+1)/2^24 in (0,1] there so that ln(R) is finite
  double lnu() {
    if (mode == ORACLE_RNG_MT19937) return std::log(dis(mt));
    return std::log((double)(draw24() + 1u) * (1.0 / 16777216.0));
  }
  // the uniforms of a call f(R(),R(),R()): g++ evaluates such argument lists
  // right to left, the Philox layout defines them left to right
  void u3(double& a, double& b, double& c) {
    if (mode == ORACLE_RNG_MT19937 && rtl) { c = u(); b = u(); a = u(); }
    else { a = u(); b = u(); c = u(); }
  }
  void u2(double& a, double& b) {
    if (mode == ORACLE_RNG_MT19937 && rtl) { b = u(); a = u(); }
    else { a = u(); b = u(); }
  }
};

// ---------------------------------------------------------------------------
// geometry and scoring: Simulation/src/Geom.cc (slab), plus a voxel phantom
// ---------------------------------------------------------------------------
struct Geometry {
  int mode = ORACLE_GEOM_SLAB;
  double L = 1, invL = 1, hL = 0.5;
  int geomIndex = 0;
  int nx = 0, ny = 0, nz = 0;
  double org[3] = {0, 0, 0};
  const uint8_t* vmat = nullptr;
  const float* vrho = nullptr;
  const double kTolerance = 1.0E-4;
  const double kExtent = 100.0;
  // scoring
  std::vector<double>* edep = nullptr;
  std::vector<double>* steps = nullptr;
  int histSize = 20;
  int maxIz = -1;
  double* batchBuf = nullptr;
  std::vector<int> touched;
  double histSum = 0;
  double scored = 0;
  uint64_t nScore = 0, nVisit = 0;

  // Geom::DistanceToBoundary (Geom.cc:20-117); the recursion on a boundary push
  // is unrolled into a loop that re-enters from the extent check
  double slab_distance(double* r, const double* v, int* i) const {
    const double kHalfTolerance = 0.5 * kTolerance;
    for (;;) {
      if (std::abs(r[0]) > kExtent || std::abs(r[1]) > kExtent || r[2] > kExtent || r[2] < -L) return -1.0;
      i[0] = (int)std::floor((r[0] + hL) * invL);
      i[1] = (int)std::floor((r[1] + hL) * invL);
      i[2] = (int)std::floor((r[2] + hL) * invL);
      const double tr[3] = {r[0] - i[0] * L, r[1] - i[1] * L, r[2] - i[2] * L};
      double snext = 1.0E+20;
      int push = -1; double sgn = 0;
      for (int a = 0; a < 3 && push < 0; ++a) {
        if (v[a] > 0.0) {
          const double pdist = hL - tr[a];
          if (pdist < kHalfTolerance) { push = a; sgn = 1.0; }
          else { const double s = pdist / v[a]; if (a == 0 || s < snext) snext = s; }
        } else if (v[a] < 0.0) {
          const double pdist = hL + tr[a];
          if (pdist < kHalfTolerance) { push = a; sgn = -1.0; }
          else { const double s = -pdist / v[a]; if (a == 0 || s < snext) snext = s; }
        }
      }
      if (push < 0) return snext;
      if (sgn > 0) r[push] += kTolerance; else r[push] -= kTolerance;
    }
  }

  // voxel phantom: voxel (ix,iy,iz) spans org + [i, i+1) * L; same boundary
  // logic (tolerance push) as the slab; leaving the phantom returns -1
  double vox_distance(double* r, const double* v, int* i) const {
    const double kHalfTolerance = 0.5 * kTolerance;
    const int n[3] = {nx, ny, nz};
    for (;;) {
      double tr[3];
      for (int a = 0; a < 3; ++a) {
        const double uu = r[a] - org[a];
        const double fl = std::floor(uu * invL);
        if (!(fl >= 0.0) || fl >= (double)n[a]) return -1.0;
        i[a] = (int)fl;
        tr[a] = uu - ((double)i[a] + 0.5) * L;
      }
      double snext = 1.0E+20;
      int push = -1; double sgn = 0;
      for (int a = 0; a < 3 && push < 0; ++a) {
        if (v[a] > 0.0) {
          const double pdist = hL - tr[a];
          if (pdist < kHalfTolerance) { push = a; sgn = 1.0; }
          else { const double s = pdist / v[a]; if (a == 0 || s < snext) snext = s; }
        } else if (v[a] < 0.0) {
          const double pdist = hL + tr[a];
          if (pdist < kHalfTolerance) { push = a; sgn = -1.0; }
          else { const double s = -pdist / v[a]; if (a == 0 || s < snext) snext = s; }
        }
      }
      if (push < 0) return snext;
      if (sgn > 0) r[push] += kTolerance; else r[push] -= kTolerance;
    }
  }

  double distance(double* r, const double* v, int* i) const {
    return mode == ORACLE_GEOM_SLAB ? slab_distance(r, v, i) : vox_distance(r, v, i);
  }

  // Geom::GetMaterialIndex (Geom.cc:133-164)
  int material(const int* i) {
    ++nVisit;
    if (mode == ORACLE_GEOM_VOXEL) return vmat[((size_t)i[2] * ny + i[1]) * nx + i[0]];
    const int iz = i[2];
    if (iz < 0) return -1;
    switch (geomIndex) {
      case 0: return 0;
      case 1: if (iz < 10.0 * invL) return 0; else if (iz < 30.0 * invL) return 1; return 0;
      case 2: if (iz < 20.0 * invL) return 0; else if (iz < 40.0 * invL) return 1; return 0;
      case 3: if (iz < 10.0 * invL) return 0; else if (iz < 15.0 * invL) return 1; else if (iz < 25.0 * invL) return 2; return 0;
      default: return 0;
    }
  }
  // Geom::GetVoxelMaterialDensity (Geom.hh:85-87); per-voxel density in the phantom
  double density(int imat, const int* i) const {
    if (mode == ORACLE_GEOM_VOXEL) return (double)vrho[((size_t)i[2] * ny + i[1]) * nx + i[0]];
    return imat > -1 ? g.rho[imat] : 1.0E-40;
  }

  // Geom::Score (Geom.cc:120-130) including its histogram growth rule
  void score(double e, const int* i) {
    ++nScore;
    if (mode == ORACLE_GEOM_SLAB) {
      const int iz = i[2];
      if (iz < 0) return;
      if (iz >= histSize) histSize = iz + 10;
      if (iz > maxIz) maxIz = iz;
      if (iz < (int)edep->size()) { (*edep)[iz] += e; if (steps) (*steps)[iz] += 1.0; }
      scored += e; histSum += e;
      return;
    }
    const size_t k = ((size_t)i[2] * ny + i[1]) * nx + i[0];
    (*edep)[k] += e;
    if (batchBuf) { if (batchBuf[k] == 0.0) touched.push_back((int)k); batchBuf[k] += e; }
    scored += e; histSum += e;
  }
};

// ---------------------------------------------------------------------------
// transport: Simulation/src/SimDPMLike.cc
// ---------------------------------------------------------------------------
struct Sim {
  Geometry geo;
  Rng rng;
  std::vector<Track> stack;
  double majScale = 1.0;
  double escaped = 0;
  uint64_t nPair = 0, nAnnih = 0;
  uint64_t nJaw = 0;   // primaries absorbed in the jaws of a collimated field
  uint64_t nSB = 0, nMo = 0, nKN = 0, nGS = 0, nSec = 0, nEStep = 0, nGStep = 0;

  Track& push(Track& parent) {
    stack.emplace_back();
    Track& t = stack.back();
    t.hist = parent.hist;
    t.serial = child_serial(parent.serial, parent.nchild++);
    ++nSec;
    return t;
  }

  // RotateToLabFrame (SimDPMLike.cc:638-656)
  static void rotate(double& u, double& v, double& w, double u1, double u2, double u3) {
    double up = u1 * u1 + u2 * u2;
    if (up > 0.) {
      up = std::sqrt(up);
      const double px = u, py = v, pz = w;
      u = (u1 * u3 * px - u2 * py) / up + u1 * pz;
      v = (u2 * u3 * px + u1 * py) / up + u2 * pz;
      w = -up * px + u3 * pz;
    } else if (u3 < 0.) {
      u = -u;
      w = -w;
    }
  }

  // PerformAnnihilation (SimDPMLike.cc:753-791)
  void annihilate(Track& tr) {
    const double kPI = 3.1415926535897932;
    const double kEMC2 = 0.510991;
    ++nAnnih;
    const double cost = 1.0 - 2.0 * rng.u();
    const double sint = std::sqrt((1.0 - cost) * (1.0 + cost));
    const double phi = 2.0 * kPI * rng.u();
    const double rx = sint * cos(phi), ry = sint * sin(phi), rz = cost;
    for (int k = 0; k < 2; ++k) {
      Track& a = push(tr);
      a.type = 0; a.ekin = kEMC2; a.mat = tr.mat;
      for (int j = 0; j < 3; ++j) { a.pos[j] = tr.pos[j]; a.box[j] = tr.box[j]; }
      const double sg = k == 0 ? 1.0 : -1.0;
      a.dir[0] = sg * rx; a.dir[1] = sg * ry; a.dir[2] = sg * rz;
    }
  }

  // PerformBrem (SimDPMLike.cc:661-695)
  void brem(Track& tr) {
    const double kPI = 3.1415926535897932;
    const double kEMC2 = 0.510991;
    const double kHalfSqrt2EMC2 = kEMC2 * 0.7071067812;
    double r1, r2, r3;
    rng.u3(r1, r2, r3);
    ++nSB;
    const double eGamma = sample_sb(tr.ekin, tr.mat, r1, r2, r3, nullptr, nullptr);
    Track& a = push(tr);
    a.type = 0; a.ekin = eGamma; a.mat = tr.mat;
    for (int j = 0; j < 3; ++j) { a.pos[j] = tr.pos[j]; a.box[j] = tr.box[j]; }
    const double dum0 = kHalfSqrt2EMC2 / (tr.ekin + kEMC2);
    const double cost = std::max(-1.0, std::min(1.0, 1.0 - dum0 * dum0));
    const double sint = std::sqrt((1.0 + cost) * (1.0 - cost));
    const double phi = 2.0 * kPI * rng.u();
    a.dir[0] = sint * std::cos(phi);
    a.dir[1] = sint * std::sin(phi);
    a.dir[2] = cost;
    rotate(a.dir[0], a.dir[1], a.dir[2], tr.dir[0], tr.dir[1], tr.dir[2]);
    tr.ekin = tr.ekin - eGamma;
  }

  // PerformMoller (SimDPMLike.cc:699-728)
  void moller(Track& tr) {
    const double kPI = 3.1415926535897932;
    const double kEMC2 = 0.510991;
    const double k2EMC2 = 2.0 * kEMC2;
    double r1, r2, r3;
    rng.u3(r1, r2, r3);
    ++nMo;
    const double secEkin = sample_moller(tr.ekin, r1, r2, r3, nullptr, nullptr);
    const double cost = std::sqrt(secEkin * (tr.ekin + k2EMC2) / (tr.ekin * (secEkin + k2EMC2)));
    const double secCost = std::min(1.0, cost);
    Track& a = push(tr);
    a.type = -1; a.ekin = secEkin; a.mat = tr.mat;
    for (int j = 0; j < 3; ++j) { a.pos[j] = tr.pos[j]; a.box[j] = tr.box[j]; }
    const double sint = std::sqrt((1.0 + secCost) * (1.0 - secCost));
    const double phi = 2.0 * kPI * rng.u();
    a.dir[0] = sint * std::cos(phi);
    a.dir[1] = sint * std::sin(phi);
    a.dir[2] = secCost;
    rotate(a.dir[0], a.dir[1], a.dir[2], tr.dir[0], tr.dir[1], tr.dir[2]);
    tr.ekin -= secEkin;
  }

  // PerformMSCAngularDeflection (SimDPMLike.cc:731-750)
  void msc(Track& tr, double ekin0) {
    const double kPI = 3.1415926535897932;
    double r1, r2;
    rng.u2(r1, r2);
    ++nGS;
    const double dum0 = sample_gs(ekin0, r1, r2, nullptr, nullptr);
    const double cost = std::max(-1.0, std::min(1.0, dum0));
    const double sint = std::sqrt((1.0 - cost) * (1.0 + cost));
    const double phi = 2.0 * kPI * rng.u();
    double u1 = sint * std::cos(phi), u2 = sint * std::sin(phi), u3 = cost;
    rotate(u1, u2, u3, tr.dir[0], tr.dir[1], tr.dir[2]);
    tr.dir[0] = u1; tr.dir[1] = u2; tr.dir[2] = u3;
  }

  // KeepTrackingElectron (SimDPMLike.cc:254-460)
  int keep_tracking_electron(Track& tr, double& numTr1MFP, double& numMollerMFP, double invMollerMFP, double& numBremMFP) {
    int what = 0;
    double stepGeom = geo.distance(tr.pos, tr.dir, tr.box);
    if (stepGeom < 0.0) { escaped += tr.ekin; tr.ekin = 0.0; return 4; }
    const double ecut = g.ecut;
    while (what == 0) {
      ++nEStep;
      double stepLength = stepGeom;
      what = 0;
      int imat = geo.material(tr.box);
      if (imat < 0) { escaped += tr.ekin; tr.ekin = 0.0; return 4; }
      tr.mat = imat;
      const double scalMolMFP = g.mollerScale[imat];
      const double rho = geo.density(imat, tr.box);
      double theDEDX = g.dedx[imat].value(tr.ekin) * rho;
      double midStepE = std::max(tr.ekin - 0.5 * stepLength * theDEDX, ecut);
      double delNumTr1MFP = g.itr1[imat].value(midStepE) * rho;
      const double delNumMollerMFP = invMollerMFP * scalMolMFP * rho;
      const double delNumBremMFP = g.brem[imat].value(midStepE) * rho;
      const double stepBrem = numBremMFP / delNumBremMFP;
      if (stepBrem < stepLength) { stepLength = stepBrem; what = 1; }
      const double stepMoller = numMollerMFP / delNumMollerMFP;
      if (stepMoller < stepLength) { stepLength = stepMoller; what = 2; }
      double stepElastic = numTr1MFP / delNumTr1MFP;
      if (stepElastic < stepLength) {
        stepElastic = numTr1MFP / (g.itr1[imat].value(tr.ekin) * rho);
        midStepE = std::max(tr.ekin - 0.5 * stepElastic * theDEDX, ecut);
        delNumTr1MFP = g.itr1[imat].value(midStepE) * rho;
        stepElastic = std::min(stepLength, numTr1MFP / delNumTr1MFP);
        stepLength = stepElastic;
        what = 3;
      }
      numBremMFP = (what == 1) ? 0 : numBremMFP - stepLength * delNumBremMFP;
      numMollerMFP = (what == 2) ? 0 : numMollerMFP - stepLength * delNumMollerMFP;
      numTr1MFP = (what == 3) ? 0 : numTr1MFP - stepLength * delNumTr1MFP;
      midStepE = std::max(tr.ekin - 0.5 * stepLength * theDEDX, ecut);
      theDEDX = g.dedx[imat].value(midStepE) * rho;
      const double deltE = stepLength * theDEDX;
      const double eFinal = tr.ekin - deltE;
      if (eFinal < ecut) {
        tr.edep = tr.ekin;
        tr.ekin = 0.0;
        if (tr.type == +1) {
          tr.pos[0] += tr.dir[0] * stepLength;
          tr.pos[1] += tr.dir[1] * stepLength;
          tr.pos[2] += tr.dir[2] * stepLength;
          geo.distance(tr.pos, tr.dir, tr.box);
          imat = geo.material(tr.box);
          if (imat < 0) { escaped += tr.edep; return 4; }
          tr.mat = imat;
          annihilate(tr);
        }
        what = 4;
      } else {
        tr.ekin = eFinal;
        tr.edep = deltE;
      }
      tr.pos[0] += tr.dir[0] * stepLength;
      tr.pos[1] += tr.dir[1] * stepLength;
      tr.pos[2] += tr.dir[2] * stepLength;
      geo.score(tr.edep, tr.box);
      if (what == 0) {
        stepGeom = geo.distance(tr.pos, tr.dir, tr.box);
        if (stepGeom < 0.0) { escaped += tr.ekin; tr.ekin = 0.0; return 4; }
      }
    }
    return what;
  }

  // KeepTrackingPhoton (SimDPMLike.cc:463-635)
  void keep_tracking_photon(Track& tr) {
    const double kPI = 3.1415926535897932;
    const double kEMC2 = 0.510991;
    const double kInvEMC2 = 1.0 / kEMC2;
    const double ecut = g.ecut, gcut = g.gcut;
    while (tr.ekin > 0.0) {
      ++nGStep;
      double globalMaxMFP = 1.0 / (g.gmax.value(tr.ekin) * majScale);
      const double stepLength = -globalMaxMFP * rng.lnu();
      if (g_fault == 4) globalMaxMFP = 1.0 / g.gmax.value(tr.ekin);
      tr.pos[0] += tr.dir[0] * stepLength;
      tr.pos[1] += tr.dir[1] * stepLength;
      tr.pos[2] += tr.dir[2] * stepLength;
      if (geo.distance(tr.pos, tr.dir, tr.box) < 0) { escaped += tr.ekin; tr.ekin = 0.0; return; }
      const int imat = geo.material(tr.box);
      if (imat < 0) { escaped += tr.ekin; tr.ekin = 0.0; return; }
      tr.mat = imat;
      const double rho = geo.density(imat, tr.box);
      const double totalIMFP = g.ptot[imat].value(tr.ekin) * rho;
      const double r1 = rng.u();
      double theProb = 1.0 - totalIMFP * globalMaxMFP;
      if (r1 < theProb) continue;
      theProb += g.pcomp[imat].value(tr.ekin) * rho * globalMaxMFP;
      if (r1 < theProb) {
        double q1, q2, q3;
        rng.u3(q1, q2, q3);
        ++nKN;
        const double theEps = sample_kn(tr.ekin, q1, q2, q3, nullptr, nullptr);
        const double kappa = tr.ekin * kInvEMC2;
        const double phCost = 1.0 - (1.0 - theEps) / (theEps * kappa);
        const double phEner = theEps * tr.ekin;
        const double phPhi = 2.0 * kPI * rng.u();
        const double elEner = tr.ekin - phEner;
        if (elEner < ecut) {
          geo.score(elEner, tr.box);
        } else if (g_fault == 1) {
          // planted mistake: the Compton electron is lost
        } else {
          const double e0 = tr.ekin * kInvEMC2;
          const double elCost = (1.0 + e0) * std::sqrt((1.0 - theEps) / (e0 * (2.0 + e0 * (1.0 - theEps))));
          Track& a = push(tr);
          a.type = -1; a.ekin = elEner; a.mat = tr.mat;
          for (int j = 0; j < 3; ++j) { a.pos[j] = tr.pos[j]; a.box[j] = tr.box[j]; }
          const double sint = std::sqrt((1.0 + elCost) * (1.0 - elCost));
          const double phi = phPhi + kPI;
          a.dir[0] = sint * std::cos(phi);
          a.dir[1] = sint * std::sin(phi);
          a.dir[2] = elCost;
          rotate(a.dir[0], a.dir[1], a.dir[2], tr.dir[0], tr.dir[1], tr.dir[2]);
        }
        tr.ekin = phEner;
        if (tr.ekin < gcut) {
          geo.score(tr.ekin, tr.box);
          tr.ekin = 0.0;
          return;
        }
        const double phSint = std::sqrt((1.0 - phCost) * (1.0 + phCost));
        double u1 = phSint * std::cos(phPhi), u2 = phSint * std::sin(phPhi), u3 = phCost;
        rotate(u1, u2, u3, tr.dir[0], tr.dir[1], tr.dir[2]);
        tr.dir[0] = u1; tr.dir[1] = u2; tr.dir[2] = u3;
        continue;
      }
      theProb += g.ppair[imat].value(tr.ekin) * rho * globalMaxMFP;
      if (r1 < theProb) {
        ++nPair;
        const double sumEkin = tr.ekin - 2.0 * kEMC2;
        const double e1 = rng.u() * sumEkin;
        const double e2 = sumEkin - e1;
        if (e2 < ecut) {
          geo.score(e2, tr.box);
        } else {
          Track& a = push(tr);
          a.type = -1; a.ekin = e2; a.mat = tr.mat;
          for (int j = 0; j < 3; ++j) { a.pos[j] = tr.pos[j]; a.box[j] = tr.box[j]; a.dir[j] = tr.dir[j]; }
        }
        if (e1 < ecut) {
          geo.score(e1, tr.box);
          annihilate(tr);
        } else {
          Track& a = push(tr);
          a.type = +1; a.ekin = e1; a.mat = tr.mat;
          for (int j = 0; j < 3; ++j) { a.pos[j] = tr.pos[j]; a.box[j] = tr.box[j]; a.dir[j] = tr.dir[j]; }
        }
        tr.ekin = 0.0;
        return;
      }
      geo.score(tr.ekin, tr.box);
      tr.ekin = 0.0;
    }
  }

  // body of the per-track loop of Simulate (SimDPMLike.cc:81-246)
  void track_one(Track& tr) {
    geo.distance(tr.pos, tr.dir, tr.box);
    int imat = geo.material(tr.box);
    if (imat < 0) { escaped += tr.ekin; return; }
    tr.mat = imat;
    if (tr.type == 0) { keep_tracking_photon(tr); return; }
    bool isMSCHinge = true;
    double theEkin0 = tr.ekin;
    double numTr1MFP = g.k1.value(tr.ekin);
    double numTr1MFP0 = rng.u() * numTr1MFP;
    numTr1MFP -= numTr1MFP0;
    double numMollerMFP = -rng.lnu();
    double invMollerMFP = g.moller.value(tr.ekin);
    double numBremMFP = -rng.lnu();
    while (tr.ekin > 0.0) {
      const int what = keep_tracking_electron(tr, numTr1MFP, numMollerMFP, invMollerMFP, numBremMFP);
      if (tr.mat < 0) continue;
      switch (what) {
        case 1:
          if (tr.ekin > g.gcut) brem(tr);
          if (tr.ekin < g.ecut) {
            geo.score(tr.ekin, tr.box);
            tr.ekin = 0.0;
            if (tr.type == +1) annihilate(tr);
            break;
          }
          numBremMFP = -rng.lnu();
          break;
        case 2:
          if (tr.ekin > 2. * g.ecut) moller(tr);
          numMollerMFP = -rng.lnu();
          invMollerMFP = g.moller.value(tr.ekin);
          break;
        case 3:
          if (isMSCHinge) {
            msc(tr, theEkin0);
            numTr1MFP = numTr1MFP0;
            isMSCHinge = false;
          } else {
            theEkin0 = tr.ekin;
            numTr1MFP = g.k1.value(theEkin0);
            numTr1MFP0 = rng.u() * numTr1MFP;
            numTr1MFP -= numTr1MFP0;
            isMSCHinge = true;
          }
          break;
        default:
          break;
      }
    }
  }
};

}  // namespace

extern "C" {

const char* oracle_error(void) { return g_err.c_str(); }
int oracle_num_materials(void) { return g.loaded ? g.nmat : 0; }

int oracle_load(const char* dir) {
  g = Data();
  const std::string d(dir);
  if (!load_material(d)) return 1;
  if (!load_mat_splines(d + "/el_itr1mfp.dat", 4, 3, g.itr1)) return 1;
  if (!load_ref_spline(d + "/el_scatStrength.dat", 4, g.k1)) return 1;
  if (!load_ref_spline(d + "/imfp_moller.dat", 3, g.moller)) return 1;
  if (!load_mat_splines(d + "/imfp_brem.dat", 3, 2, g.brem)) return 1;
  if (!load_mat_splines(d + "/eloss_rdedx.dat", 2, 2, g.dedx)) return 1;
  if (!load_energy_alias(d + "/ioni_MollerDtrData.dat", g.mo)) return 1;
  if (!load_sb(d)) return 1;
  if (!load_gs(d)) return 1;
  if (!load_gmax(d)) return 1;
  if (!load_photon(d + "/imfp_total.dat", g.ptot)) return 1;
  if (!load_photon(d + "/imfp_compton.dat", g.pcomp)) return 1;
  if (!load_photon(d + "/imfp_pairp.dat", g.ppair)) return 1;
  if (!load_energy_alias(d + "/compton_KNDtrData.dat", g.kn)) return 1;
  g.loaded = true;
  return 0;
}

// Source extensions of the voxel phantom (not in the reference, whose
// primaries are a pencil at the slab's front face, SimDPMLike.cc:64-75).
// cos/sin of a gantry angle in degrees, exact at multiples of 90.
void gantry_cos_sin(double deg, double* c, double* s) {
  const double q = deg / 90.0;
  if (q == std::floor(q)) {
    const int k = ((int)std::fmod(q, 4.0) + 4) % 4;
    const double cs[4] = {1.0, 0.0, -1.0, 0.0}, sn[4] = {0.0, 1.0, 0.0, -1.0};
    *c = cs[k];
    *s = sn[k];
    return;
  }
  const double a = deg * (3.14159265358979323846 / 180.0);
  *c = std::cos(a);
  *s = std::sin(a);
}

// energy from a histogram spectrum by inverse CDF with one uniform u: the
// bin i whose cumulative weight interval [cdf[i], cdf[i+1]) holds u (cdf
// normalised by the total weight, its last entry 1), then the same fraction
// of bin i's width
double table_energy(double u, const std::vector<double>& cdf, const double* edges) {
  const int n = (int)cdf.size() - 1;
  int lo = 0;
  for (int i = 1; i < n; ++i)
    if (cdf[i] <= u) lo = i;
  const double f = (u - cdf[lo]) / (cdf[lo + 1] - cdf[lo]);
  return edges[lo] + f * (edges[lo + 1] - edges[lo]);
}

// inverse CDF of the triangular distribution (lo, mode, hi)
double triangular(double u, double lo, double mode, double hi) {
  const double fc = (mode - lo) / (hi - lo);
  if (u < fc) return lo + std::sqrt(u * (hi - lo) * (mode - lo));
  return hi - std::sqrt((1.0 - u) * (hi - lo) * (hi - mode));
}

// Field point (fu, fv) on the plane through the phantom centre perpendicular
// to the beam direction d = (sa, 0, ca), with axes a = (ca, 0, -sa) and
// b = (0, 1, 0): the start point is where the ray of direction d through it
// enters the box [-h, h] (slab method), moved 1e-3 mm inside; false if the
// ray misses the box.
bool gantry_entry(double fu, double fv, double ca, double sa, const double h[3], double pos[3], double dir[3]) {
  const double q[3] = {fu * ca, fv, -fu * sa};
  const double d[3] = {sa, 0.0, ca};
  double tin = -1e300, tout = 1e300;
  for (int i = 0; i < 3; ++i) {
    if (d[i] == 0.0) {
      if (q[i] <= -h[i] || q[i] >= h[i]) return false;
      continue;
    }
    const double t1 = (-h[i] - q[i]) / d[i], t2 = (h[i] - q[i]) / d[i];
    tin = std::max(tin, std::min(t1, t2));
    tout = std::min(tout, std::max(t1, t2));
  }
  if (!(tin + 2e-3 < tout)) return false;
  const double t = tin + 1e-3;
  for (int i = 0; i < 3; ++i) {
    pos[i] = q[i] + t * d[i];
    dir[i] = d[i];
  }
  return true;
}

// Collimated gantry field (an extension, no counterpart in the reference):
// the field point q = (fu ca, fv, -fu sa) on the plane through the phantom
// centre perpendicular to d = (sa, 0, ca). With sad > 0 the particle leaves
// the point source s = -sad d along r = (q - s)/|q - s| (else along d);
// tmin = -|q - s| marks the source on the line q + t r, cos theta = r.d.
// Returns the number of jaw pairs whose aperture the field point (du, dv),
// relative to the field centre, misses ([x1, x2) x [y1, y2); 0 without jaws).
int collimated_ray(double fu, double fv, double du, double dv, double ca, double sa, const oracle_cfg* c,
                   double q[3], double r[3], double* tmin, double* cost) {
  q[0] = fu * ca; q[1] = fv; q[2] = -fu * sa;
  const double d[3] = {sa, 0.0, ca};
  for (int i = 0; i < 3; ++i) r[i] = d[i];
  *tmin = -1e30;
  *cost = 1.0;
  if (c->sad > 0.0) {
    for (int i = 0; i < 3; ++i) r[i] = q[i] + c->sad * d[i];
    const double len = std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
    for (int i = 0; i < 3; ++i) r[i] = r[i] / len;
    *tmin = -len;
    *cost = r[0] * d[0] + r[2] * d[2];
  }
  if (!c->jaws) return 0;
  return (int)(du < c->jaw_x1 || du >= c->jaw_x2) + (int)(dv < c->jaw_y1 || dv >= c->jaw_y2);
}

// where the line q + t r (t >= tmin) enters the box [-h, h] (slab method),
// 1e-3 mm inside; false if it misses
bool ray_entry(const double q[3], const double r[3], double tmin, const double h[3], double pos[3], double dir[3]) {
  double tin = -1e30, tout = 1e30;
  for (int i = 0; i < 3; ++i) {
    if (r[i] == 0.0) {
      if (q[i] <= -h[i] || q[i] >= h[i]) return false;
      continue;
    }
    const double t1 = (-h[i] - q[i]) / r[i], t2 = (h[i] - q[i]) / r[i];
    tin = std::max(tin, std::min(t1, t2));
    tout = std::min(tout, std::max(t1, t2));
  }
  if (!(tin >= tmin && tin + 2e-3 < tout)) return false;
  const double t = tin + 1e-3;
  for (int i = 0; i < 3; ++i) {
    pos[i] = q[i] + t * r[i];
    dir[i] = r[i];
  }
  return true;
}

// a phase-space record outside the voxel phantom (its faces included) moves
// along its direction to where it enters the phantom, 1e-3 mm inside; false
// if the ray misses
bool phsp_entry(const double h[3], double pos[3], const double dir[3]) {
  if (std::fabs(pos[0]) <= h[0] && std::fabs(pos[1]) <= h[1] && std::fabs(pos[2]) <= h[2]) return true;
  double tin = -1e30, tout = 1e30;
  for (int i = 0; i < 3; ++i) {
    if (dir[i] == 0.0) {
      if (pos[i] <= -h[i] || pos[i] >= h[i]) return false;
      continue;
    }
    const double t1 = (-h[i] - pos[i]) / dir[i], t2 = (h[i] - pos[i]) / dir[i];
    tin = std::max(tin, std::min(t1, t2));
    tout = std::min(tout, std::max(t1, t2));
  }
  if (!(tin >= 0.0 && tin + 2e-3 < tout)) return false;
  const double t = tin + 1e-3;
  for (int i = 0; i < 3; ++i) pos[i] = pos[i] + t * dir[i];
  return true;
}

// the state of a primary as generated (x y z u v w E)
void store_primary(const Track& p, double* o) {
  for (int j = 0; j < 3; ++j) { o[j] = p.pos[j]; o[3 + j] = p.dir[j]; }
  o[6] = p.ekin;
}

int oracle_run(const oracle_cfg* cfg, oracle_out* out) {
  if (!g.loaded) { g_err = "tables not loaded"; return 1; }
  Sim s;
  s.rng.mode = cfg->rng_mode;
  s.rng.rtl = cfg->arg_rtl != 0;
  s.rng.mt.seed((std::mt19937::result_type)cfg->seed);
  s.rng.key[0] = (uint32_t)cfg->seed;
  s.rng.key[1] = (uint32_t)(cfg->seed >> 32);
  s.majScale = cfg->majorant_scale > 0 ? cfg->majorant_scale : 1.0;
  Geometry& geo = s.geo;
  geo.mode = cfg->geom_mode;
  geo.L = cfg->lbox; geo.invL = 1. / cfg->lbox; geo.hL = 0.5 * cfg->lbox;
  geo.geomIndex = cfg->geom_index;
  std::vector<double> edep, steps, batch;
  size_t nvox = 0;
  if (geo.mode == ORACLE_GEOM_VOXEL) {
    geo.nx = cfg->nx; geo.ny = cfg->ny; geo.nz = cfg->nz;
    geo.vmat = cfg->vox_mat; geo.vrho = cfg->vox_rho;
    geo.org[0] = -0.5 * cfg->nx * cfg->lbox;
    geo.org[1] = -0.5 * cfg->ny * cfg->lbox;
    geo.org[2] = -0.5 * cfg->nz * cfg->lbox;
    nvox = (size_t)cfg->nx * cfg->ny * cfg->nz;
    for (size_t k = 0; k < nvox; ++k)
      if (cfg->vox_mat[k] >= (uint8_t)g.nmat) { g_err = "voxel material index out of range"; return 1; }
    edep.assign(nvox, 0.0);
    if (cfg->nbatch > 0) { batch.assign(nvox, 0.0); geo.batchBuf = batch.data(); }
  } else {
    edep.assign(out->hist_cap, 0.0);
    steps.assign(out->hist_cap, 0.0);
    geo.steps = &steps;
  }
  geo.edep = &edep;
  if (cfg->rng_mode == ORACLE_RNG_MT19937 && (cfg->src_kind != ORACLE_SRC_PENCIL || geo.mode != ORACLE_GEOM_SLAB)) {
    g_err = "mt19937 mode reproduces the reference pencil beam in the slab only";
    return 1;
  }
  if (cfg->src_kind == ORACLE_SRC_GANTRY && geo.mode != ORACLE_GEOM_VOXEL) {
    g_err = "gantry fields need a voxel phantom";
    return 1;
  }
  const double zstart = geo.mode == ORACLE_GEOM_SLAB ? -0.5 * cfg->lbox : geo.org[2];
  // table spectrum: cumulative weights normalised by the total, last entry 1
  std::vector<double> spec_cdf;
  if (cfg->spectrum == ORACLE_SPEC_TABLE) {
    if (cfg->spec_n < 1 || !cfg->spec_edges || !cfg->spec_weights) {
      g_err = "spectrum table needs bins, edges and weights";
      return 1;
    }
    spec_cdf.assign(cfg->spec_n + 1, 0.0);
    for (int i = 0; i < cfg->spec_n; ++i) spec_cdf[i + 1] = spec_cdf[i] + cfg->spec_weights[i];
    const double total = spec_cdf[cfg->spec_n];
    for (int i = 1; i < cfg->spec_n; ++i) spec_cdf[i] /= total;
    spec_cdf[cfg->spec_n] = 1.0;
  }
  double ca = 1.0, sa = 0.0;
  gantry_cos_sin(cfg->angle_deg, &ca, &sa);
  int curBatch = 0;
  double ep = 0.0;
  for (uint64_t ih = 0; ih < cfg->nhist; ++ih) {
    if (cfg->nbatch > 0 && geo.mode == ORACLE_GEOM_VOXEL) {
      const uint64_t ib = g_fault == 2 ? std::min(ih + 1, cfg->nhist - 1) : ih;
      const int b = (int)((ib * (uint64_t)cfg->nbatch) / cfg->nhist);
      if (b != curBatch) {
        for (int k : geo.touched) { out->batch_sum[k] += batch[k]; out->batch_sum2[k] += batch[k] * batch[k]; batch[k] = 0.0; }
        geo.touched.clear();
        curBatch = b;
      }
    }
    geo.histSum = 0;
    s.stack.clear();
    s.stack.emplace_back();
    Track& p = s.stack.back();
    p.hist = (uint32_t)(cfg->hist_begin + ih);
    p.serial = 0;
    p.type = cfg->ptype;
    p.pos[0] = cfg->x0; p.pos[1] = cfg->y0; p.pos[2] = zstart;
    p.dir[0] = 0.0; p.dir[1] = 0.0; p.dir[2] = 1.0;
    p.ekin = cfg->e0;
    if (cfg->src_kind == ORACLE_SRC_PHSP) {
      const double* r = cfg->phsp + 8 * ih;
      p.type = (int)r[0];
      p.ekin = r[1];
      for (int j = 0; j < 3; ++j) { p.pos[j] = r[2 + j]; p.dir[j] = r[5 + j]; }
      ep += p.ekin;
      if (geo.mode == ORACLE_GEOM_VOXEL) {
        const double h[3] = {-geo.org[0], -geo.org[1], -geo.org[2]};
        if (!phsp_entry(h, p.pos, p.dir)) {
          s.escaped += p.ekin;
          if (out->hist_edep) out->hist_edep[ih] = 0.0;
          if (out->prim) store_primary(p, out->prim + 7 * ih);
          continue;
        }
      }
    } else if (cfg->rng_mode == ORACLE_RNG_PHILOX) {
      s.rng.cur = &p;
      double du = 0.0, dv = 0.0;
      if (cfg->src_kind == ORACLE_SRC_FIELD || cfg->src_kind == ORACLE_SRC_GANTRY) {
        const double ux = s.rng.u(), uy = s.rng.u();
        const double wf = g_fault == 3 ? 0.5 : 1.0;
        du = (ux - 0.5) * (wf * cfg->fx);
        dv = (uy - 0.5) * (wf * cfg->fy);
        p.pos[0] = cfg->x0 + du;
        p.pos[1] = cfg->y0 + dv;
      }
      if (cfg->spectrum == ORACLE_SPEC_TABLE)
        p.ekin = table_energy(s.rng.u(), spec_cdf, cfg->spec_edges);
      else if (cfg->emax > cfg->emin)
        p.ekin = cfg->spectrum == ORACLE_SPEC_TRIANGULAR ? triangular(s.rng.u(), cfg->emin, cfg->emode, cfg->emax)
                                                         : cfg->emin + s.rng.u() * (cfg->emax - cfg->emin);
      ep += p.ekin;
      if (cfg->src_kind == ORACLE_SRC_GANTRY && (cfg->sad > 0.0 || cfg->jaws)) {
        const double h[3] = {-geo.org[0], -geo.org[1], -geo.org[2]};
        double q[3], r[3], tmin, cost;
        const int n = collimated_ray(p.pos[0], p.pos[1], du, dv, ca, sa, cfg, q, r, &tmin, &cost);
        bool absorbed = false;
        if (n > 0) {
          // one uniform decides the unscattered passage through n jaws
          const double u = s.rng.u();
          absorbed = !(u < std::exp(-(double)n * cfg->jaw_tau / (g_fault == 5 ? 1.0 : cost)));
        }
        double pos[3], dir[3];
        if (absorbed || !ray_entry(q, r, tmin, h, pos, dir)) {
          // absorbed in the jaws or missing the phantom: no interaction here
          s.escaped += p.ekin;
          if (absorbed) ++s.nJaw;
          if (out->hist_edep) out->hist_edep[ih] = 0.0;
          if (out->prim) store_primary(p, out->prim + 7 * ih);
          continue;
        }
        for (int j = 0; j < 3; ++j) { p.pos[j] = pos[j]; p.dir[j] = dir[j]; }
      } else if (cfg->src_kind == ORACLE_SRC_GANTRY) {
        const double h[3] = {-geo.org[0], -geo.org[1], -geo.org[2]};
        if (!gantry_entry(p.pos[0], p.pos[1], ca, sa, h, p.pos, p.dir)) {
          // the ray misses the phantom: the primary leaves without interacting
          s.escaped += p.ekin;
          if (out->hist_edep) out->hist_edep[ih] = 0.0;
          if (out->prim) store_primary(p, out->prim + 7 * ih);
          continue;
        }
      }
    }
    if (out->prim) store_primary(p, out->prim + 7 * ih);
    while (!s.stack.empty()) {
      Track tr = s.stack.back();
      s.stack.pop_back();
      s.rng.cur = &tr;
      s.track_one(tr);
    }
    if (out->hist_edep) out->hist_edep[ih] = geo.histSum;
  }
  if (cfg->nbatch > 0 && geo.mode == ORACLE_GEOM_VOXEL) {
    for (int k : geo.touched) { out->batch_sum[k] += batch[k]; out->batch_sum2[k] += batch[k] * batch[k]; }
  }
  if (geo.mode == ORACLE_GEOM_SLAB) {
    for (int i = 0; i < out->hist_cap; ++i) { out->edep[i] = edep[i]; if (out->steps) out->steps[i] = steps[i]; }
  } else {
    std::memcpy(out->edep, edep.data(), nvox * sizeof(double));
  }
  out->hist_size = geo.histSize;
  out->max_iz = geo.maxIz;
  out->e_escaped = s.escaped;
  out->e_primary = cfg->rng_mode == ORACLE_RNG_PHILOX ? ep : cfg->e0 * (double)cfg->nhist;
  out->e_scored = geo.scored;
  out->n_pair = s.nPair;
  out->n_jaw = s.nJaw;
  out->n_annih = s.nAnnih;
  const uint64_t c[9] = {geo.nVisit, geo.nScore, s.nSB, s.nMo, s.nKN, s.nGS, s.nSec, s.nEStep, s.nGStep};
  std::memcpy(out->cnt, c, sizeof(c));
  return 0;
}

double oracle_sample_sb(double e, int imat, double r1, double r2, double r3, int* ebin, int* abin) {
  return sample_sb(e, imat, r1, r2, r3, ebin, abin);
}
double oracle_sample_moller(double e, double r1, double r2, double r3, int* ebin, int* abin) {
  return sample_moller(e, r1, r2, r3, ebin, abin);
}
double oracle_sample_kn(double e, double r1, double r2, double r3, int* ebin, int* abin) {
  return sample_kn(e, r1, r2, r3, ebin, abin);
}
double oracle_sample_gs(double e, double r1, double r2, int* ebin, int* cbin) {
  return sample_gs(e, r1, r2, ebin, cbin);
}
double oracle_sample_gs_omu(double e, double r1, double r2, int* ebin, int* cbin) {
  return sample_gs_omu(e, r1, r2, ebin, cbin);
}

// n samples at one energy with explicit FP64 uniforms (kind 0 SB, 1 Moller,
// 2 KN, 3 GS: returns cos(theta)); for histograms of the model tests
int oracle_sample_many(int kind, int n, double e, int imat, const double* r1, const double* r2, const double* r3,
                       double* out) {
  for (int i = 0; i < n; ++i) {
    switch (kind) {
      case 0: out[i] = sample_sb(e, imat, r1[i], r2[i], r3[i], nullptr, nullptr); break;
      case 1: out[i] = sample_moller(e, r1[i], r2[i], r3[i], nullptr, nullptr); break;
      case 2: out[i] = sample_kn(e, r1[i], r2[i], r3[i], nullptr, nullptr); break;
      case 3: out[i] = sample_gs(e, r1[i], r2[i], nullptr, nullptr); break;
      default: g_err = "unknown sampler"; return -1;
    }
  }
  return 0;
}

// which: 0 itr1(E,m) 1 dedx(E,m) 2 brem(E,m) 3 moller(E) 4 K1(E) 5 gmax(E)
//        6 total(E,m) 7 compton(E,m) 8 pair(E,m)
double oracle_eval(int which, double e, int imat) {
  switch (which) {
    case 0: return g.itr1[imat].value(e);
    case 1: return g.dedx[imat].value(e);
    case 2: return g.brem[imat].value(e);
    case 3: return g.moller.value(e);
    case 4: return g.k1.value(e);
    case 5: return g.gmax.value(e);
    case 6: return g.ptot[imat].value(e);
    case 7: return g.pcomp[imat].value(e);
    case 8: return g.ppair[imat].value(e);
    default: return 0.0;
  }
}

void oracle_philox(uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) { philox10(ctr, key, out); }

// Geometry::distance on a standalone geometry: the reference slab
// (Geom.cc:20-117, boxes of size lbox centred on i*lbox, extent +-100 mm) or a
// centred nx*ny*nz voxel phantom; r is updated by the tolerance push
double oracle_distance(int geom_mode, double lbox, int nx, int ny, int nz, double* r, const double* v, int* box) {
  Geometry geo;
  geo.mode = geom_mode;
  geo.L = lbox; geo.invL = 1. / lbox; geo.hL = 0.5 * lbox;
  geo.nx = nx; geo.ny = ny; geo.nz = nz;
  geo.org[0] = -0.5 * nx * lbox; geo.org[1] = -0.5 * ny * lbox; geo.org[2] = -0.5 * nz * lbox;
  return geo.distance(r, v, box);
}

// the table spectrum's energy for uniforms u[m] (cumulative weights built as
// for a run)
void oracle_table_energy(int n, const double* edges, const double* weights, int m, const double* u, double* e) {
  std::vector<double> cdf(n + 1, 0.0);
  for (int i = 0; i < n; ++i) cdf[i + 1] = cdf[i] + weights[i];
  const double total = cdf[n];
  for (int i = 1; i < n; ++i) cdf[i] /= total;
  cdf[n] = 1.0;
  for (int j = 0; j < m; ++j) e[j] = table_energy(u[j], cdf, edges);
}

// RotateToLabFrame (SimDPMLike.cc:638-656) on uvw[3] with the frame's axis (u1, u2, u3)
void oracle_rotate(double* uvw, double u1, double u2, double u3) { Sim::rotate(uvw[0], uvw[1], uvw[2], u1, u2, u3); }

void oracle_set_fault(int f) { g_fault = f; }

}  // extern "C"
