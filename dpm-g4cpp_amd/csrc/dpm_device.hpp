// dpm_device.hpp -- device-side tables, random numbers, samplers and geometry
// of the MI355X DPM engine (gfx950, wave64).
//
// Everything here is templated on the transport real type R (float for the
// production path, double for the FP64 check path). Citations name the
// reference function each piece re-implements.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dpm {

// ---------------------------------------------------------------------------
// constants of the reference transport (SimDPMLike.cc:464-466, 662-664)
// ---------------------------------------------------------------------------
constexpr double kPI = 3.1415926535897932;
constexpr double kEMC2 = 0.510991;
constexpr double kTolerance = 1.0E-4;   // Geom.hh:107
constexpr double kExtent = 100.0;       // Geom.hh:109
constexpr int kMaxMat = 16;

// ---------------------------------------------------------------------------
// math dispatch
// ---------------------------------------------------------------------------
__device__ __forceinline__ float rlog(float x) { return logf(x); }
__device__ __forceinline__ double rlog(double x) { return log(x); }
__device__ __forceinline__ float rexp(float x) { return expf(x); }
__device__ __forceinline__ double rexp(double x) { return exp(x); }
__device__ __forceinline__ float rsqrt_(float x) { return sqrtf(x); }
__device__ __forceinline__ double rsqrt_(double x) { return sqrt(x); }
__device__ __forceinline__ float rfloor(float x) { return floorf(x); }
__device__ __forceinline__ double rfloor(double x) { return floor(x); }
__device__ __forceinline__ float rabs(float x) { return fabsf(x); }
__device__ __forceinline__ double rabs(double x) { return fabs(x); }
__device__ __forceinline__ void rsincos(float x, float* s, float* c) { sincosf(x, s, c); }
__device__ __forceinline__ void rsincos(double x, double* s, double* c) { *s = sin(x); *c = cos(x); }
template <typename R> __device__ __forceinline__ R rmin(R a, R b) { return b < a ? b : a; }
template <typename R> __device__ __forceinline__ R rmax(R a, R b) { return a < b ? b : a; }

// ---------------------------------------------------------------------------
// table packs (built on the host by dpm_tables_create)
// ---------------------------------------------------------------------------
template <typename R>
struct alignas(sizeof(R) * 8) SplineNode {   // SimDataSpline::OnePoint (SimDataSpline.hh:139-146)
  R x, y, b, c, d, pad0, pad1, pad2;
};

template <typename R>
struct SplineTab {              // one quantity, all materials share the log grid
  const SplineNode<R>* node;    // [nmat * n]
  int n;
  R logMinX, invLogDelta;       // SimDataSpline::fLogMinX, fInvLogDelta
  R emin, emaxm;                // clamp [Emin, Emax - 1e-6] of the thin wrappers
};

template <typename R>
struct LinTab {                 // SimDataLinear, all materials share the grid
  const R* xy;                  // [nmat * n * 2] (x, y) pairs
  int n;
  R minX, invDelta, emin, emaxm;
};

template <typename R>
struct alignas(sizeof(R) * 2) AliasEntry {   // SimLinAliasData::OnePoint
  R x, y;
  uint32_t thr;                 // integer form of fAliasW: keep bin iff frac24 < thr
  uint32_t alias;
};

template <typename R>
struct AliasSet {               // SimSBTables / SimMollerTables / SimKNTables
  const AliasEntry<R>* e;       // [(nmat * ne + ie) * ns + is]
  int ne, ns;
  R minE, logMinE, invLogDelta; // transport precision
  double logMinE_d, invLogDelta_d;  // exact bin selection
};

template <typename R>
struct alignas(sizeof(R) * 4) GSEntry {
  R u, a, b;
  R v;                          // 1 - u, rounded from FP64 (FP32 inversion near u = 1)
};

template <typename R>
struct GSSet {                  // SimGSTables
  const GSEntry<R>* e;          // [ie * ns + is]
  const R* egrid;               // [ne]
  const R* transf;              // [ne]
  int ne, ns;
  R logMinE, invLogDelta, deltaCum;
  double logMinE_d, invLogDelta_d;
};

template <typename R>
struct Tables {
  SplineTab<R> itr1, dedx, brem, moller, k1;
  LinTab<R> gmax, ptot, pcomp, ppair;
  AliasSet<R> sb, mo, kn;
  GSSet<R> gs;
  R rho[kMaxMat];
  R mscale[kMaxMat];
  R ecut, gcut;
  int nmat;
  int lin_shared_grid;          // the maximum and per-material photon IMFPs share one energy grid
};

// ---------------------------------------------------------------------------
// interpolation (SimDataSpline.hh:38-50, SimDataLinear.hh:38-45 and the
// clamping wrappers, e.g. SimStoppingPower.hh:36-42)
// ---------------------------------------------------------------------------
template <typename R>
__device__ __forceinline__ R spline_value(const SplineTab<R>& t, R e, int imat) {
  const R ee = rmin(t.emaxm, rmax(t.emin, e));
  const R lx = rlog(ee);
  int i = (int)((lx - t.logMinX) * t.invLogDelta);
  i = max(0, min(t.n - 2, i));
  const SplineNode<R> nd = t.node[imat * t.n + i];
  const R dx = ee - nd.x;
  return rmax((R)1.0E-20, nd.y + dx * (nd.b + dx * (nd.c + dx * nd.d)));
}

template <typename R>
__device__ __forceinline__ R linear_value(const LinTab<R>& t, R e, int imat) {
  const R ee = rmin(t.emaxm, rmax(t.emin, e));
  int i = (int)((ee - t.minX) * t.invDelta);
  i = max(0, min(t.n - 2, i));
  const R* p = t.xy + 2 * (imat * t.n + i);
  const R x0 = p[0], y0 = p[1], x1 = p[2], y1 = p[3];
  return rmax((R)1.0E-20, (y1 - y0) * (ee - x0) / (x1 - x0) + y0);
}

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11), one stream per particle:
// key = seed, counter = (block, serial_lo, serial_hi, history)
// ---------------------------------------------------------------------------
struct Philox {
  __device__ __forceinline__ static void block(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0,
                                               uint32_t k1, uint32_t out[4]) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
      // full 64-bit products: one v_mad_u64_u32 each instead of mul_hi + mul_lo
      const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
      const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
      const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
      const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
      c0 = n0; c1 = (uint32_t)p1; c2 = n2; c3 = (uint32_t)p0;
      k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
  }
};

__host__ __device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
// serial of the n-th secondary of a particle with serial `parent`
__host__ __device__ __forceinline__ uint64_t child_serial(uint64_t parent, uint32_t n) {
  return mix64(parent ^ (0x9E3779B97F4A7C15ull * (uint64_t)(n + 1)));
}

// Words of one particle stream, buffered up to two Philox blocks ahead. The
// transport loops call topup() at wave-uniform points (after a refill and
// before each step), so the 10 rounds run there for all lanes at once; a draw
// site only selects a register. A draw that finds the buffer empty (more
// draws in one step than the buffer held) generates its block in place.
// Buffering changes when a block is computed, never which words are consumed.
struct Rng {
  uint32_t k0, k1;        // key (seed)
  uint32_t hist;
  uint32_t slo, shi;      // serial
  uint32_t ndraw;         // words consumed so far
  uint32_t nchild;
  uint32_t nblk;          // buffered blocks: b = block ndraw>>2 if nblk >= 1, c = the next if nblk == 2
  uint32_t b0, b1, b2, b3;
  uint32_t c0, c1, c2, c3;
  __device__ __forceinline__ void topup() {
    while (nblk < 2u) {
      uint32_t o[4];
      Philox::block((ndraw >> 2) + nblk, slo, shi, hist, k0, k1, o);
      if (nblk == 0u) { b0 = o[0]; b1 = o[1]; b2 = o[2]; b3 = o[3]; }
      else { c0 = o[0]; c1 = o[1]; c2 = o[2]; c3 = o[3]; }
      ++nblk;
    }
  }
  __device__ __forceinline__ uint32_t draw24() {
    if (nblk == 0u) {
      uint32_t o[4];
      Philox::block(ndraw >> 2, slo, shi, hist, k0, k1, o);
      b0 = o[0]; b1 = o[1]; b2 = o[2]; b3 = o[3];
      nblk = 1u;
    }
    const uint32_t j = ndraw & 3u;
    // the empty asm keeps the words in registers: without it the select chain
    // is folded into an indexed load, which moves the whole track to scratch
    uint32_t x0 = b0, x1 = b1, x2 = b2, x3 = b3;
    asm("" : "+v"(x0), "+v"(x1), "+v"(x2), "+v"(x3));
    const uint32_t w = j == 0u ? x0 : (j == 1u ? x1 : (j == 2u ? x2 : x3));
    ++ndraw;
    if (j == 3u) { b0 = c0; b1 = c1; b2 = c2; b3 = c3; --nblk; }
    return w >> 8;
  }
  template <typename R> __device__ __forceinline__ R uni() { return (R)draw24() * (R)(1.0 / 16777216.0); }
  // ln(R) of the exponential draws, R=(k+1)/2^24 in (0,1]
  template <typename R> __device__ __forceinline__ R lnu() { return rlog((R)(draw24() + 1u) * (R)(1.0 / 16777216.0)); }
  __device__ __forceinline__ uint64_t next_child() {
    const uint64_t s = ((uint64_t)shi << 32) | slo;
    return child_serial(s, nchild++);
  }
};

// ---------------------------------------------------------------------------
// samplers with explicit 24-bit uniforms. The energy-bin pick
// (rndm1 < phigher, e.g. SimSBTables.cc:96-102) is taken in FP64 whenever the
// FP32 value is within 1e-4 of a decision edge, so that the chosen table is the
// one of the FP64 reference; the alias bin is selected with integer
// arithmetic on the 24-bit uniform against an integer form of the keep
// probability (exact for uniforms on the 2^-24 lattice).
// ---------------------------------------------------------------------------
// FP64 re-evaluation of an energy-bin pick that the FP32 value cannot decide;
// returns the chosen bin and the lower grid index (out of line: rare path)
__device__ __noinline__ int2 pick_energy_bin_fp64(double e, double logMinE_d, double invLogDelta_d, uint32_t k1);

template <typename R>
__device__ __forceinline__ int pick_energy_bin(R lpe, R logMinE, R invLogDelta, double logMinE_d, double invLogDelta_d,
                                               R e, uint32_t k1, int* ielow) {
  const R ph = (lpe - logMinE) * invLogDelta;
  int ie = (int)ph;
  R frac = ph - (R)ie;
  const R r1 = (R)k1 * (R)(1.0 / 16777216.0);
  if (sizeof(R) == 4) {
    const R d = frac - r1;
    if (rabs(d) < (R)1e-4 || frac < (R)1e-4 || frac > (R)(1.0 - 1e-4)) {
      const int2 r = pick_energy_bin_fp64((double)e, logMinE_d, invLogDelta_d, k1);
      if (ielow) *ielow = r.y;
      return r.x;
    }
  }
  if (ielow) *ielow = ie;
  return (r1 < frac) ? ie + 1 : ie;
}

// inversion of the linear pdf within the chosen alias bin [a, b]
// (SimLinAliasData.cc:32-41)
template <typename R>
__device__ __forceinline__ R alias_invert(const AliasEntry<R>& a, const AliasEntry<R>& b, uint32_t k3) {
  const R r2 = (R)k3 * (R)(1.0 / 16777216.0);
  const R xv = a.x;
  const R xd = b.x - xv;
  if (a.y > (R)0) {
    const R dum = (b.y - a.y) / a.y;
    if (rabs(dum) > (R)0.1) {
      if (sizeof(R) == 8) return xv - xd / dum * ((R)1.0 - rsqrt_((R)1.0 + r2 * dum * (dum + (R)2.0)));
      // 1 - sqrt(1+t) = -t / (1 + sqrt(1+t)), free of cancellation
      const R t = r2 * dum * (dum + (R)2.0);
      return xv + xd * r2 * (dum + (R)2.0) / ((R)1.0 + rsqrt_((R)1.0 + t));
    }
    return xv + r2 * xd * ((R)1.0 - (R)0.5 * dum * (r2 - (R)1.0) * ((R)1.0 + dum * r2));
  }
  return xv + xd * rsqrt_(r2);
}

// SimLinAliasData::Sample (SimLinAliasData.cc:25-42). The bin's entry and
// its successor are loaded together: when the bin is kept they are the pdf's
// end points, so only an aliased bin costs a second round trip.
template <typename R>
__device__ __forceinline__ R alias_sample(const AliasEntry<R>* tab, int ns, uint32_t k2, uint32_t k3, int* abin) {
  const uint32_t rest = k2 * (uint32_t)(ns - 1);
  uint32_t i = rest >> 24;
  const uint32_t fr = rest & 0xFFFFFFu;
  AliasEntry<R> a = tab[i];
  AliasEntry<R> b = tab[i + 1];
  if (fr >= a.thr) {
    i = a.alias;
    a = tab[i];
    b = tab[i + 1];
  }
  if (abin) *abin = (int)i;
  return alias_invert(a, b, k3);
}

// exp(xi * (ln E - ln Emin)) of the SB and Moller energy transfers in FP32:
// the argument (up to ~6) is formed in FP64 from the FP32 ln E, so that its
// rounding does not add ~5e-7 relative error to the result
__device__ __forceinline__ float transfer_exp(float xi, float lpe, double logMinE_d) {
  return rexp((float)((double)xi * ((double)lpe - logMinE_d)));
}

// SimSBTables::SampleEnergyTransfer (SimSBTables.cc:93-114): photon energy
template <typename R>
__device__ __forceinline__ R sample_sb(const AliasSet<R>& s, R e, int imat, uint32_t k1, uint32_t k2, uint32_t k3,
                                       int* ebin, int* abin) {
  const R lpe = rlog(e);
  const int ie = pick_energy_bin(lpe, s.logMinE, s.invLogDelta, s.logMinE_d, s.invLogDelta_d, e, k1, (int*)nullptr);
  if (ebin) *ebin = ie;
  const R xi = alias_sample(s.e + (size_t)(imat * s.ne + ie) * s.ns, s.ns, k2, k3, abin);
  if (sizeof(R) == 8) {
    const R kappac = s.minE / e;
    const R kappa = kappac * rexp(-xi * (s.logMinE - lpe));
    return kappa * e;
  }
  return s.minE * transfer_exp((float)xi, (float)lpe, s.logMinE_d);
}

// SimMollerTables::SampleEnergyTransfer (SimMollerTables.cc:69-89): secondary energy
template <typename R>
__device__ __forceinline__ R sample_moller(const AliasSet<R>& s, R e, uint32_t k1, uint32_t k2, uint32_t k3, int* ebin,
                                           int* abin) {
  const R lpe = rlog(e);
  const int ie = pick_energy_bin(lpe, s.logMinE, s.invLogDelta, s.logMinE_d, s.invLogDelta_d, e, k1, (int*)nullptr);
  if (ebin) *ebin = ie;
  const R xi = alias_sample(s.e + (size_t)ie * s.ns, s.ns, k2, k3, abin);
  if (sizeof(R) == 4) return transfer_exp((float)xi, (float)lpe, s.logMinE_d) * (R)0.5 * s.minE;
  const R dum1 = lpe - s.logMinE;
  return rexp(xi * dum1) * (R)0.5 * s.minE;
}

// SimKNTables::SampleEnergyTransfer (SimKNTables.cc:69-90): eps = E1/E0
template <typename R>
__device__ __forceinline__ R sample_kn(const AliasSet<R>& s, R e, uint32_t k1, uint32_t k2, uint32_t k3, int* ebin,
                                       int* abin) {
  const R lpe = rlog(e);
  const int ie = pick_energy_bin(lpe, s.logMinE, s.invLogDelta, s.logMinE_d, s.invLogDelta_d, e, k1, (int*)nullptr);
  if (ebin) *ebin = ie;
  const R xi = alias_sample(s.e + (size_t)ie * s.ns, s.ns, k2, k3, abin);
  const R kappa = e / (R)kEMC2;
  if (sizeof(R) == 4) return rexp((R)((double)rlog((R)1. + (R)2. * kappa) * ((double)xi - 1.0)));
  return rexp(rlog((R)1. + (R)2. * kappa) * (xi - (R)1.));
}

// the arithmetic of gs_rational on loaded table values: cumulative-bin
// entries g0, g1 and the transform parameters and grid energies at ielow and
// ielow + 1. Returns 1-cos(theta) (exact form 2aU/(1-U+a)).
template <typename R>
__device__ __forceinline__ R gs_invert(const GSEntry<R>& g0, const GSEntry<R>& g1, R a0, R a1, R e0, R e1, R deltaCum,
                                       int ns, uint32_t k2, R e) {
  const uint32_t rest = k2 * (uint32_t)(ns - 1);
  const int il = (int)(rest >> 24);
  if (sizeof(R) == 8) {
    const R r2 = (R)k2 * (R)(1.0 / 16777216.0);
    const R resid = r2 - (R)il * deltaCum;
    const R dum1 = ((R)1.0 + g0.a + g0.b) * deltaCum * resid;
    const R dum2 = deltaCum * deltaCum + g0.a * deltaCum * resid + g0.b * resid * resid;
    const R theU = g0.u + dum1 / dum2 * (g1.u - g0.u);
    const R parTransf = (a1 - a0) / (e1 - e0) * (e - e0) + a0;
    return (R)2. * parTransf * theU / ((R)1. - theU + parTransf);
  }
  // FP32 tables, arithmetic after the loads in FP64 (full rate on this GPU).
  // 1-U is interpolated from the stored complements instead of formed as
  // 1 - U, which cancels as U -> 1, and the fraction within the cumulative
  // bin is kept in FP64: in the last bins (u1 - u0 ~ 0.03, a ~ 1e-3) an
  // FP32 fraction alone moves cos(theta) by several 1e-6.
  const double dc = (double)deltaCum;
  const double resid = (double)(rest & 0xFFFFFFu) * (1.0 / 16777216.0) * dc;
  const double ga = (double)g0.a, gb = (double)g0.b;
  const double f = (1.0 + ga + gb) * dc * resid / (dc * dc + ga * dc * resid + gb * resid * resid);
  const double U = (double)g0.u + f * ((double)g1.u - (double)g0.u);
  const double omU = (double)g0.v + f * ((double)g1.v - (double)g0.v);
  const double pa = ((double)a1 - (double)a0) / ((double)e1 - (double)e0) * ((double)e - (double)e0) + (double)a0;
  return (R)(2.0 * pa * U / (omU + pa));
}

// Rational inverse of SimGSTables::SampleAngularDeflection (SimGSTables.cc:
// 96-119) for the picked energy table `ie` (and lower grid index `ielow` of
// the transform parameter). Returns 1-cos(theta).
template <typename R>
__device__ __forceinline__ R gs_rational(const GSSet<R>& s, int ie, int ielow, R e, uint32_t k2, int* cbin) {
  // cumulative bin (int)(r2/deltaCum) == floor(r2*(ns-1)) exactly on the lattice
  const int il = (int)((k2 * (uint32_t)(s.ns - 1)) >> 24);
  if (cbin) *cbin = il;
  return gs_invert(s.e[ie * s.ns + il], s.e[ie * s.ns + il + 1], s.transf[ielow], s.transf[ielow + 1],
                   s.egrid[ielow], s.egrid[ielow + 1], s.deltaCum, s.ns, k2, e);
}

// SimGSTables::SampleAngularDeflection (SimGSTables.cc:84-119). Returns
// 1-cos(theta); cos(theta) = 1 - result.
template <typename R>
__device__ __forceinline__ R sample_gs_omu(const GSSet<R>& s, R e, uint32_t k1, uint32_t k2, int* ebin, int* cbin) {
  const R lpe = rlog(e);
  int ielow = 0;
  const int ie = pick_energy_bin(lpe, s.logMinE, s.invLogDelta, s.logMinE_d, s.invLogDelta_d, e, k1, &ielow);
  if (ebin) *ebin = ie;
  return gs_rational(s, ie, ielow, e, k2, cbin);
}

// per-lane choice between two spline tables (one evaluation site for both)
template <typename R>
__device__ __forceinline__ SplineTab<R> pick_spline(bool c, const SplineTab<R>& a, const SplineTab<R>& b) {
  SplineTab<R> r;
  r.node = c ? a.node : b.node;
  r.n = c ? a.n : b.n;
  r.logMinX = c ? a.logMinX : b.logMinX;
  r.invLogDelta = c ? a.invLogDelta : b.invLogDelta;
  r.emin = c ? a.emin : b.emin;
  r.emaxm = c ? a.emaxm : b.emaxm;
  return r;
}

// RotateToLabFrame (SimDPMLike.cc:638-656)
template <typename R>
__device__ __forceinline__ void rotate_to_lab(R& u, R& v, R& w, R u1, R u2, R u3) {
  R up = u1 * u1 + u2 * u2;
  if (up > (R)0.) {
    up = rsqrt_(up);
    const R px = u, py = v, pz = w;
    u = (u1 * u3 * px - u2 * py) / up + u1 * pz;
    v = (u2 * u3 * px + u1 * py) / up + u2 * pz;
    w = -up * px + u3 * pz;
  } else if (u3 < (R)0.) {
    u = -u;
    w = -w;
  }
}

// ---------------------------------------------------------------------------
// geometry (Geom.cc:20-117, 133-164) -- slab of the reference, or voxel phantom
// ---------------------------------------------------------------------------
enum { GEOM_SLAB = 0, GEOM_VOXEL = 1 };

template <typename R>
struct Geometry {
  R L, invL, hL;
  int geomIndex;                  // slab
  int nx, ny, nz;                 // voxel
  R org[3];                       // voxel: corner of voxel (0,0,0)
  const uint8_t* vmat;            // [nz][ny][nx]
  const float* vrho;
  R rho[kMaxMat];                 // nominal densities (slab)
  unsigned long long* guard;      // counts tolerance-push loops that end without a decision (or null)
};

// one pass of the boundary computation; returns the push axis (0..2) and
// direction via *sgn when the point sits within half tolerance of a face
template <typename R>
__device__ __forceinline__ int local_distance(const R tr[3], const R v[3], R hL, R* snext, R* sgn) {
  const R kHalfTolerance = (R)(0.5 * kTolerance);
  R s = (R)1.0E+20;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    if (v[a] > (R)0.0) {
      const R pdist = hL - tr[a];
      if (pdist < kHalfTolerance) { *sgn = (R)1; return a; }
      const R t = pdist / v[a];
      if (a == 0 || t < s) s = t;
    } else if (v[a] < (R)0.0) {
      const R pdist = hL + tr[a];
      if (pdist < kHalfTolerance) { *sgn = (R)-1; return a; }
      const R t = -pdist / v[a];
      if (a == 0 || t < s) s = t;
    }
  }
  *snext = s;
  return -1;
}

// Geom::DistanceToBoundary; the recursive tolerance push becomes a bounded
// loop re-entering from the extent check (the reference recursion).
template <typename R, int GEOM>
__device__ __forceinline__ R distance_to_boundary(const Geometry<R>& g, R r[3], const R v[3], int box[3]) {
  for (int it = 0; it < 16; ++it) {
    R tr[3];
    if (GEOM == GEOM_SLAB) {
      if (rabs(r[0]) > (R)kExtent || rabs(r[1]) > (R)kExtent || r[2] > (R)kExtent || r[2] < -g.L) return (R)-1.0;
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        box[a] = (int)rfloor((r[a] + g.hL) * g.invL);
        tr[a] = r[a] - (R)box[a] * g.L;
      }
    } else {
      const int n[3] = {g.nx, g.ny, g.nz};
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        const R uu = r[a] - g.org[a];
        const R fl = rfloor(uu * g.invL);
        if (!(fl >= (R)0.0) || fl >= (R)n[a]) return (R)-1.0;
        box[a] = (int)fl;
        tr[a] = uu - ((R)box[a] + (R)0.5) * g.L;
      }
    }
    R snext, sgn;
    const int push = local_distance(tr, v, g.hL, &snext, &sgn);
    if (push < 0) return snext;
    // explicit selects keep r[] in registers (no dynamic indexing)
    const R d = sgn * (R)kTolerance;
    if (push == 0) r[0] += d;
    else if (push == 1) r[1] += d;
    else r[2] += d;
  }
  // pathological: treat as leaving the geometry, and count it
  if (g.guard) atomicAdd(g.guard, 1ull);
  return (R)-1.0;
}

// DistanceToBoundary where only the box, the tolerance push and the extent
// test are used (Woodcock photon steps, track starts, the stopping
// positron's position): the loop of distance_to_boundary -- extent test and
// box, then the first axis in x, y, z order whose face ahead lies within half
// the tolerance pushes the position by the tolerance and the loop re-enters --
// without the face distances. False where distance_to_boundary returns -1.
template <typename R, int GEOM>
__device__ __forceinline__ bool locate(const Geometry<R>& g, R r[3], const R v[3], int box[3]) {
  const R kHalfTolerance = (R)(0.5 * kTolerance);
  for (int it = 0; it < 16; ++it) {
    R tr[3];
    if (GEOM == GEOM_SLAB) {
      if (rabs(r[0]) > (R)kExtent || rabs(r[1]) > (R)kExtent || r[2] > (R)kExtent || r[2] < -g.L) return false;
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        box[a] = (int)rfloor((r[a] + g.hL) * g.invL);
        tr[a] = r[a] - (R)box[a] * g.L;
      }
    } else {
      const int n[3] = {g.nx, g.ny, g.nz};
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        const R uu = r[a] - g.org[a];
        const R fl = rfloor(uu * g.invL);
        if (!(fl >= (R)0.0) || fl >= (R)n[a]) return false;
        box[a] = (int)fl;
        tr[a] = uu - ((R)box[a] + (R)0.5) * g.L;
      }
    }
    int push = -1;
    R sgn = (R)0;
#pragma unroll
    for (int a = 2; a >= 0; --a) {   // the lowest qualifying axis wins, as in local_distance
      if (v[a] > (R)0.0 && g.hL - tr[a] < kHalfTolerance) { push = a; sgn = (R)1; }
      else if (v[a] < (R)0.0 && g.hL + tr[a] < kHalfTolerance) { push = a; sgn = (R)-1; }
    }
    if (push < 0) return true;
    const R d = sgn * (R)kTolerance;
    if (push == 0) r[0] += d;
    else if (push == 1) r[1] += d;
    else r[2] += d;
  }
  // pathological: treat as leaving the geometry, and count it
  if (g.guard) atomicAdd(g.guard, 1ull);
  return false;
}

// box of a position without the extent check (the reference keeps the
// parent's box when DistanceToBoundary returns -1 at a track's start)
template <typename R, int GEOM>
__device__ __forceinline__ void box_of(const Geometry<R>& g, const R r[3], int box[3]) {
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    if (GEOM == GEOM_SLAB) box[a] = (int)rfloor((r[a] + g.hL) * g.invL);
    else box[a] = (int)rfloor((r[a] - g.org[a]) * g.invL);
  }
}

// Geom::GetMaterialIndex (Geom.cc:133-164)
template <typename R, int GEOM>
__device__ __forceinline__ int material_index(const Geometry<R>& g, const int box[3]) {
  if (GEOM == GEOM_VOXEL) {
    if (box[0] < 0 || box[1] < 0 || box[2] < 0 || box[0] >= g.nx || box[1] >= g.ny || box[2] >= g.nz) return -1;
    return (int)g.vmat[((size_t)box[2] * g.ny + box[1]) * g.nx + box[0]];
  }
  const int iz = box[2];
  if (iz < 0) return -1;
  const R fiz = (R)iz;
  switch (g.geomIndex) {
    case 1: return fiz < (R)10.0 * g.invL ? 0 : (fiz < (R)30.0 * g.invL ? 1 : 0);
    case 2: return fiz < (R)20.0 * g.invL ? 0 : (fiz < (R)40.0 * g.invL ? 1 : 0);
    case 3: return fiz < (R)10.0 * g.invL ? 0 : (fiz < (R)15.0 * g.invL ? 1 : (fiz < (R)25.0 * g.invL ? 2 : 0));
    default: return 0;
  }
}

// Geom::GetVoxelMaterialDensity (Geom.hh:85-87); per-voxel in the phantom
template <typename R, int GEOM>
__device__ __forceinline__ R voxel_density(const Geometry<R>& g, int imat, const int box[3]) {
  if (GEOM == GEOM_VOXEL) return (R)g.vrho[((size_t)box[2] * g.ny + box[1]) * g.nx + box[0]];
  return g.rho[imat];
}

}  // namespace dpm
