// dpm_engine.hpp -- structures shared by the host engine and the kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dpm_device.hpp"

namespace dpm {

// counters accumulated by the kernels (the first ten are exported through
// dpm_stats.counters; deposits counts scoring calls, tally atomics the FP64
// atomics actually issued after write-combining)
enum {
  kCntVoxelVisits = 0, kCntDeposits, kCntSB, kCntMoller, kCntKN, kCntGS, kCntSecondaries, kCntESteps, kCntGSteps,
  kCntTallyAtomics, kCntPair, kCntAnnih, kNumCounters
};

// counters per kernel kind: 0 photons, 1 electrons
struct StatsDev {
  unsigned long long cnt[2][kNumCounters];
  double escaped;
  double e_primary;
  double scored;            // sum of the returned tally (set after the run)
  int maxiz;
  int pad;
};

// SoA particle queue; tt packs type (2 bits) | draws already used (6) | tag (24)
template <typename R>
struct QueueView {
  R *x, *y, *z, *u, *v, *w, *e;
  uint32_t *hist, *slo, *shi, *tt;
  int* count;
  int cap;
};

template <typename R>
struct KParams {
  Tables<R> T;
  Geometry<R> G;
  uint32_t k0, k1;          // Philox key
  R majScale;               // Woodcock majorant scale
  double* tally;            // slab: [nbins]; voxel: [rows][nvox]
  double* steps;            // slab: [nbins] scoring-call counts
  int64_t nvox;
  int nbins;
  double* hist_edep;        // optional per-history tally
  uint64_t hist_base;
  StatsDev* stats;
  int* overflow;
  int kind;                 // 0 photon kernel, 1 electron kernel (counter slot)
};

struct SourceParams {
  int kind;                 // 0 pencil, 1 rectangular field
  int particle;             // -1 e-, 0 gamma, +1 e+
  double energy, emin, emax;
  double x0, y0, fx, fy, z0;
  int nbatch;               // >1: tally row = batch index - batch_first
  uint32_t batch_first;
  uint64_t hist_begin, nhist_total;
  int nbeamlet;             // >0: beamlet source, tally row = beamlet
  const uint64_t* bl_offset;  // [nbeamlet] first global history of each beamlet
  const double* bl_x;
  const double* bl_y;
  const double* bl_e;
  const int* bl_type;
};

#define DPM_DECLARE_LAUNCHERS(R, SFX)                                                                              \
  hipError_t launch_source##SFX(const KParams<R>& P, const SourceParams& S, uint64_t h0, int n,                     \
                                const QueueView<R>& qe, const QueueView<R>& qg, hipStream_t st);                   \
  hipError_t launch_photons##SFX(const KParams<R>& P, int geom, int blocks, const QueueView<R>& in, int* head,     \
                                 const QueueView<R>& qe, const QueueView<R>& qg, hipStream_t st);                  \
  hipError_t launch_electrons##SFX(const KParams<R>& P, int geom, int blocks, const QueueView<R>& in, int* head,   \
                                   const QueueView<R>& qe, const QueueView<R>& qg, hipStream_t st);                \
  hipError_t launch_sample##SFX(const Tables<R>& T, int kind, int n, const double* e, const int* imat,             \
                                const uint32_t* r1, const uint32_t* r2, const uint32_t* r3, double* out,           \
                                int* ebin, int* abin, hipStream_t st);                                             \
  hipError_t launch_eval##SFX(const Tables<R>& T, int which, int n, const double* e, const int* imat,              \
                              double* out, hipStream_t st);

DPM_DECLARE_LAUNCHERS(float, _f32)
DPM_DECLARE_LAUNCHERS(double, _f64)

}  // namespace dpm
