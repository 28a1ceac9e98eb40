// dpm_engine.hpp -- structures shared by the host engine and the kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dpm_device.hpp"

namespace dpm {

// counters accumulated by the kernels (the first ten are exported through
// dpm_stats.counters; deposits counts scoring calls, tally atomics the FP64
// atomics actually issued after write-combining; guard counts particles ended
// by the per-particle iteration cap)
enum {
  kCntVoxelVisits = 0, kCntDeposits, kCntSB, kCntMoller, kCntKN, kCntGS, kCntSecondaries, kCntESteps, kCntGSteps,
  kCntTallyAtomics, kCntPair, kCntAnnih, kCntLaneSlots, kCntGuard, kNumCounters
};

// counters per kernel kind: 0 photons, 1 electrons
struct StatsDev {
  unsigned long long cnt[2][kNumCounters];
  double escaped;
  double e_primary;
  double scored;            // sum of the returned tally (set after the run)
  int maxiz;
  int pad;
  unsigned long long guard_push;   // boundary / locate loops that ran out of tolerance pushes
};

// Particle queue in blocks of 64 records (AoSoA): block b holds slots
// [64b, 64b + 64) as 7 runs of 64 transport-precision values (x y z u v w E)
// followed by 4 runs of 64 u32 (history, serial lo, serial hi, tt). A wave's
// 64 consecutive slots read or write one contiguous run per field, every
// field of a slot sits at a fixed offset from its first one (an instruction
// immediate: at most 2.75 KB in FP32), and a queue is one base pointer.
// tt packs type (2 bits) | draws already used (6) | tag (24).
// Electron queues are filled from both ends: Compton electrons below
// kBackBelow MeV from slot cap - 1 downwards (the back run), every other
// record from slot 0 upwards (the front run); consumers take the front run
// first, so that the particles still in flight when a kernel's input runs
// out are short-lived ones and the kernel's tail is short. One 64-bit counter holds
// both run lengths (front: low 32 bits, back: high 32 bits), so a
// reservation sees the other run and the two can never overlap.
constexpr int kQLanes = 64;
template <typename R>
struct QueueView {
  unsigned char* base;          // cap / 64 blocks of kQBlockBytes<R>
  unsigned long long* count;    // front | back << 32 (slots reserved)
  int cap;                      // multiple of 64
};
constexpr float kBackBelow = 1.0f;
// control words of a run (int index into the device array; the four packed
// queue counters G0, G1, E0, E1 take ints 0..7)
constexpr int kCtlHeadG = 8, kCtlOverflow = 9, kCtlHeadE = 10, kCtlInts = 16;
template <typename R> constexpr int kQRealRun = kQLanes * (int)sizeof(R);
template <typename R> constexpr int kQU32 = 7 * kQRealRun<R>;            // offset of the u32 runs in a block
template <typename R> constexpr int kQBlockBytes = kQU32<R> + 4 * kQLanes * 4;

// capacity of the LDS copy of the electron kernel's spline tables (nodes;
// 20 B each): 3 * nmat * n + 2 * n = 1408 for 3 materials of 128 nodes
constexpr int kLdsSplineNodes = 1536;
// capacity of the LDS copy of the photon kernel's linear tables (floats): the
// shared energy grid, the maximum IMFP and 3 IMFPs per material, n nodes each:
// (2 + 3 * nmat) * n = 5120 for one material of 1024 nodes, 8192 for two;
// 32 KB, so that four blocks of the photon kernel fit in a CU's 160 KB. The
// large copy (48 KB, three blocks per CU) takes three materials (11264): the
// lookups it saves outweigh the lost block on heterogeneous phantoms, whose
// Woodcock tracking makes many null steps.
constexpr int kLdsLinearFloats = 8192;
constexpr int kLdsLinearFloatsLarge = 12288;

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
  int refill_min;           // idle lanes that trigger a wave refill (1..64)
  int tally_mode;           // voxel tally: 0 per-lane write-combining, 1 one atomic per deposit
  int lds_tables;           // FP32 electron kernel: spline tables copied into LDS (they fit)
  int lds_photon;           // FP32 photon kernel: linear tables in LDS: 0 no, 1 small copy, 2 large copy
};

// Primaries of one particle kind, enumerated in list order: entry i owns
// this kind's history indices [cum[i], cum[i+1]), whose global history ids
// (Philox stream ids) start at first[i]. The plain source is one entry; the
// beamlet form has one entry per beamlet of the kind.
struct PrimList {
  int n;
  const uint64_t* cum;      // [n + 1]
  const uint64_t* first;    // [n]
  const double* x;          // [n] entry point (mm)
  const double* y;
  const double* e;          // [n] energy (MeV) unless a spectrum is sampled
  const int* particle;      // [n] -1 e-, 0 gamma, +1 e+
  const int* row;           // [n] tally row, or -1: batch row (or row 0)
  const double* ca;         // [n] gantry fields: cos and sin of the angle
  const double* sa;
  const double* z;          // [n] phase-space records: entry z (mm) and direction
  const double* u;
  const double* v;
  const double* w;
};

// Source of one run, resident in device memory for its duration
struct SourceParams {
  int kind;                 // 0 pencil, 1 rectangular field around the entry point, 2 gantry field,
                            // 3 phase-space records (one history per list entry, its own position,
                            // direction, energy and particle)
  int phsp_project;         // 3, voxel phantom: records outside the phantom move along their direction
                            // to where they enter it
  int nbatch;               // > 1: tally row = batch index - first batch of the chunk
  double emin, emax;        // emax > emin: energy from the spectrum over [emin, emax)
  int spectrum;             // 0 flat, 1 triangular with mode emode, 2 table
  double emode;
  int spec_n;               // table: bins, their edges [spec_n + 1] and the
  const double* spec_edges; // normalised cumulative weights [spec_n + 1]
  const double* spec_cdf;   // (0 ... 1)
  double fx, fy, z0;
  // gantry-field collimation: sad > 0 a point source at that distance
  // upstream of the phantom centre; jaws != 0 the aperture [jx1, jx2) x
  // [jy1, jy2) of field coordinates relative to the field centre, outside
  // which a particle goes on with probability exp(-n jtau / cos theta)
  double sad;
  int jaws;
  double jx1, jx2, jy1, jy2, jtau;
  uint64_t hist_begin, nhist_total;
  PrimList list[2];         // [0] photons, [1] electrons and positrons
  uint64_t perm[2];         // > 0: index v of list k is processed as (v * perm[k]) mod (its count), a
                            // bijection (perm prime, not a divisor of the count) that interleaves the
                            // entries' histories; 0: list order
  uint64_t count[2];        // histories of each list
};

// primaries a transport launch generates after its input queue: this kind's
// indices [beg, end)
struct PrimRange {
  uint64_t beg, end;
  uint32_t batch_first;
};

#define DPM_DECLARE_LAUNCHERS(R, SFX)                                                                              \
  hipError_t launch_photons##SFX(const KParams<R>& P, const SourceParams* S, const PrimRange& pr, int geom,       \
                                 int blocks, const QueueView<R>& in, int* head, const QueueView<R>& qe,            \
                                 const QueueView<R>& qg, hipStream_t st);                                          \
  hipError_t launch_electrons##SFX(const KParams<R>& P, const SourceParams* S, const PrimRange& pr, int geom,     \
                                   int blocks, const QueueView<R>& in, int* head, const QueueView<R>& qe,          \
                                   const QueueView<R>& qg, hipStream_t st);                                        \
  hipError_t launch_perform##SFX(const Tables<R>& T, int kind, int n, double energy, int imat, uint64_t seed,     \
                                 double* out_v, double* out_c, hipStream_t st);                                    \
  hipError_t launch_sample##SFX(const Tables<R>& T, int kind, int n, const double* e, const int* imat,             \
                                const uint32_t* r1, const uint32_t* r2, const uint32_t* r3, double* out,           \
                                int* ebin, int* abin, hipStream_t st);                                             \
  hipError_t launch_eval##SFX(const Tables<R>& T, int which, int n, const double* e, const int* imat,              \
                              double* out, hipStream_t st);

DPM_DECLARE_LAUNCHERS(float, _f32)
DPM_DECLARE_LAUNCHERS(double, _f64)

}  // namespace dpm
