// dpm_kernels.hip -- event-based transport kernels of the MI355X DPM engine.
//
// Compiled twice: with DPM_REAL=float (production FP32 transport, FP64 tally)
// and DPM_REAL=double (FP64-everywhere check path, no FMA contraction so its
// arithmetic follows the CPU restatement operation for operation).
//
// Work decomposition (replaces the history loop and LIFO TrackStack of
// Simulate, SimDPMLike.cc:57-247, TrackStack.hh:17-59):
//   - secondaries live in SoA queues (one for photons, one for e-/e+);
//   - k_photons / k_electrons are persistent kernels: each wave keeps 64
//     particles in flight and refills finished lanes from the input queue,
//     then from the primaries of the current chunk (generated in registers),
//     in per-wave ranges of 64 taken with one atomic; secondaries are
//     appended to the output queues in per-wave blocks with ballot + prefix
//     counts;
//   - the host alternates photon/electron rounds until both queues drain.
// Every particle carries its own Philox stream (seed, history, serial), so
// results do not depend on the processing order.
#include "dpm_device.hpp"
#include "dpm_engine.hpp"

#ifndef DPM_REAL
#error "DPM_REAL must be float or double"
#endif

namespace dpm {

using R = DPM_REAL;

constexpr int kBlock = 256;

// FP32 kernels: photons held to 128 VGPRs (4 waves per SIMD), electrons to
// 168 (3 waves) without spills; the FP64 build would spill under such bounds
#if DPM_PREC == 32
#define DPM_PHOTON_WAVES __attribute__((amdgpu_waves_per_eu(4)))
#define DPM_ELECTRON_WAVES __attribute__((amdgpu_waves_per_eu(3)))
#else
#define DPM_PHOTON_WAVES
#define DPM_ELECTRON_WAVES
#endif

__device__ __noinline__ int2 pick_energy_bin_fp64(double e, double logMinE_d, double invLogDelta_d, uint32_t k1) {
  const double phd = (log(e) - logMinE_d) * invLogDelta_d;
  const int ie = (int)phd;
  const double fr = phd - (double)ie;
  return make_int2(((double)k1 * (1.0 / 16777216.0) < fr) ? ie + 1 : ie, ie);
}
constexpr int kSlabLds = 1024;         // depth bins privatised in LDS
struct NoLds {                         // placeholder of an LDS table a kernel variant does not use
  int unused;
};
constexpr uint32_t kMaxIter = 200000;  // per-particle iteration cap (termination guard)

// ---------------------------------------------------------------------------
// wave helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & 63u); }
__device__ __forceinline__ uint64_t lanes_below() { return (1ull << lane_id()) - 1ull; }

// Output queues are appended in per-wave blocks of kBlk consecutive slots: a
// wave takes a new block from the queue counter only when its current block
// is full, so the shared counter sees one atomic per kBlk secondaries instead
// of one per wave and loop trip. The block state (base, used) of each wave
// lives in LDS because appends happen inside divergent branches. Slots a wave
// reserved but did not fill are marked empty at the end of the kernel and
// skipped by the consumer.
constexpr int kBlk = 64;
constexpr uint32_t kEmptySlot = 3u;   // type code of an unfilled queue slot

template <typename Q>
struct OutQ {
  Q q;
  volatile int* st;    // this wave's [base, used] of the front run in LDS
  volatile int* stb;   // the same for the back run (electron queues), or null
};

// Slots for the active lanes, each in the front run (back = false) or the
// back run (only where the queue has one: stb not null), counted from that
// run's end of the queue; a slot >= cap when the queue is full. Both runs go
// through one code path, and one 64-bit atomic takes new blocks for both;
// a reservation that would make the runs meet fails.
__device__ __forceinline__ int wave_slot(unsigned long long* counter, int cap, volatile int* st, volatile int* stb,
                                         bool back) {
  const uint64_t act = __ballot(1);
  const uint64_t bm = stb ? __ballot(back) : 0ull;
  const uint64_t fm = act & ~bm;
  const int kf = __popcll(fm), kb = __popcll(bm);
  const int leader = __ffsll((long long)act) - 1;
  const int fbase = st[0], fused = st[1];
  const int bbase = stb ? stb[0] : 0, bused = stb ? stb[1] : 0;
  const int froom = kBlk - fused, broom = kBlk - bused;
  const bool fnew = kf > froom, bnew = kb > broom;
  int fnb = 0, bnb = 0;
  if (fnew || bnew) {   // the same in every lane
    if (lane_id() == leader) {
      const unsigned long long inc =
          (fnew ? (unsigned long long)kBlk : 0ull) | (bnew ? (unsigned long long)kBlk << 32 : 0ull);
      const unsigned long long old = atomicAdd(counter, inc);
      const uint64_t f = old & 0xffffffffull, b = old >> 32;
      const bool ok = f + b + (fnew ? kBlk : 0) + (bnew ? kBlk : 0) <= (uint64_t)cap;
      fnb = ok ? (int)f : cap;
      bnb = ok ? (int)b : cap;
    }
    fnb = __shfl(fnb, leader);
    bnb = __shfl(bnb, leader);
  }
  const int rank = (int)__popcll((back ? bm : fm) & lanes_below());
  const int slot = back ? (rank < broom ? bbase + bused + rank : bnb + (rank - broom))
                        : (rank < froom ? fbase + fused + rank : fnb + (rank - froom));
  if (lane_id() == leader) {
    if (fnew) { st[0] = fnb; st[1] = kf - froom; }
    else st[1] = fused + kf;
    if (stb) {
      if (bnew) { stb[0] = bnb; stb[1] = kb - broom; }
      else stb[1] = bused + kb;
    }
  }
  return slot;
}

// ---------------------------------------------------------------------------
// per-block accumulators (LDS), flushed once per block
// ---------------------------------------------------------------------------
template <int NS>
struct BlockAccT {
  double escaped;
  double eprim;                 // energy of the primaries generated by the block
  unsigned cnt[kNumCounters];
  int maxiz;
  double slab_e[NS];
  unsigned slab_n[NS];
};
template <int GEOM> struct AccOf { using type = BlockAccT<(GEOM == GEOM_SLAB) ? kSlabLds : 1>; };

template <class BlockAcc>
__device__ __forceinline__ void acc_init(BlockAcc& a, bool slab) {
  for (int i = threadIdx.x; i < kNumCounters; i += blockDim.x) a.cnt[i] = 0u;
  if (threadIdx.x == 0) { a.escaped = 0.0; a.eprim = 0.0; a.maxiz = -1; }
  if (slab)
    for (int i = threadIdx.x; i < (int)(sizeof(a.slab_n) / sizeof(unsigned)); i += blockDim.x) {
      a.slab_e[i] = 0.0;
      a.slab_n[i] = 0u;
    }
  __syncthreads();
}

template <class BlockAcc>
__device__ __forceinline__ void acc_flush(BlockAcc& a, const KParams<R>& P, bool slab) {
  __syncthreads();
  for (int i = threadIdx.x; i < kNumCounters; i += blockDim.x)
    if (a.cnt[i]) atomicAdd((unsigned long long*)&P.stats->cnt[P.kind][i], (unsigned long long)a.cnt[i]);
  if (threadIdx.x == 0) {
    if (a.escaped != 0.0) unsafeAtomicAdd(&P.stats->escaped, a.escaped);
    if (a.eprim != 0.0) unsafeAtomicAdd(&P.stats->e_primary, a.eprim);
    if (a.maxiz >= 0) atomicMax(&P.stats->maxiz, a.maxiz);
  }
  if (slab) {
    const int ns = (int)(sizeof(a.slab_n) / sizeof(unsigned));
    const int nb = P.nbins < ns ? P.nbins : ns;
    for (int i = threadIdx.x; i < nb; i += blockDim.x) {
      if (a.slab_n[i]) {
        unsafeAtomicAdd(&P.tally[i], a.slab_e[i]);
        unsafeAtomicAdd(&P.steps[i], (double)a.slab_n[i]);
      }
    }
  }
}

// one LDS atomic per wave (active lanes counted with a ballot)
template <class BlockAcc>
__device__ __forceinline__ void count(BlockAcc& a, int which) {
  const uint64_t act = __ballot(1);
  if (lane_id() == __ffsll((long long)act) - 1) atomicAdd(&a.cnt[which], (unsigned)__popcll(act));
}
// two counters that every active lane increments (one ballot)
template <class BlockAcc>
__device__ __forceinline__ void count2(BlockAcc& a, int w1, int w2) {
  const uint64_t act = __ballot(1);
  if (lane_id() == __ffsll((long long)act) - 1) {
    const unsigned n = (unsigned)__popcll(act);
    atomicAdd(&a.cnt[w1], n);
    atomicAdd(&a.cnt[w2], n);
  }
}

// Geom::Score (Geom.cc:120-130) for the slab; voxel tally row `tag`
template <int GEOM, class BlockAcc>
__device__ __forceinline__ void score(BlockAcc& a, const KParams<R>& P, R e, const int box[3], uint32_t tag,
                                      uint32_t hist) {
  count(a, kCntDeposits);
  if (GEOM == GEOM_SLAB) {
    const int iz = box[2];
    if (iz < 0) return;
    atomicMax(&a.maxiz, iz);
    if (iz < (int)(sizeof(a.slab_n) / sizeof(unsigned)) && iz < kSlabLds) {
      atomicAdd(&a.slab_e[iz], (double)e);
      atomicAdd(&a.slab_n[iz], 1u);
    } else if (iz < P.nbins) {
      unsafeAtomicAdd(&P.tally[iz], (double)e);
      unsafeAtomicAdd(&P.steps[iz], 1.0);
    }
  } else {
    const size_t k = ((size_t)box[2] * P.G.ny + box[1]) * P.G.nx + box[0];
    count(a, kCntTallyAtomics);
    unsafeAtomicAdd(&P.tally[(size_t)tag * P.nvox + k], (double)e);
  }
  if (P.hist_edep) unsafeAtomicAdd(&P.hist_edep[hist - P.hist_base], (double)e);
}

// Write-combined voxel scoring of the electron kernel: consecutive deposits of
// a lane into the same voxel are summed in an FP64 register and reach the
// tally with one atomic when the voxel changes or the particle ends.
struct WComb {
  int64_t vox;     // pending voxel (-1: none)
  double e;        // pending energy
};

// Per-wave LDS privatisation of hot tally cells (voxel tally mode 2): each
// wave owns a direct-mapped table of kHotSlots (cell, FP64 sum) entries in
// front of the global FP64 atomics. A write-combined sum goes to the slot of
// its cell (row * nvox + voxel): it is added there when the slot holds that
// cell; otherwise the lane takes the slot with a compare-and-swap on the key
// and sends the evicted sum to HBM, or, when another lane of the wave took
// the slot for a different cell in the same instruction, sends its own sum to
// HBM. Only the owning wave touches its table, so wavefront-scope fences order
// the steps; the wave flushes the table when it ends.
constexpr int kHotSlots = 256;
constexpr unsigned long long kHotEmpty = ~0ull;
struct HotTable {
  unsigned long long key[kBlock / 64][kHotSlots];
  double val[kBlock / 64][kHotSlots];
};
// only the kernel variant launched for tally mode 2 holds the table (16 KB of
// LDS per block would otherwise cost the other variants a block per CU)
template <bool HOT> struct HotOf { using type = NoLds; };
template <> struct HotOf<true> { using type = HotTable; };

__device__ __forceinline__ void wave_fence() { __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront"); }

__device__ __forceinline__ void hot_init(HotTable& H, int wave) {
  for (int i = lane_id(); i < kHotSlots; i += 64) {
    H.key[wave][i] = kHotEmpty;
    H.val[wave][i] = 0.0;
  }
  wave_fence();
}

template <class BlockAcc>
__device__ __forceinline__ void hot_add(BlockAcc& a, const KParams<R>& P, HotTable& H, int wave, uint64_t k,
                                        double e) {
  unsigned long long* key = H.key[wave];
  double* val = H.val[wave];
  const int s = (int)((k * 0x9E3779B97F4A7C15ull) >> 56);   // 8 bits: kHotSlots = 256
  const unsigned long long K = __hip_atomic_load(&key[s], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
  const bool hit = K == k;
  if (hit) atomicAdd(&val[s], e);
  wave_fence();
  unsigned long long old = K;
  if (!hit) old = atomicCAS(&key[s], K, (unsigned long long)k);
  const bool won = !hit && old == K;
  wave_fence();
  double ov = 0.0;
  if (won) ov = atomicExch(&val[s], e);
  wave_fence();
  const bool join = !hit && !won && old == k;
  if (join) atomicAdd(&val[s], e);
  if (won && K != kHotEmpty) {
    count(a, kCntTallyAtomics);
    unsafeAtomicAdd(&P.tally[K], ov);
  }
  if (!hit && !won && !join) {
    count(a, kCntTallyAtomics);
    unsafeAtomicAdd(&P.tally[k], e);
  }
}

// end of a wave: every cached sum to HBM (the wave's lanes have all finished)
template <class BlockAcc>
__device__ __forceinline__ void hot_flush(BlockAcc& a, const KParams<R>& P, HotTable& H, int wave) {
  wave_fence();
  for (int i = lane_id(); i < kHotSlots; i += 64) {
    const unsigned long long K = H.key[wave][i];
    if (K != kHotEmpty) {
      count(a, kCntTallyAtomics);
      unsafeAtomicAdd(&P.tally[K], H.val[wave][i]);
    }
  }
}

// a write-combined sum to the tally: global atomic, or the wave's LDS table
template <class BlockAcc, class Hot>
__device__ __forceinline__ void tally_add(BlockAcc& a, const KParams<R>& P, Hot& H, int wave, int64_t k, double e) {
  if constexpr (sizeof(Hot) == sizeof(HotTable)) {
    if (P.tally_mode == 2) {
      hot_add(a, P, H, wave, (uint64_t)k, e);
      return;
    }
  }
  count(a, kCntTallyAtomics);
  unsafeAtomicAdd(&P.tally[k], e);
}

template <int GEOM, class BlockAcc, class Hot>
__device__ __forceinline__ void score_wc(BlockAcc& a, const KParams<R>& P, WComb& w, Hot& H, int wave, R e,
                                         const int box[3], uint32_t tag, uint32_t hist) {
  if (GEOM == GEOM_SLAB || P.tally_mode == 1) {
    score<GEOM>(a, P, e, box, tag, hist);
    return;
  }
  count(a, kCntDeposits);
  const int64_t k = (int64_t)((size_t)tag * P.nvox + ((size_t)box[2] * P.G.ny + box[1]) * P.G.nx + box[0]);
  if (k != w.vox) {
    if (w.vox >= 0) tally_add(a, P, H, wave, w.vox, w.e);
    w.vox = k;
    w.e = 0.0;
  }
  w.e += (double)e;
  if (P.hist_edep) unsafeAtomicAdd(&P.hist_edep[hist - P.hist_base], (double)e);
}

template <class BlockAcc, class Hot>
__device__ __forceinline__ void flush_wc(BlockAcc& a, const KParams<R>& P, WComb& w, Hot& H, int wave) {
  if (w.vox >= 0) {
    tally_add(a, P, H, wave, w.vox, w.e);
    w.vox = -1;
    w.e = 0.0;
  }
}

template <class BlockAcc>
__device__ __forceinline__ void escape(BlockAcc& a, R e) {
  if (e != (R)0) atomicAdd(&a.escaped, (double)e);
}

// ---------------------------------------------------------------------------
// queue records
// ---------------------------------------------------------------------------
// type code: 0 photon, 1 electron, 2 positron; tt = type | ndraw << 2 | tag << 8
__device__ __forceinline__ uint32_t make_tt(int type, uint32_t ndraw, uint32_t tag) {
  return (uint32_t)type | (ndraw << 2) | (tag << 8);
}

// a slot's record: field f of the 7 real fields at r + f * kQRealRun, u32
// field g at u + g * 256 (the same 64-bit address plus an immediate in FP32)
struct QSlot {
  unsigned char* r;
  unsigned char* u;
};
__device__ __forceinline__ QSlot qslot(const QueueView<R>& q, int slot) {
  unsigned char* blk = q.base + (uint64_t)(uint32_t)(slot >> 6) * (uint32_t)kQBlockBytes<R>;
  const uint32_t l = (uint32_t)slot & 63u;
  return QSlot{blk + l * sizeof(R), blk + kQU32<R> + l * 4u};
}
__device__ __forceinline__ R& qreal(const QSlot& s, int f) { return *(R*)(s.r + f * kQRealRun<R>); }
__device__ __forceinline__ uint32_t& qu32(const QSlot& s, int g) { return *(uint32_t*)(s.u + g * (kQLanes * 4)); }
enum { kQHist = 0, kQSlo = 1, kQShi = 2, kQTT = 3 };

// SPLIT: a record below kBackBelow MeV goes to the back run. Only the
// Compton electrons of the photon kernel are split: they are most of the
// electron kernel's input, and the extra code runs in almost every trip of
// the photon kernel, so the rarer pair and Moller electrons keep the
// front-only path.
template <bool SPLIT = false>
__device__ __forceinline__ void push(const OutQ<QueueView<R>>& o, int* overflow, const R pos[3], R u, R v, R w, R e,
                                     uint32_t hist, uint64_t serial, int type, uint32_t tag) {
  const QueueView<R>& q = o.q;
  const bool back = SPLIT && e < (R)kBackBelow;
  const int slot = wave_slot(q.count, q.cap, o.st, SPLIT ? o.stb : nullptr, back);
  if (slot >= q.cap) { atomicAdd(overflow, 1); return; }
  const QSlot s = qslot(q, back ? q.cap - 1 - slot : slot);
  qreal(s, 0) = pos[0]; qreal(s, 1) = pos[1]; qreal(s, 2) = pos[2];
  qreal(s, 3) = u; qreal(s, 4) = v; qreal(s, 5) = w; qreal(s, 6) = e;
  qu32(s, kQHist) = hist;
  qu32(s, kQSlo) = (uint32_t)serial;
  qu32(s, kQShi) = (uint32_t)(serial >> 32);
  qu32(s, kQTT) = make_tt(type, 0u, tag);
}

__device__ __forceinline__ void close_run(const QueueView<R>& q, volatile int* st, bool back) {
  const int base = st[0], used = st[1];
  const int l = lane_id();
  if (used < kBlk && l >= used && l < kBlk && base + l < q.cap)
    qu32(qslot(q, back ? q.cap - 1 - (base + l) : base + l), kQTT) = kEmptySlot;
}
__device__ __forceinline__ void close_block(const OutQ<QueueView<R>>& o) {
  close_run(o.q, o.st, false);
  if (o.stb) close_run(o.q, o.stb, true);
}

// records in a queue: the front run, then the back run
struct InCount {
  int front, n;
};
__device__ __forceinline__ InCount in_count(const QueueView<R>& q) {
  const unsigned long long c = *q.count;
  const int f = (int)min(c & 0xffffffffull, (unsigned long long)q.cap);
  const int b = (int)min(c >> 32, (unsigned long long)(q.cap - f));
  return InCount{f, f + b};
}
// queue slot of input index i < n
__device__ __forceinline__ int in_slot(const QueueView<R>& q, const InCount& c, int i) {
  return i < c.front ? i : q.cap - 1 - (i - c.front);
}

struct Track {
  R pos[3];
  R dir[3];
  R E;
  int box[3];
  int type;        // 0 photon, 1 e-, 2 e+
  uint32_t tag;
  uint32_t iter;
  Rng rng;
};

__device__ __forceinline__ void load_track(const QueueView<R>& q, int i, const KParams<R>& P, Track& t) {
  const QSlot s = qslot(q, i);
  t.pos[0] = qreal(s, 0); t.pos[1] = qreal(s, 1); t.pos[2] = qreal(s, 2);
  t.dir[0] = qreal(s, 3); t.dir[1] = qreal(s, 4); t.dir[2] = qreal(s, 5);
  t.E = qreal(s, 6);
  const uint32_t tt = qu32(s, kQTT);
  t.type = (int)(tt & 3u);
  t.tag = tt >> 8;
  t.iter = 0;
  t.rng.k0 = P.k0; t.rng.k1 = P.k1;
  t.rng.hist = qu32(s, kQHist);
  t.rng.slo = qu32(s, kQSlo);
  t.rng.shi = qu32(s, kQShi);
  t.rng.ndraw = (tt >> 2) & 63u;
  t.rng.nchild = 0;
  // nothing buffered: the first top-up (or draw) computes block ndraw>>2, so a
  // stream resumed mid-block (draws taken by the source) continues in place
  t.rng.nblk = 0;
}

// PerformAnnihilation (SimDPMLike.cc:753-791): two back-to-back 0.511 photons
template <class BlockAcc>
__device__ __forceinline__ void annihilate(BlockAcc& a, const KParams<R>& P, Track& t, const OutQ<QueueView<R>>& qg) {
  count(a, kCntAnnih);
  const R cost = (R)1.0 - (R)2.0 * t.rng.uni<R>();
  const R sint = rsqrt_(((R)1.0 - cost) * ((R)1.0 + cost));
  const R phi = (R)(2.0 * kPI) * t.rng.uni<R>();
  R sp, cp;
  rsincos(phi, &sp, &cp);
  const R rx = sint * cp, ry = sint * sp, rz = cost;
  count(a, kCntSecondaries);
  push(qg, P.overflow, t.pos, rx, ry, rz, (R)kEMC2, t.rng.hist, t.rng.next_child(), 0, t.tag);
  count(a, kCntSecondaries);
  push(qg, P.overflow, t.pos, -rx, -ry, -rz, (R)kEMC2, t.rng.hist, t.rng.next_child(), 0, t.tag);
}

// ---------------------------------------------------------------------------
// primaries (SimDPMLike.cc:64-75, extended by fields, spectra and beamlets),
// generated by the transport kernels themselves when their input queue is
// drained: history index v of a kind's list -> entry by binary search, then
// the source draws (field position x, y; spectrum energy) from the history's
// own stream, in that order
// ---------------------------------------------------------------------------
// Gantry field (voxel phantom): the field point (fu, fv) lies on the plane
// through the phantom centre perpendicular to the beam direction
// d = (sa, 0, ca), with axes a = (ca, 0, -sa) and b = (0, 1, 0); the particle
// starts where the ray of direction d through it enters the phantom (slab
// method), 1e-3 mm inside. False if the ray misses the phantom.
__device__ __forceinline__ bool gantry_entry(R fu, R fv, R ca, R sa, const Geometry<R>& G, R pos[3], R dir[3]) {
  const R q[3] = {fu * ca, fv, -fu * sa};
  const R d[3] = {sa, (R)0, ca};
  const R h[3] = {-G.org[0], -G.org[1], -G.org[2]};
  R tin = (R)-1e30, tout = (R)1e30;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    if (d[i] == (R)0) {
      if (q[i] <= -h[i] || q[i] >= h[i]) return false;
      continue;
    }
    const R t1 = (-h[i] - q[i]) / d[i], t2 = (h[i] - q[i]) / d[i];
    tin = rmax(tin, rmin(t1, t2));
    tout = rmin(tout, rmax(t1, t2));
  }
  if (!(tin + (R)2e-3 < tout)) return false;
  const R t = tin + (R)1e-3;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    pos[i] = q[i] + t * d[i];
    dir[i] = d[i];
  }
  return true;
}

// Collimated gantry field: the field point q = (fu ca, fv, -fu sa) on the
// isocentre plane; with sad > 0 the particle comes from the point source
// s = -sad d and travels along r = (q - s) / |q - s|, else along d. With jaws,
// a particle aimed outside the aperture ([jx1, jx2) x [jy1, jy2) of the field
// point relative to the field centre, (du, dv)) crosses one jaw per pair it
// misses (n = 1, 2) and goes on with probability exp(-n jtau / cos theta),
// cos theta = r.d, drawing one uniform; otherwise it is absorbed. The start
// point is where the line q + t r enters the phantom, ahead of the source
// (t >= -|q - s|), 1e-3 mm inside. False if absorbed or if the ray misses.
__device__ __forceinline__ bool collimated_entry(R fu, R fv, R du, R dv, R ca, R sa, const SourceParams* __restrict__ S,
                                                 const Geometry<R>& G, Rng& rng, R pos[3], R dir[3]) {
  const R q[3] = {fu * ca, fv, -fu * sa};
  const R d[3] = {sa, (R)0, ca};
  R r[3] = {d[0], d[1], d[2]};
  R tmin = (R)-1e30, cost = (R)1;
  if (S->sad > 0.0) {
    const R sad = (R)S->sad;
#pragma unroll
    for (int i = 0; i < 3; ++i) r[i] = q[i] + sad * d[i];
    const R len = rsqrt_(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
#pragma unroll
    for (int i = 0; i < 3; ++i) r[i] = r[i] / len;
    tmin = -len;
    cost = r[0] * d[0] + r[2] * d[2];
  }
  if (S->jaws) {
    const int n = (int)(du < (R)S->jx1 || du >= (R)S->jx2) + (int)(dv < (R)S->jy1 || dv >= (R)S->jy2);
    if (n > 0) {
      const R u = rng.uni<R>();
      if (!(u < rexp(-(R)n * (R)S->jtau / cost))) return false;
    }
  }
  const R h[3] = {-G.org[0], -G.org[1], -G.org[2]};
  R tin = (R)-1e30, tout = (R)1e30;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    if (r[i] == (R)0) {
      if (q[i] <= -h[i] || q[i] >= h[i]) return false;
      continue;
    }
    const R t1 = (-h[i] - q[i]) / r[i], t2 = (h[i] - q[i]) / r[i];
    tin = rmax(tin, rmin(t1, t2));
    tout = rmin(tout, rmax(t1, t2));
  }
  if (!(tin >= tmin && tin + (R)2e-3 < tout)) return false;
  const R t = tin + (R)1e-3;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    pos[i] = q[i] + t * r[i];
    dir[i] = r[i];
  }
  return true;
}

// inverse CDF of the triangular distribution (lo, mode, hi)
__device__ __forceinline__ R triangular(R u, R lo, R mode, R hi) {
  const R fc = (mode - lo) / (hi - lo);
  if (u < fc) return lo + rsqrt_(u * (hi - lo) * (mode - lo));
  return hi - rsqrt_(((R)1.0 - u) * (hi - lo) * (hi - mode));
}

// energy from a tabulated spectrum (histogram bins of relative weight):
// the bin whose cumulative-weight interval [cdf[i], cdf[i+1]) holds the
// uniform u = k 2^-24, then the same fraction of the bin's width (one uniform,
// in FP64 for every precision)
__device__ __forceinline__ double spectrum_table(int n, const double* __restrict__ edges,
                                                 const double* __restrict__ cdf, uint32_t k) {
  const double u = (double)k * (1.0 / 16777216.0);
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (cdf[mid] <= u) lo = mid;
    else hi = mid - 1;
  }
  const double f = (u - cdf[lo]) / (cdf[lo + 1] - cdf[lo]);
  return edges[lo] + f * (edges[lo + 1] - edges[lo]);
}

// A phase-space record outside the voxel phantom moves along its direction
// to where that ray enters the phantom (slab method), 1e-3 mm inside; false
// if the ray misses the phantom. A record inside or on a face stays where it
// is (a record on the source plane starts as the pencil beam does).
__device__ __forceinline__ bool phsp_entry(const Geometry<R>& G, R pos[3], const R dir[3]) {
  const R h[3] = {-G.org[0], -G.org[1], -G.org[2]};
  if (rabs(pos[0]) <= h[0] && rabs(pos[1]) <= h[1] && rabs(pos[2]) <= h[2]) return true;
  R tin = (R)-1e30, tout = (R)1e30;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    if (dir[i] == (R)0) {
      if (pos[i] <= -h[i] || pos[i] >= h[i]) return false;
      continue;
    }
    const R t1 = (-h[i] - pos[i]) / dir[i], t2 = (h[i] - pos[i]) / dir[i];
    tin = rmax(tin, rmin(t1, t2));
    tout = rmin(tout, rmax(t1, t2));
  }
  if (!(tin >= (R)0 && tin + (R)2e-3 < tout)) return false;
  const R t = tin + (R)1e-3;
#pragma unroll
  for (int i = 0; i < 3; ++i) pos[i] = pos[i] + t * dir[i];
  return true;
}

// returns false when a gantry ray or a phase-space record misses the phantom
// (t.E then holds the primary's energy, which the caller counts as escaped)
__device__ __forceinline__ bool gen_primary(const KParams<R>& P, const SourceParams* __restrict__ S,
                                            const PrimList& L, uint32_t batch_first, uint64_t v, Track& t,
                                            double& eprim, int k) {
  // beamlets: neighbouring lanes take histories of different entries, so that
  // one pencil beam's axis voxels do not receive a wave's worth of atomics
  if (S->perm[k]) v = (v * S->perm[k]) % S->count[k];
  int lo = 0, hi = L.n - 1;
  if (S->kind == 3) {
    lo = (int)v;   // one history per record
  } else {
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (L.cum[mid] <= v) lo = mid;
      else hi = mid - 1;
    }
  }
  const uint64_t hid = L.first[lo] + (v - L.cum[lo]);
  t.rng.k0 = P.k0; t.rng.k1 = P.k1;
  t.rng.hist = (uint32_t)hid;
  t.rng.slo = 0; t.rng.shi = 0;
  t.rng.ndraw = 0; t.rng.nchild = 0; t.rng.nblk = 0;
  // no top-up here: the first draw fills a block, and the transport loop
  // tops the buffer up at its next wave-uniform point
  const double x0 = L.x[lo], y0 = L.y[lo];
  t.pos[0] = (R)x0;
  t.pos[1] = (R)y0;
  t.pos[2] = S->kind == 3 ? (R)L.z[lo] : (R)S->z0;
  R du = (R)0, dv = (R)0;
  if (S->kind == 1 || S->kind == 2) {
    const R ux = t.rng.uni<R>(), uy = t.rng.uni<R>();
    du = (ux - (R)0.5) * (R)S->fx;
    dv = (uy - (R)0.5) * (R)S->fy;
    t.pos[0] = (R)x0 + du;
    t.pos[1] = (R)y0 + dv;
  }
  R E = (R)L.e[lo];
  if (S->kind == 3) {
    // a phase-space record: its own energy and direction, no source draws
  } else if (S->spectrum == 2)
    E = (R)spectrum_table(S->spec_n, S->spec_edges, S->spec_cdf, t.rng.draw24());
  else if (S->emax > S->emin)
    E = S->spectrum == 1 ? triangular(t.rng.uni<R>(), (R)S->emin, (R)S->emode, (R)S->emax)
                         : (R)S->emin + t.rng.uni<R>() * (R)(S->emax - S->emin);
  eprim += (double)E;
  t.dir[0] = (R)0; t.dir[1] = (R)0; t.dir[2] = (R)1;
  bool inside = true;
  if (S->kind == 2) {
    if (S->sad > 0.0 || S->jaws)
      inside = collimated_entry(t.pos[0], t.pos[1], du, dv, (R)L.ca[lo], (R)L.sa[lo], S, P.G, t.rng, t.pos, t.dir);
    else
      inside = gantry_entry(t.pos[0], t.pos[1], (R)L.ca[lo], (R)L.sa[lo], P.G, t.pos, t.dir);
  }
  if (S->kind == 3) {
    t.dir[0] = (R)L.u[lo]; t.dir[1] = (R)L.v[lo]; t.dir[2] = (R)L.w[lo];
    if (S->phsp_project) inside = phsp_entry(P.G, t.pos, t.dir);
  }
  int row = L.row[lo];
  if (row < 0)
    row = S->nbatch > 1
              ? (int)((uint32_t)(((hid - S->hist_begin) * (uint64_t)S->nbatch) / S->nhist_total) - batch_first)
              : 0;
  const int pt = L.particle[lo];
  t.type = pt == 0 ? 0 : (pt < 0 ? 1 : 2);
  t.tag = (uint32_t)row;
  t.iter = 0;
  t.E = E;
  return inside;
}

// wave sum of the lanes' primary energies into the block accumulator
template <class BlockAcc>
__device__ __forceinline__ void flush_eprim(BlockAcc& a, double eprim) {
  for (int o = 32; o > 0; o >>= 1) eprim += __shfl_xor(eprim, o);
  if (lane_id() == 0 && eprim != 0.0) atomicAdd(&a.eprim, eprim);
}

// The photon kernel's linear tables (SimDataLinear::GetValue, SimDataLinear.hh:
// 38-45, behind SimIMFPMaxPhoton.hh:36-40 and SimIMFPPhoton.hh:40-46), read
// from the device pack or from a copy in LDS (FP32 build, when the maximum and
// the per-material IMFPs share one energy grid): the grid, then the maximum
// IMFP, then the total, Compton and pair IMFPs of every material. Same
// arithmetic in both forms.
template <int N>
struct LdsLinear {
  float v[N];
};
// LT: 0 tables in HBM, 1 the small copy (kLdsLinearFloats: four blocks of
// the kernel fit in a CU), 2 the large one (kLdsLinearFloatsLarge: three)
template <int LT> struct LdsLinOf { using type = NoLds; };
template <> struct LdsLinOf<1> { using type = LdsLinear<kLdsLinearFloats>; };
template <> struct LdsLinOf<2> { using type = LdsLinear<kLdsLinearFloatsLarge>; };

template <bool LDS>
struct PTables {
  const Tables<R>& T;
  const float* L;

  __device__ __forceinline__ R value(const LinTab<R>& t, int ybase, R e, int imat) const {
    if (!LDS) return linear_value(t, e, imat);
    const R ee = rmin(t.emaxm, rmax(t.emin, e));
    int i = (int)((ee - t.minX) * t.invDelta);
    i = max(0, min(t.n - 2, i));
    const float* y = L + ybase + imat * t.n;
    const R x0 = (R)L[i], x1 = (R)L[i + 1], y0 = (R)y[i], y1 = (R)y[i + 1];
    return rmax((R)1.0E-20, (y1 - y0) * (ee - x0) / (x1 - x0) + y0);
  }
  __device__ __forceinline__ R gmax(R e) const { return value(T.gmax, T.gmax.n, e, 0); }
  __device__ __forceinline__ R total(R e, int m) const { return value(T.ptot, 2 * T.gmax.n, e, m); }
  __device__ __forceinline__ R compton(R e, int m) const {
    return value(T.pcomp, 2 * T.gmax.n + T.nmat * T.ptot.n, e, m);
  }
  __device__ __forceinline__ R pair(R e, int m) const {
    return value(T.ppair, 2 * T.gmax.n + 2 * T.nmat * T.ptot.n, e, m);
  }
};

// copy of the photon tables into the block's LDS (all threads)
__device__ __forceinline__ void load_lds_linear(const Tables<R>& T, float* L) {
  const int n = T.gmax.n, nm = T.nmat * n;
  const int total = 2 * n + 3 * nm;
  for (int i = threadIdx.x; i < total; i += blockDim.x) {
    float v;
    if (i < n) v = (float)T.gmax.xy[2 * i];
    else if (i < 2 * n) v = (float)T.gmax.xy[2 * (i - n) + 1];
    else if (i < 2 * n + nm) v = (float)T.ptot.xy[2 * (i - 2 * n) + 1];
    else if (i < 2 * n + 2 * nm) v = (float)T.pcomp.xy[2 * (i - 2 * n - nm) + 1];
    else v = (float)T.ppair.xy[2 * (i - 2 * n - 2 * nm) + 1];
    L[i] = v;
  }
}

// ---------------------------------------------------------------------------
// photons: KeepTrackingPhoton (SimDPMLike.cc:463-635), one Woodcock step per
// loop trip; e-/e+ go to qe (this round's electron queue), annihilation
// photons to qg (next round's photon queue)
// ---------------------------------------------------------------------------
template <int GEOM, bool LDS, class BlockAcc>
__device__ __forceinline__ bool photon_step(BlockAcc& a, const KParams<R>& P, const PTables<LDS>& PT, Track& t,
                                            const OutQ<QueueView<R>>& qe, const OutQ<QueueView<R>>& qg) {
  const Tables<R>& T = P.T;
  const R kInvEMC2 = (R)(1.0 / kEMC2);
  count(a, kCntGSteps);
  const R globalMaxMFP = (R)1.0 / (PT.gmax(t.E) * P.majScale);
  const R stepLength = -globalMaxMFP * t.rng.lnu<R>();
  t.pos[0] += t.dir[0] * stepLength;
  t.pos[1] += t.dir[1] * stepLength;
  t.pos[2] += t.dir[2] * stepLength;
  if (!locate<R, GEOM>(P.G, t.pos, t.dir, t.box)) { escape(a, t.E); return false; }
  count(a, kCntVoxelVisits);
  const int imat = material_index<R, GEOM>(P.G, t.box);
  if (imat < 0) { escape(a, t.E); return false; }
  const R rho = voxel_density<R, GEOM>(P.G, imat, t.box);
  const R totalIMFP = PT.total(t.E, imat) * rho;
  const R r1 = t.rng.uni<R>();
  R theProb = (R)1.0 - totalIMFP * globalMaxMFP;
  if (r1 < theProb) return true;
  theProb += PT.compton(t.E, imat) * rho * globalMaxMFP;
  if (r1 < theProb) {
    // Compton (SimDPMLike.cc:512-570)
    count(a, kCntKN);
    const uint32_t q1 = t.rng.draw24(), q2 = t.rng.draw24(), q3 = t.rng.draw24();
    const R theEps = sample_kn(T.kn, t.E, q1, q2, q3, (int*)nullptr, (int*)nullptr);
    const R kappa = t.E * kInvEMC2;
    const R phCost = (R)1.0 - ((R)1.0 - theEps) / (theEps * kappa);
    const R phEner = theEps * t.E;
    const R phPhi = (R)(2.0 * kPI) * t.rng.uni<R>();
    const R elEner = t.E - phEner;
    if (elEner < T.ecut) {
      score<GEOM>(a, P, elEner, t.box, t.tag, t.rng.hist);
    } else {
      const R e0 = t.E * kInvEMC2;
      const R elCost = ((R)1.0 + e0) * rsqrt_(((R)1.0 - theEps) / (e0 * ((R)2.0 + e0 * ((R)1.0 - theEps))));
      // FP32 rounding can put elCost a few ulp above 1 at eps = eps_min:
      // the clamp keeps the sine real and leaves FP64 values unchanged
      const R sint = rsqrt_(rmax((R)0.0, ((R)1.0 + elCost) * ((R)1.0 - elCost)));
      const R phi = phPhi + (R)kPI;
      R sp, cp;
      rsincos(phi, &sp, &cp);
      R du = sint * cp, dv = sint * sp, dw = elCost;
      rotate_to_lab(du, dv, dw, t.dir[0], t.dir[1], t.dir[2]);
      count(a, kCntSecondaries);
      push<true>(qe, P.overflow, t.pos, du, dv, dw, elEner, t.rng.hist, t.rng.next_child(), 1, t.tag);
    }
    t.E = phEner;
    if (t.E < T.gcut) {
      score<GEOM>(a, P, t.E, t.box, t.tag, t.rng.hist);
      return false;
    }
    const R phSint = rsqrt_(rmax((R)0.0, ((R)1.0 - phCost) * ((R)1.0 + phCost)));
    R sp, cp;
    rsincos(phPhi, &sp, &cp);
    R u1 = phSint * cp, u2 = phSint * sp, u3 = phCost;
    rotate_to_lab(u1, u2, u3, t.dir[0], t.dir[1], t.dir[2]);
    t.dir[0] = u1; t.dir[1] = u2; t.dir[2] = u3;
    return true;
  }
  theProb += PT.pair(t.E, imat) * rho * globalMaxMFP;
  if (r1 < theProb) {
    // pair production (SimDPMLike.cc:573-627)
    count(a, kCntPair);
    const R sumEkin = t.E - (R)(2.0 * kEMC2);
    const R e1 = t.rng.uni<R>() * sumEkin;
    const R e2 = sumEkin - e1;
    if (e2 < T.ecut) {
      score<GEOM>(a, P, e2, t.box, t.tag, t.rng.hist);
    } else {
      count(a, kCntSecondaries);
      push(qe, P.overflow, t.pos, t.dir[0], t.dir[1], t.dir[2], e2, t.rng.hist, t.rng.next_child(), 1, t.tag);
    }
    if (e1 < T.ecut) {
      score<GEOM>(a, P, e1, t.box, t.tag, t.rng.hist);
      annihilate(a, P, t, qg);
    } else {
      count(a, kCntSecondaries);
      push(qe, P.overflow, t.pos, t.dir[0], t.dir[1], t.dir[2], e1, t.rng.hist, t.rng.next_child(), 2, t.tag);
    }
    return false;
  }
  // photoelectric absorption (SimDPMLike.cc:629-633)
  score<GEOM>(a, P, t.E, t.box, t.tag, t.rng.hist);
  return false;
}

// start of a popped track (SimDPMLike.cc:86-95); false if it ends here
template <int GEOM, class BlockAcc>
__device__ __forceinline__ bool start_track(BlockAcc& a, const KParams<R>& P, Track& t) {
  if (t.type == (int)kEmptySlot) return false;
  if (!locate<R, GEOM>(P.G, t.pos, t.dir, t.box)) box_of<R, GEOM>(P.G, t.pos, t.box);
  if constexpr (GEOM == GEOM_VOXEL) {
    // every voxel holds a valid material (checked when the phantom was
    // created), so the material lookup fails exactly outside the phantom:
    // a bounds test gives the same answer without a dependent load
    const bool inside = t.box[0] >= 0 && t.box[1] >= 0 && t.box[2] >= 0 && t.box[0] < P.G.nx &&
                        t.box[1] < P.G.ny && t.box[2] < P.G.nz;
    if (!inside) { escape(a, t.E); return false; }
    return true;
  }
  count(a, kCntVoxelVisits);
  const int imat = material_index<R, GEOM>(P.G, t.box);
  if (imat < 0) { escape(a, t.E); return false; }
  return true;
}

// Input queues are consumed in per-wave ranges of kBlk entries: a wave takes a
// range from the queue head with one atomic and hands its entries to lanes
// whose particle has ended (ballot rank), so the head sees one atomic per
// kBlk particles. Runs with the whole wave converged (loop top).
struct InRange {
  int next, end;
};

// The indices are handed out first (over as many ranges as it takes) and the
// start code then runs once for all refilled lanes.
template <class Start>
__device__ __forceinline__ void refill(bool& has, bool& exhausted, InRange& r, int* head, int n_in, Start start) {
  uint64_t need = __ballot(!has);
  int idx = -1;
  while (need != 0ull && !exhausted) {
    if (r.next >= r.end) {
      int b = 0;
      if (lane_id() == 0) b = atomicAdd(head, kBlk);
      b = __shfl(b, 0);
      if (b >= n_in) { exhausted = true; break; }
      r.next = b;
      r.end = min(b + kBlk, n_in);
    }
    const int avail = r.end - r.next;
    const bool mine = ((need >> lane_id()) & 1ull) != 0ull;
    const int rank = (int)__popcll(need & lanes_below());
    const bool take = mine && rank < avail;
    if (take) idx = r.next + rank;
    r.next += min(avail, (int)__popcll(need));
    need &= ~__ballot(take);
  }
  if (idx >= 0) has = start(idx);
}

template <int GEOM, int LT>
__global__ void __launch_bounds__(kBlock) DPM_PHOTON_WAVES
k_photons(KParams<R> P, const SourceParams* __restrict__ S, PrimRange pr, QueueView<R> in, int* head, QueueView<R> qe,
          QueueView<R> qg) {
  // a launch with less work than the grid holds lanes keeps only the blocks
  // it needs (the others would find the input exhausted); an empty drain
  // round costs one load of the queue count per block
  const InCount cin = in_count(in);
  const int n_in = cin.n;
  const int n_all = n_in + (int)(pr.end - pr.beg);
  if ((int64_t)blockIdx.x * kBlock >= (int64_t)n_all) return;
  __shared__ typename AccOf<GEOM>::type acc;
  __shared__ int wout[kBlock / 64][3][2];
  constexpr bool LDS = LT != 0;
  __shared__ typename LdsLinOf<LT>::type lin;
  if constexpr (LDS) load_lds_linear(P.T, lin.v);   // made visible by acc_init's barrier
  PTables<LDS> PT{P.T, nullptr};
  if constexpr (LDS) PT.L = lin.v;
  const int wave = threadIdx.x >> 6;
  if (lane_id() == 0) {
    wout[wave][0][0] = 0; wout[wave][0][1] = kBlk;
    wout[wave][1][0] = 0; wout[wave][1][1] = kBlk;
    wout[wave][2][0] = 0; wout[wave][2][1] = kBlk;
  }
  acc_init(acc, GEOM == GEOM_SLAB);
  const OutQ<QueueView<R>> oe{qe, &wout[wave][0][0], &wout[wave][2][0]};
  const OutQ<QueueView<R>> og{qg, &wout[wave][1][0], nullptr};
  Track t;
  double eprim = 0.0;
  bool has = false, exhausted = false;
  InRange rng_in{0, 0};
  uint32_t trips = 0;   // transport-loop trips of this wave (wave-uniform)
  for (;;) {
    // compaction interval: refill once refill_min lanes are idle (or none is busy)
    if (!exhausted && (__popcll(__ballot(!has)) >= P.refill_min || !__any(has)))
      refill(has, exhausted, rng_in, head, n_all, [&](int idx) {
        if (idx < n_in) load_track(in, in_slot(in, cin, idx), P, t);
        else if (!gen_primary(P, S, S->list[0], pr.batch_first, pr.beg + (uint64_t)(idx - n_in), t, eprim, 0)) {
          escape(acc, t.E);
          return false;
        }
        // the buffer is filled by the step's top-up (one pass over the
        // wave for new and running lanes alike)
        return start_track<GEOM>(acc, P, t);
      });
    if (!__any(has)) {
      if (exhausted) break;
      continue;
    }
    ++trips;   // the wave is converged here
    if (has) {
      t.rng.topup();
      has = photon_step<GEOM, LDS>(acc, P, PT, t, oe, og);
      if (has && ++t.iter > kMaxIter) {
        count(acc, kCntGuard);
        escape(acc, t.E);
        has = false;
      }
    }
  }
  if (lane_id() == 0) atomicAdd(&acc.cnt[kCntLaneSlots], 64u * trips);
  flush_eprim(acc, eprim);
  close_block(oe);
  close_block(og);
  acc_flush(acc, P, GEOM == GEOM_SLAB);
}

// ---------------------------------------------------------------------------
// electrons / positrons: the loop of Simulate (SimDPMLike.cc:113-245) around
// KeepTrackingElectron (SimDPMLike.cc:254-460), flattened so that one trip of
// the kernel loop is one condensed-history step plus its interaction
// ---------------------------------------------------------------------------
struct EState {
  R nTr1, nTr10, nMol, invMol, nBrem, E0;
  bool hinge;
};

// The five spline tables of the electron step (SimDataSpline, with the
// clamps of SimITr1MFPElastic.hh:35-47, SimStoppingPower.hh:36-54,
// SimIMFPBrem.hh:35-53, SimIMFPMoller.hh:35-47, SimMaxScatStrength.hh:39-51),
// read either from the device table pack or from a copy in LDS. The LDS copy
// (FP32 build) holds the node abscissae and (y, b, c, d) of ITr1, dE/dx and
// brem IMFP (nmat materials each), then Moller IMFP and K1; each block loads
// it once when the kernel starts, and a lookup is then one ds_read_b32 plus
// one ds_read_b128 instead of two loads through L1/L2. The arithmetic is the
// same in both forms, so they give identical results.
struct LdsSplines {
  float x[kLdsSplineNodes];
  float4 q[kLdsSplineNodes];
};
template <bool LDS> struct LdsOf { using type = NoLds; };
template <> struct LdsOf<true> { using type = LdsSplines; };

// LDS node offsets of the five tables (order above)
__device__ __forceinline__ int4 lds_spline_offsets(const Tables<R>& T) {
  const int o_dedx = T.nmat * T.itr1.n;
  const int o_brem = o_dedx + T.nmat * T.dedx.n;
  const int o_mol = o_brem + T.nmat * T.brem.n;
  return make_int4(o_dedx, o_brem, o_mol, o_mol + T.moller.n);
}

template <bool LDS>
struct ESplines {
  const Tables<R>& T;
  const LdsSplines* L;
  int4 off;   // dE/dx, brem, Moller, K1 (ITr1 at 0)

  __device__ __forceinline__ R value(const SplineTab<R>& t, int base, R e, int imat) const {
    if (!LDS) return spline_value(t, e, imat);
    const R ee = rmin(t.emaxm, rmax(t.emin, e));
    const R lx = rlog(ee);
    int i = (int)((lx - t.logMinX) * t.invLogDelta);
    i = max(0, min(t.n - 2, i));
    const int j = base + imat * t.n + i;
    const float x = L->x[j];
    const float4 q = L->q[j];
    const R dx = ee - (R)x;
    return rmax((R)1.0E-20, (R)q.x + dx * ((R)q.y + dx * ((R)q.z + dx * (R)q.w)));
  }
  __device__ __forceinline__ R itr1(R e, int m) const { return value(T.itr1, 0, e, m); }
  __device__ __forceinline__ R dedx(R e, int m) const { return value(T.dedx, off.x, e, m); }
  __device__ __forceinline__ R brem(R e, int m) const { return value(T.brem, off.y, e, m); }
  __device__ __forceinline__ R moller(R e) const { return value(T.moller, off.z, e, 0); }
  __device__ __forceinline__ R k1(R e) const { return value(T.k1, off.w, e, 0); }
  // Moller IMFP or K1 through one evaluation site
  __device__ __forceinline__ R moller_or_k1(bool mol, R e) const {
    if (!LDS) return spline_value(pick_spline(mol, T.moller, T.k1), e, 0);
    return value(pick_spline(mol, T.moller, T.k1), mol ? off.z : off.w, e, 0);
  }
};

// copy of the five spline tables into the block's LDS (all threads)
__device__ __forceinline__ void load_lds_splines(const Tables<R>& T, LdsSplines& L) {
  const int4 o = lds_spline_offsets(T);
  const int total = o.w + T.k1.n;
  for (int i = threadIdx.x; i < total; i += blockDim.x) {
    const SplineNode<R>* nd;
    if (i < o.x) nd = T.itr1.node + i;
    else if (i < o.y) nd = T.dedx.node + (i - o.x);
    else if (i < o.z) nd = T.brem.node + (i - o.y);
    else if (i < o.w) nd = T.moller.node + (i - o.z);
    else nd = T.k1.node + (i - o.w);
    L.x[i] = (float)nd->x;
    L.q[i] = make_float4((float)nd->y, (float)nd->b, (float)nd->c, (float)nd->d);
  }
}

template <int GEOM, bool LDS, class BlockAcc>
__device__ __forceinline__ bool electron_init(BlockAcc& a, const KParams<R>& P, const ESplines<LDS>& SP, Track& t,
                                              EState& s) {
  if constexpr (GEOM == GEOM_VOXEL) {
    // The DistanceToBoundary call at a track's start (SimDPMLike.cc:86-95)
    // only pushes a position lying on a face and tests the phantom's bounds;
    // the step's own call (SimDPMLike.cc:257), made next on the same
    // position, does the same push and the same test (escape with the
    // track's energy), so in a voxel phantom it is left to the step.
    if (t.type == (int)kEmptySlot) return false;
  } else if (!start_track<GEOM>(a, P, t)) {
    return false;
  }
  s.hinge = true;
  s.E0 = t.E;
  s.nTr1 = SP.k1(t.E);
  s.nTr10 = t.rng.uni<R>() * s.nTr1;
  s.nTr1 -= s.nTr10;
  s.nMol = -t.rng.lnu<R>();
  s.invMol = SP.moller(t.E);
  s.nBrem = -t.rng.lnu<R>();
  return true;
}

template <int GEOM, bool LDS, class BlockAcc, class Hot>
__device__ __forceinline__ bool electron_step(BlockAcc& a, const KParams<R>& P, const ESplines<LDS>& SP, Track& t,
                                              EState& s, WComb& wc, Hot& H, int wave, const OutQ<QueueView<R>>& qe,
                                              const OutQ<QueueView<R>>& qg) {
  const Tables<R>& T = P.T;
  // distance to boundary at the start of every step (SimDPMLike.cc:257, 451)
  const R stepGeom = distance_to_boundary<R, GEOM>(P.G, t.pos, t.dir, t.box);
  if (stepGeom < (R)0.0) { escape(a, t.E); return false; }
  count2(a, kCntESteps, kCntVoxelVisits);
  int imat = material_index<R, GEOM>(P.G, t.box);
  if (imat < 0) { escape(a, t.E); return false; }
  const R ecut = T.ecut;
  const R scalMolMFP = T.mscale[imat];
  const R rho = voxel_density<R, GEOM>(P.G, imat, t.box);
  R stepLength = stepGeom;
  int what = 0;
  R theDEDX = SP.dedx(t.E, imat) * rho;
  R midStepE = rmax(t.E - (R)0.5 * stepLength * theDEDX, ecut);
  R delNumTr1MFP = SP.itr1(midStepE, imat) * rho;
  const R delNumMollerMFP = s.invMol * scalMolMFP * rho;
  const R delNumBremMFP = SP.brem(midStepE, imat) * rho;
  const R stepBrem = s.nBrem / delNumBremMFP;
  if (stepBrem < stepLength) { stepLength = stepBrem; what = 1; }
  const R stepMoller = s.nMol / delNumMollerMFP;
  if (stepMoller < stepLength) { stepLength = stepMoller; what = 2; }
  R stepElastic = s.nTr1 / delNumTr1MFP;
  if (stepElastic < stepLength) {
    stepElastic = s.nTr1 / (SP.itr1(t.E, imat) * rho);
    midStepE = rmax(t.E - (R)0.5 * stepElastic * theDEDX, ecut);
    delNumTr1MFP = SP.itr1(midStepE, imat) * rho;
    stepElastic = rmin(stepLength, s.nTr1 / delNumTr1MFP);
    stepLength = stepElastic;
    what = 3;
  }
  s.nBrem = (what == 1) ? (R)0 : s.nBrem - stepLength * delNumBremMFP;
  s.nMol = (what == 2) ? (R)0 : s.nMol - stepLength * delNumMollerMFP;
  s.nTr1 = (what == 3) ? (R)0 : s.nTr1 - stepLength * delNumTr1MFP;
  midStepE = rmax(t.E - (R)0.5 * stepLength * theDEDX, ecut);
  theDEDX = SP.dedx(midStepE, imat) * rho;
  const R deltE = stepLength * theDEDX;
  const R eFinal = t.E - deltE;
  R edep;
  if (eFinal < ecut) {
    edep = t.E;
    t.E = (R)0.0;
    if (t.type == 2) {
      // the stopping e+ annihilates one step ahead (SimDPMLike.cc:406-421)
      t.pos[0] += t.dir[0] * stepLength;
      t.pos[1] += t.dir[1] * stepLength;
      t.pos[2] += t.dir[2] * stepLength;
      locate<R, GEOM>(P.G, t.pos, t.dir, t.box);
      count(a, kCntVoxelVisits);
      imat = material_index<R, GEOM>(P.G, t.box);
      if (imat < 0) { escape(a, edep); return false; }
      annihilate(a, P, t, qg);
    }
    what = 4;
  } else {
    t.E = eFinal;
    edep = deltE;
  }
  t.pos[0] += t.dir[0] * stepLength;
  t.pos[1] += t.dir[1] * stepLength;
  t.pos[2] += t.dir[2] * stepLength;
  score_wc<GEOM>(a, P, wc, H, wave, edep, t.box, t.tag, t.rng.hist);
  if (what == 0) return true;
  if (what == 4) return false;
  // The interaction switch of Simulate (SimDPMLike.cc:169-244). PerformBrem
  // (:661-695), PerformMoller (:699-728) and PerformMSCAngularDeflection
  // (:731-750) share one pass through the log-energy table pick, one alias
  // or GS inversion, the azimuth, the rotation to the lab frame and the #mfp
  // re-sample, so a wave whose lanes took different interactions runs each
  // part once. Every lane still draws its own uniforms in the reference's
  // order: sampler (3 for SB/Moller, 2 for GS), azimuth, then re-sample.
  const bool brem = what == 1 && t.E > T.gcut;
  const bool mol = what == 2 && t.E > (R)2. * ecut;
  const bool hinge = what == 3 && s.hinge;
  if (brem || mol || hinge) {
    const uint32_t k1 = t.rng.draw24();
    const uint32_t k2 = t.rng.draw24();
    const uint32_t k3 = hinge ? 0u : t.rng.draw24();
    const R es = hinge ? s.E0 : t.E;  // GS is sampled at the pre-step energy (:220)
    const R lpe = rlog(es);
    const R logMinE = brem ? T.sb.logMinE : (mol ? T.mo.logMinE : T.gs.logMinE);
    const R invLogDelta = brem ? T.sb.invLogDelta : (mol ? T.mo.invLogDelta : T.gs.invLogDelta);
    const double logMinE_d = brem ? T.sb.logMinE_d : (mol ? T.mo.logMinE_d : T.gs.logMinE_d);
    const double invLogDelta_d = brem ? T.sb.invLogDelta_d : (mol ? T.mo.invLogDelta_d : T.gs.invLogDelta_d);
    int ielow = 0;
    const int ie = pick_energy_bin(lpe, logMinE, invLogDelta, logMinE_d, invLogDelta_d, es, k1, &ielow);
    R cost, q, esec = (R)0;
    // FP32 tables: the first loads of both samplers are issued together for
    // every interacting lane -- the two GS entries of the cumulative bin (MSC
    // hinge) or the alias entry of the bin and its successor, which are the
    // pdf's end points when the bin is kept (SB / Moller) -- so one memory
    // round trip serves the wave; a lane whose bin is aliased loads its two
    // entries after the test, as alias_sample does. Same values, same
    // arithmetic as gs_rational / alias_sample.
    AliasEntry<R> aa, ab;
    GSEntry<R> g0, g1;
    R ga0 = (R)0, ga1 = (R)0, ge0 = (R)0, ge1 = (R)0;
    const AliasEntry<R>* tab = brem ? T.sb.e + (size_t)(imat * T.sb.ne + ie) * T.sb.ns : T.mo.e + (size_t)ie * T.mo.ns;
    if constexpr (sizeof(R) == 4) {
      const int nsel = hinge ? T.gs.ns : (brem ? T.sb.ns : T.mo.ns);
      const uint32_t rest = k2 * (uint32_t)(nsel - 1);
      const int i0 = (int)(rest >> 24);
      const uint4* src = hinge ? (const uint4*)(T.gs.e + ie * T.gs.ns + i0) : (const uint4*)(tab + i0);
      const uint4 w0 = src[0], w1 = src[1];
      if (hinge) {
        ga0 = T.gs.transf[ielow]; ga1 = T.gs.transf[ielow + 1];
        ge0 = T.gs.egrid[ielow]; ge1 = T.gs.egrid[ielow + 1];
        g0 = *reinterpret_cast<const GSEntry<R>*>(&w0);
        g1 = *reinterpret_cast<const GSEntry<R>*>(&w1);
      } else {
        aa = *reinterpret_cast<const AliasEntry<R>*>(&w0);
        ab = *reinterpret_cast<const AliasEntry<R>*>(&w1);
        if ((rest & 0xFFFFFFu) >= aa.thr) {
          const uint32_t ia = aa.alias;
          aa = tab[ia];
          ab = tab[ia + 1];
        }
      }
    }
    if (hinge) {
      R omu;
      if constexpr (sizeof(R) == 4) omu = gs_invert(g0, g1, ga0, ga1, ge0, ge1, T.gs.deltaCum, T.gs.ns, k2, es);
      else omu = gs_rational(T.gs, ie, ielow, es, k2, (int*)nullptr);
      if (sizeof(R) == 8) {
        cost = rmax((R)-1.0, rmin((R)1.0, (R)1. - omu));
        q = ((R)1.0 - cost) * ((R)1.0 + cost);
      } else {
        const R o = rmax((R)0.0, rmin((R)2.0, omu));
        cost = (R)1.0 - o;
        q = o * ((R)2.0 - o);
      }
    } else {
      // SimSBTables.cc:93-114 (photon energy) or SimMollerTables.cc:69-89
      const int ns = brem ? T.sb.ns : T.mo.ns;
      const R minE = brem ? T.sb.minE : T.mo.minE;
      R xi;
      if constexpr (sizeof(R) == 4) xi = alias_invert(aa, ab, k3);
      else xi = alias_sample(tab, ns, k2, k3, (int*)nullptr);
      if (sizeof(R) == 8) {
        const R ex = rexp(brem ? -xi * (logMinE - lpe) : xi * (lpe - logMinE));
        esec = brem ? minE / es * ex * es : ex * (R)0.5 * minE;
      } else {
        const R ex = transfer_exp((float)xi, (float)lpe, logMinE_d);
        esec = brem ? minE * ex : ex * (R)0.5 * minE;
      }
      if (brem) {
        // photon angle of the DPM mean-angle approximation (:673-684)
        const R dum0 = (R)(kEMC2 * 0.7071067812) / (t.E + (R)kEMC2);
        cost = rmax((R)-1.0, rmin((R)1.0, (R)1.0 - dum0 * dum0));
      } else {
        const R k2EMC2 = (R)(2.0 * kEMC2);
        cost = rmin((R)1.0, rsqrt_(esec * (t.E + k2EMC2) / (t.E * (esec + k2EMC2))));
      }
      q = ((R)1.0 + cost) * ((R)1.0 - cost);
    }
    const R sint = rsqrt_(q);
    const R phi = (R)(2.0 * kPI) * t.rng.uni<R>();
    R sp, cp;
    rsincos(phi, &sp, &cp);
    R du = sint * cp, dv = sint * sp, dw = cost;
    rotate_to_lab(du, dv, dw, t.dir[0], t.dir[1], t.dir[2]);
    if (hinge) {
      count(a, kCntGS);
      t.dir[0] = du; t.dir[1] = dv; t.dir[2] = dw;
      s.nTr1 = s.nTr10;
      s.hinge = false;
    } else {
      if (brem) count(a, kCntSB);  // count() needs one counter per call site
      else count(a, kCntMoller);
      count(a, kCntSecondaries);
      if (brem) push(qg, P.overflow, t.pos, du, dv, dw, esec, t.rng.hist, t.rng.next_child(), 0, t.tag);
      else push(qe, P.overflow, t.pos, du, dv, dw, esec, t.rng.hist, t.rng.next_child(), 1, t.tag);
      t.E = t.E - esec;
    }
  }
  if (what == 1 && t.E < ecut) {
    // below the cut after the brem: deposit the rest (:179-188)
    score_wc<GEOM>(a, P, wc, H, wave, t.E, t.box, t.tag, t.rng.hist);
    t.E = (R)0.0;
    if (t.type == 2) annihilate(a, P, t, qg);
    return false;
  }
  if (what == 1 || what == 2) {
    const R n = -t.rng.lnu<R>();
    if (what == 1) s.nBrem = n;
    else s.nMol = n;
  }
  // Moller IMFP after a Moller step, or K1 at the end of an MSC step: one
  // spline evaluation at the current energy (both tables are material-free)
  if (what == 2 || (what == 3 && !hinge)) {
    const R v = SP.moller_or_k1(what == 2, t.E);
    if (what == 2) {
      s.invMol = v;
    } else {
      s.E0 = t.E;
      s.nTr1 = v;
      s.nTr10 = t.rng.uni<R>() * s.nTr1;
      s.nTr1 -= s.nTr10;
      s.hinge = true;
    }
  }
  return true;
}

template <int GEOM, bool LDS, bool HOT>
__global__ void __launch_bounds__(kBlock) DPM_ELECTRON_WAVES
k_electrons(KParams<R> P, const SourceParams* __restrict__ S, PrimRange pr, QueueView<R> in, int* head,
            QueueView<R> qe, QueueView<R> qg) {
  // blocks beyond the launch's work leave at once (see k_photons)
  const InCount cin = in_count(in);
  const int n_in = cin.n;
  const int n_all = n_in + (int)(pr.end - pr.beg);
  if ((int64_t)blockIdx.x * kBlock >= (int64_t)n_all) return;
  __shared__ typename AccOf<GEOM>::type acc;
  __shared__ int wout[kBlock / 64][3][2];
  __shared__ typename LdsOf<LDS>::type lsp;
  if constexpr (LDS) load_lds_splines(P.T, lsp);   // made visible by acc_init's barrier
  ESplines<LDS> SP{P.T, nullptr, make_int4(0, 0, 0, 0)};
  if constexpr (LDS) {
    SP.L = &lsp;
    SP.off = lds_spline_offsets(P.T);
  }
  const int wave = threadIdx.x >> 6;
  __shared__ typename HotOf<HOT>::type hot;
  if constexpr (HOT) hot_init(hot, wave);
  if (lane_id() == 0) {
    wout[wave][0][0] = 0; wout[wave][0][1] = kBlk;
    wout[wave][1][0] = 0; wout[wave][1][1] = kBlk;
    wout[wave][2][0] = 0; wout[wave][2][1] = kBlk;
  }
  acc_init(acc, GEOM == GEOM_SLAB);
  const OutQ<QueueView<R>> oe{qe, &wout[wave][0][0], nullptr};   // Moller electrons: front run only
  const OutQ<QueueView<R>> og{qg, &wout[wave][1][0], nullptr};
  Track t;
  EState s;
  WComb wc{-1, 0.0};
  double eprim = 0.0;
  bool has = false, exhausted = false;
  InRange rng_in{0, 0};
  uint32_t trips = 0;   // transport-loop trips of this wave (wave-uniform)
  for (;;) {
    if (!exhausted && (__popcll(__ballot(!has)) >= P.refill_min || !__any(has)))
      refill(has, exhausted, rng_in, head, n_all, [&](int idx) {
        if (idx < n_in) load_track(in, in_slot(in, cin, idx), P, t);
        else if (!gen_primary(P, S, S->list[1], pr.batch_first, pr.beg + (uint64_t)(idx - n_in), t, eprim, 1)) {
          escape(acc, t.E);
          return false;
        }
        // electron_init's first draw fills one block; the step's top-up
        // adds the second
        return electron_init<GEOM, LDS>(acc, P, SP, t, s);
      });
    if (!__any(has)) {
      if (exhausted) break;
      continue;
    }
    ++trips;   // the wave is converged here
    if (has) {
      t.rng.topup();
      has = electron_step<GEOM, LDS>(acc, P, SP, t, s, wc, hot, wave, oe, og);
      if (has && ++t.iter > kMaxIter) {
        count(acc, kCntGuard);
        escape(acc, t.E);
        has = false;
      }
      if (!has) flush_wc(acc, P, wc, hot, wave);
    }
  }
  if constexpr (HOT) hot_flush(acc, P, hot, wave);
  if (lane_id() == 0) atomicAdd(&acc.cnt[kCntLaneSlots], 64u * trips);
  flush_eprim(acc, eprim);
  close_block(oe);
  close_block(og);
  acc_flush(acc, P, GEOM == GEOM_SLAB);
}

// ---------------------------------------------------------------------------
// samplers and interpolations with explicit inputs (parity entry points)
// ---------------------------------------------------------------------------
__global__ void k_sample(Tables<R> T, int kind, int n, const double* e, const int* imat, const uint32_t* r1,
                         const uint32_t* r2, const uint32_t* r3, double* out, int* ebin, int* abin) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const R E = (R)e[i];
  int eb = -1, ab = -1;
  double v = 0;
  if (kind == 0) v = (double)sample_sb(T.sb, E, imat[i], r1[i], r2[i], r3[i], &eb, &ab);
  else if (kind == 1) v = (double)sample_moller(T.mo, E, r1[i], r2[i], r3[i], &eb, &ab);
  else if (kind == 2) v = (double)sample_kn(T.kn, E, r1[i], r2[i], r3[i], &eb, &ab);
  else if (kind == 3) v = 1.0 - (double)sample_gs_omu(T.gs, E, r1[i], r2[i], &eb, &ab);
  else if (kind == 4) v = (double)sample_gs_omu(T.gs, E, r1[i], r2[i], &eb, &ab);
  out[i] = v;
  ebin[i] = eb;
  abin[i] = ab;
}

// Final states of single interactions of a primary of energy E moving along
// +z in material imat, one Philox stream per sample (key = seed, history =
// sample index): the GPU form of the reference's model tests
// (testAPI.cc:92-199). out_v / out_c per kind:
//   0 PerformBrem   (SimDPMLike.cc:661-695): photon energy, photon cos(theta)
//   1 PerformMoller (SimDPMLike.cc:699-728): secondary e- energy, its cos(theta)
//   2 PerformMSCAngularDeflection (SimDPMLike.cc:731-750): 1-cos(theta), cos(theta)
//   3 Compton       (SimDPMLike.cc:512-570): scattered photon energy, its cos(theta)
__global__ void k_perform(Tables<R> T, int kind, int n, double energy, int imat, uint32_t k0, uint32_t k1,
                          double* out_v, double* out_c) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  Rng rng;
  rng.k0 = k0; rng.k1 = k1;
  rng.hist = (uint32_t)i;
  rng.slo = 0; rng.shi = 0; rng.ndraw = 0; rng.nchild = 0; rng.nblk = 0;
  rng.topup();
  const R E = (R)energy;
  R v = (R)0, cost = (R)1, sint = (R)0;
  double omu_out = 0.0;
  if (kind == 0) {
    const uint32_t q1 = rng.draw24(), q2 = rng.draw24(), q3 = rng.draw24();
    v = sample_sb(T.sb, E, imat, q1, q2, q3, (int*)nullptr, (int*)nullptr);
    const R dum0 = (R)(kEMC2 * 0.7071067812) / (E + (R)kEMC2);
    cost = rmax((R)-1.0, rmin((R)1.0, (R)1.0 - dum0 * dum0));
    sint = rsqrt_(((R)1.0 + cost) * ((R)1.0 - cost));
  } else if (kind == 1) {
    const uint32_t q1 = rng.draw24(), q2 = rng.draw24(), q3 = rng.draw24();
    v = sample_moller(T.mo, E, q1, q2, q3, (int*)nullptr, (int*)nullptr);
    const R k2EMC2 = (R)(2.0 * kEMC2);
    cost = rmin((R)1.0, rsqrt_(v * (E + k2EMC2) / (E * (v + k2EMC2))));
    sint = rsqrt_(((R)1.0 + cost) * ((R)1.0 - cost));
  } else if (kind == 2) {
    const uint32_t q1 = rng.draw24(), q2 = rng.draw24();
    const R omu = sample_gs_omu(T.gs, E, q1, q2, (int*)nullptr, (int*)nullptr);
    const R o = rmax((R)0.0, rmin((R)2.0, omu));
    cost = (R)1.0 - o;
    sint = rsqrt_(o * ((R)2.0 - o));
    omu_out = (double)o;
  } else {
    const uint32_t q1 = rng.draw24(), q2 = rng.draw24(), q3 = rng.draw24();
    const R eps = sample_kn(T.kn, E, q1, q2, q3, (int*)nullptr, (int*)nullptr);
    cost = (R)1.0 - ((R)1.0 - eps) / (eps * E * (R)(1.0 / kEMC2));
    v = eps * E;
    sint = rsqrt_(rmax((R)0.0, ((R)1.0 - cost) * ((R)1.0 + cost)));
  }
  // azimuth and rotation to the frame of a primary along +z (the identity)
  const R phi = (R)(2.0 * kPI) * rng.uni<R>();
  R sp, cp;
  rsincos(phi, &sp, &cp);
  R du = sint * cp, dv = sint * sp, dw = cost;
  rotate_to_lab(du, dv, dw, (R)0, (R)0, (R)1);
  out_v[i] = kind == 2 ? omu_out : (double)v;
  out_c[i] = (double)dw;
}

__global__ void k_eval(Tables<R> T, int which, int n, const double* e, const int* imat, double* out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const R E = (R)e[i];
  const int m = imat ? imat[i] : 0;
  R v = 0;
  switch (which) {
    case 0: v = spline_value(T.itr1, E, m); break;
    case 1: v = spline_value(T.dedx, E, m); break;
    case 2: v = spline_value(T.brem, E, m); break;
    case 3: v = spline_value(T.moller, E, 0); break;
    case 4: v = spline_value(T.k1, E, 0); break;
    case 5: v = linear_value(T.gmax, E, 0); break;
    case 6: v = linear_value(T.ptot, E, m); break;
    case 7: v = linear_value(T.pcomp, E, m); break;
    case 8: v = linear_value(T.ppair, E, m); break;
    default: break;
  }
  out[i] = (double)v;
}

// ---------------------------------------------------------------------------
// launchers (host side of this precision)
// ---------------------------------------------------------------------------
#define DPM_CAT2(a, b) a##b
#define DPM_CAT(a, b) DPM_CAT2(a, b)
#if DPM_PREC == 32
#define DPM_SFX _f32
#else
#define DPM_SFX _f64
#endif

hipError_t DPM_CAT(launch_photons, DPM_SFX)(const KParams<R>& P, const SourceParams* S, const PrimRange& pr, int geom,
                                            int blocks, const QueueView<R>& in, int* head, const QueueView<R>& qe,
                                            const QueueView<R>& qg, hipStream_t st) {
#if DPM_PREC == 32
  if (P.lds_photon == 1) {
    if (geom == GEOM_SLAB)
      hipLaunchKernelGGL((k_photons<GEOM_SLAB, 1>), dim3(blocks), dim3(kBlock), 0, st, P, S, pr, in, head, qe, qg);
    else
      hipLaunchKernelGGL((k_photons<GEOM_VOXEL, 1>), dim3(blocks), dim3(kBlock), 0, st, P, S, pr, in, head, qe, qg);
    return hipGetLastError();
  }
  if (P.lds_photon == 2) {
    if (geom == GEOM_SLAB)
      hipLaunchKernelGGL((k_photons<GEOM_SLAB, 2>), dim3(blocks), dim3(kBlock), 0, st, P, S, pr, in, head, qe, qg);
    else
      hipLaunchKernelGGL((k_photons<GEOM_VOXEL, 2>), dim3(blocks), dim3(kBlock), 0, st, P, S, pr, in, head, qe, qg);
    return hipGetLastError();
  }
#endif
  if (geom == GEOM_SLAB)
    hipLaunchKernelGGL((k_photons<GEOM_SLAB, 0>), dim3(blocks), dim3(kBlock), 0, st, P, S, pr, in, head, qe, qg);
  else
    hipLaunchKernelGGL((k_photons<GEOM_VOXEL, 0>), dim3(blocks), dim3(kBlock), 0, st, P, S, pr, in, head, qe, qg);
  return hipGetLastError();
}

template <int GEOM, bool LDS, bool HOT>
static hipError_t launch_el(const KParams<R>& P, const SourceParams* S, const PrimRange& pr, int blocks,
                            const QueueView<R>& in, int* head, const QueueView<R>& qe, const QueueView<R>& qg,
                            hipStream_t st) {
  hipLaunchKernelGGL((k_electrons<GEOM, LDS, HOT>), dim3(blocks), dim3(kBlock), 0, st, P, S, pr, in, head, qe, qg);
  return hipGetLastError();
}

hipError_t DPM_CAT(launch_electrons, DPM_SFX)(const KParams<R>& P, const SourceParams* S, const PrimRange& pr,
                                              int geom, int blocks, const QueueView<R>& in, int* head,
                                              const QueueView<R>& qe, const QueueView<R>& qg, hipStream_t st) {
  // the slab tally is privatised in LDS by depth bin; tally mode 2 is for voxels
  const bool hot = geom == GEOM_VOXEL && P.tally_mode == 2;
#if DPM_PREC == 32
  if (P.lds_tables) {
    if (geom == GEOM_SLAB) return launch_el<GEOM_SLAB, true, false>(P, S, pr, blocks, in, head, qe, qg, st);
    return hot ? launch_el<GEOM_VOXEL, true, true>(P, S, pr, blocks, in, head, qe, qg, st)
               : launch_el<GEOM_VOXEL, true, false>(P, S, pr, blocks, in, head, qe, qg, st);
  }
#endif
  if (geom == GEOM_SLAB) return launch_el<GEOM_SLAB, false, false>(P, S, pr, blocks, in, head, qe, qg, st);
  return hot ? launch_el<GEOM_VOXEL, false, true>(P, S, pr, blocks, in, head, qe, qg, st)
             : launch_el<GEOM_VOXEL, false, false>(P, S, pr, blocks, in, head, qe, qg, st);
}

hipError_t DPM_CAT(launch_perform, DPM_SFX)(const Tables<R>& T, int kind, int n, double energy, int imat,
                                            uint64_t seed, double* out_v, double* out_c, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_perform, dim3((n + 255) / 256), dim3(256), 0, st, T, kind, n, energy, imat, (uint32_t)seed,
                     (uint32_t)(seed >> 32), out_v, out_c);
  return hipGetLastError();
}

hipError_t DPM_CAT(launch_sample, DPM_SFX)(const Tables<R>& T, int kind, int n, const double* e, const int* imat,
                                           const uint32_t* r1, const uint32_t* r2, const uint32_t* r3, double* out,
                                           int* ebin, int* abin, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_sample, dim3((n + 255) / 256), dim3(256), 0, st, T, kind, n, e, imat, r1, r2, r3, out, ebin,
                     abin);
  return hipGetLastError();
}

hipError_t DPM_CAT(launch_eval, DPM_SFX)(const Tables<R>& T, int which, int n, const double* e, const int* imat,
                                         double* out, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_eval, dim3((n + 255) / 256), dim3(256), 0, st, T, which, n, e, imat, out);
  return hipGetLastError();
}

}  // namespace dpm
