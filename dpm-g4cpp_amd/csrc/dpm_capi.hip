// dpm_capi.hip -- host engine and C ABI of the MI355X DPM transport library.
//
// Owns the device table packs, phantoms and queue workspaces, and drives the
// transport kernels (dpm_kernels.hip) chunk by chunk. See include/dpm_amd.h
// for the entry points and the reference interfaces they replace.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <deque>
#include <string>
#include <vector>

#include "dpm_amd.h"
#include "dpm_engine.hpp"
#include "../host/SimData.hh"

using namespace dpm;

// Device scratch of dpm_simulate_noalloc: the source description of a run
// (the first kScratchSrcBytes) and its four particle queues (the rest). One
// copy per GPU, held from the library's load on that GPU until the process
// ends. In host code `scratch` only names the symbol: the device address of a
// GPU's copy comes from hipGetSymbolAddress with that GPU current
// (dpm_tables_create keeps it in the tables handle).
constexpr size_t kScratchBytes = (size_t)128 << 20;
constexpr size_t kScratchSrcBytes = (size_t)64 << 10;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

thread_local std::string g_last_error;

int set_error(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

#define HIP_TRY(expr)                                                                              \
  do {                                                                                             \
    hipError_t e_ = (expr);                                                                        \
    if (e_ != hipSuccess)                                                                          \
      return set_error(DPM_E_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_));           \
  } while (0)

// integer keep-threshold of an alias entry: (w < f * 2^-24) <=> (f >= thr)
uint32_t alias_threshold(double w) {
  if (!(w >= 0.0)) return 0u;
  const double t = std::floor(std::ldexp(w, 24)) + 1.0;
  return t > 16777216.0 ? 16777216u : (uint32_t)t;
}

// byte buffer assembled on the host, uploaded once
struct Blob {
  std::vector<unsigned char> bytes;
  template <typename T>
  size_t add(const T* p, size_t n) {
    size_t off = (bytes.size() + 255) & ~(size_t)255;
    bytes.resize(off + n * sizeof(T));
    std::memcpy(bytes.data() + off, p, n * sizeof(T));
    return off;
  }
};

template <typename R>
struct PackBuilder {
  Blob blob;
  // offsets of every array; patched into pointers after the upload
  size_t o_itr1, o_dedx, o_brem, o_moller, o_k1, o_gmax, o_ptot, o_pcomp, o_ppair, o_sb, o_mo, o_kn, o_gs, o_gse,
      o_gst;
  Tables<R> T;

  size_t add_spline(const dpm_spline_desc* per, int nm, SplineTab<R>& t) {
    const int n = per[0].n;
    std::vector<SplineNode<R>> nodes((size_t)nm * n);
    for (int im = 0; im < nm; ++im) {
      std::vector<double> B, C, D;
      SimSplineCoefficients(per[im].x, per[im].y, n, B, C, D);
      for (int i = 0; i < n; ++i) {
        SplineNode<R>& nd = nodes[(size_t)im * n + i];
        nd.x = (R)per[im].x[i]; nd.y = (R)per[im].y[i];
        nd.b = (R)B[i]; nd.c = (R)C[i]; nd.d = (R)D[i];
        nd.pad0 = nd.pad1 = nd.pad2 = (R)0;
      }
    }
    t.n = n;
    const double lmin = std::log(per[0].x[0]);
    t.logMinX = (R)lmin;
    t.invLogDelta = (R)(1. / (std::log(per[0].x[1]) - lmin));
    t.emin = (R)per[0].x[0];
    t.emaxm = (R)(per[0].x[n - 1] - 1.0E-6);
    return blob.add(nodes.data(), nodes.size());
  }
  size_t add_linear(const dpm_linear_desc* per, int nm, LinTab<R>& t) {
    const int n = per[0].n;
    std::vector<R> xy((size_t)nm * n * 2);
    for (int im = 0; im < nm; ++im)
      for (int i = 0; i < n; ++i) {
        xy[2 * ((size_t)im * n + i)] = (R)per[im].x[i];
        xy[2 * ((size_t)im * n + i) + 1] = (R)per[im].y[i];
      }
    t.n = n;
    t.minX = (R)per[0].x[0];
    t.invDelta = (R)(1. / (per[0].x[1] - per[0].x[0]));
    t.emin = (R)per[0].x[0];
    t.emaxm = (R)(per[0].x[n - 1] - 1.0E-6);
    return blob.add(xy.data(), xy.size());
  }
  size_t add_alias(const dpm_alias_desc& d, AliasSet<R>& s) {
    const size_t n = (size_t)d.nmat * d.ne * d.ns;
    std::vector<AliasEntry<R>> e(n);
    for (size_t k = 0; k < n; ++k) {
      e[k].x = (R)d.x[k];
      e[k].y = (R)d.y[k];
      e[k].thr = alias_threshold(d.w[k]);
      // an alias is followed only when frac >= thr; keep it inside the table
      e[k].alias = (uint32_t)std::min(std::max(d.alias[k], 0), d.ns - 2);
    }
    s.ne = d.ne;
    s.ns = d.ns;
    s.minE = (R)d.energy[0];
    s.logMinE_d = std::log(d.energy[0]);
    s.invLogDelta_d = 1. / (std::log(d.energy[1]) - s.logMinE_d);
    s.logMinE = (R)s.logMinE_d;
    s.invLogDelta = (R)s.invLogDelta_d;
    return blob.add(e.data(), n);
  }
};

bool same_grid(const double* a, const double* b, int n) {
  for (int i = 0; i < n; ++i)
    if (a[i] != b[i]) return false;
  return true;
}

int validate(const dpm_table_desc* d) {
  if (!d) return set_error(DPM_E_INVALID, "null table descriptor");
  if (d->nmat < 1 || d->nmat > kMaxMat)
    return set_error(DPM_E_INVALID, "number of materials must be in [1, " + std::to_string(kMaxMat) + "]");
  const dpm_spline_desc* sp[3] = {d->itr1, d->dedx, d->brem};
  const dpm_linear_desc* ln[3] = {d->ptotal, d->pcompton, d->ppair};
  for (int k = 0; k < 3; ++k) {
    if (!sp[k] || !ln[k]) return set_error(DPM_E_INVALID, "missing per-material table");
    for (int im = 0; im < d->nmat; ++im) {
      if (sp[k][im].n < 2 || sp[k][im].n != sp[k][0].n || !sp[k][im].x || !sp[k][im].y)
        return set_error(DPM_E_INVALID, "per-material spline tables must share one grid size >= 2");
      if (!same_grid(sp[k][im].x, sp[k][0].x, sp[k][0].n))
        return set_error(DPM_E_INVALID, "per-material spline tables must share one energy grid");
      if (ln[k][im].n < 2 || ln[k][im].n != ln[k][0].n || !ln[k][im].x || !ln[k][im].y)
        return set_error(DPM_E_INVALID, "per-material photon tables must share one grid size >= 2");
      if (!same_grid(ln[k][im].x, ln[k][0].x, ln[k][0].n))
        return set_error(DPM_E_INVALID, "per-material photon tables must share one energy grid");
    }
  }
  if (d->moller.n < 2 || d->k1.n < 2 || d->gmax.n < 2) return set_error(DPM_E_INVALID, "reference table too short");
  const dpm_alias_desc* al[3] = {&d->sb, &d->moller_tab, &d->kn};
  for (int k = 0; k < 3; ++k) {
    if (al[k]->ne < 2 || al[k]->ns < 2 || al[k]->ns > 256 || !al[k]->x || !al[k]->energy)
      return set_error(DPM_E_INVALID, "malformed alias table set");
    const size_t n = (size_t)al[k]->nmat * al[k]->ne * al[k]->ns;
    for (size_t i = 0; i < n; ++i)
      if (al[k]->alias[i] < 0 || al[k]->alias[i] >= al[k]->ns)
        return set_error(DPM_E_INVALID, "alias index out of range");
  }
  if (d->sb.nmat != d->nmat) return set_error(DPM_E_INVALID, "SB tables must cover every material");
  if (d->gs.ne < 2 || d->gs.ns < 2) return set_error(DPM_E_INVALID, "malformed GS tables");
  return DPM_OK;
}

template <typename R>
int build_pack(const dpm_table_desc* d, Tables<R>& T, void** dev) {
  PackBuilder<R> pb;
  std::memset(&T, 0, sizeof(T));
  const size_t o_itr1 = pb.add_spline(d->itr1, d->nmat, T.itr1);
  const size_t o_dedx = pb.add_spline(d->dedx, d->nmat, T.dedx);
  const size_t o_brem = pb.add_spline(d->brem, d->nmat, T.brem);
  const size_t o_mol = pb.add_spline(&d->moller, 1, T.moller);
  const size_t o_k1 = pb.add_spline(&d->k1, 1, T.k1);
  const size_t o_gmax = pb.add_linear(&d->gmax, 1, T.gmax);
  const size_t o_ptot = pb.add_linear(d->ptotal, d->nmat, T.ptot);
  const size_t o_pcomp = pb.add_linear(d->pcompton, d->nmat, T.pcomp);
  const size_t o_ppair = pb.add_linear(d->ppair, d->nmat, T.ppair);
  const size_t o_sb = pb.add_alias(d->sb, T.sb);
  const size_t o_mo = pb.add_alias(d->moller_tab, T.mo);
  const size_t o_kn = pb.add_alias(d->kn, T.kn);
  // GS
  const dpm_gs_desc& g = d->gs;
  std::vector<GSEntry<R>> ge((size_t)g.ne * g.ns);
  for (size_t k = 0; k < ge.size(); ++k) {
    ge[k].u = (R)g.u[k]; ge[k].a = (R)g.a[k]; ge[k].b = (R)g.b[k]; ge[k].v = (R)(1.0 - g.u[k]);
  }
  std::vector<R> gen(g.ne), gtr(g.ne);
  for (int i = 0; i < g.ne; ++i) { gen[i] = (R)g.energy[i]; gtr[i] = (R)g.transf[i]; }
  const size_t o_gs = pb.blob.add(ge.data(), ge.size());
  const size_t o_gse = pb.blob.add(gen.data(), gen.size());
  const size_t o_gst = pb.blob.add(gtr.data(), gtr.size());
  T.gs.ne = g.ne;
  T.gs.ns = g.ns;
  T.gs.logMinE_d = std::log(g.energy[0]);
  T.gs.invLogDelta_d = 1. / (std::log(g.energy[1]) - T.gs.logMinE_d);
  T.gs.logMinE = (R)T.gs.logMinE_d;
  T.gs.invLogDelta = (R)T.gs.invLogDelta_d;
  T.gs.deltaCum = (R)(1.0 / (g.ns - 1));
  for (int im = 0; im < d->nmat; ++im) {
    T.rho[im] = (R)d->density[im];
    T.mscale[im] = (R)d->moller_scale[im];
  }
  T.ecut = (R)d->electron_cut;
  T.gcut = (R)d->gamma_cut;
  T.nmat = d->nmat;
  T.lin_shared_grid = (d->gmax.n == d->ptotal[0].n && same_grid(d->gmax.x, d->ptotal[0].x, d->gmax.n)) ? 1 : 0;
  // upload and patch pointers
  unsigned char* base = nullptr;
  HIP_TRY(hipMalloc(&base, pb.blob.bytes.size()));
  HIP_TRY(hipMemcpy(base, pb.blob.bytes.data(), pb.blob.bytes.size(), hipMemcpyHostToDevice));
  T.itr1.node = (const SplineNode<R>*)(base + o_itr1);
  T.dedx.node = (const SplineNode<R>*)(base + o_dedx);
  T.brem.node = (const SplineNode<R>*)(base + o_brem);
  T.moller.node = (const SplineNode<R>*)(base + o_mol);
  T.k1.node = (const SplineNode<R>*)(base + o_k1);
  T.gmax.xy = (const R*)(base + o_gmax);
  T.ptot.xy = (const R*)(base + o_ptot);
  T.pcomp.xy = (const R*)(base + o_pcomp);
  T.ppair.xy = (const R*)(base + o_ppair);
  T.sb.e = (const AliasEntry<R>*)(base + o_sb);
  T.mo.e = (const AliasEntry<R>*)(base + o_mo);
  T.kn.e = (const AliasEntry<R>*)(base + o_kn);
  T.gs.e = (const GSEntry<R>*)(base + o_gs);
  T.gs.egrid = (const R*)(base + o_gse);
  T.gs.transf = (const R*)(base + o_gst);
  *dev = base;
  return DPM_OK;
}

// ---------------------------------------------------------------------------
// queue workspace
// ---------------------------------------------------------------------------
template <typename R>
struct Workspace {
  void* buf = nullptr;
  size_t bytes = 0;
  int cap = 0;
  bool fixed = false;  // a view of memory owned elsewhere (see view()): never reallocated
  QueueView<R> q[4];   // G0, G1, E0, E1

  void release() {
    if (buf && !fixed) (void)hipFree(buf);
    buf = nullptr;
    bytes = 0;
    cap = 0;
  }
  // records each of the 4 queues holds when they share `total` bytes
  static int capacity_of(size_t total) {
    const size_t per = (total / 4) & ~(size_t)255;
    return (int)std::min<size_t>(per / kQBlockBytes<R> * kQLanes, (size_t)1 << 30);
  }
  // the 4 queues laid out in [base, base + total) without allocating
  void view(unsigned char* base, size_t total) {
    buf = base;
    bytes = total;
    fixed = true;
    cap = capacity_of(total);
    const size_t per = (total / 4) & ~(size_t)255;
    for (int k = 0; k < 4; ++k) {
      q[k].base = base + k * per;
      q[k].cap = cap;
    }
  }
  int ensure(int want, int* ctl) {
    if (cap >= want && buf) {
      for (int k = 0; k < 4; ++k) q[k].count = (unsigned long long*)ctl + k;
      return DPM_OK;
    }
    if (fixed)
      return set_error(DPM_E_NOMEM, "particle queues need " + std::to_string(want) + " records, the fixed workspace holds " +
                                        std::to_string(cap));
    release();
    want = (want + kQLanes - 1) / kQLanes * kQLanes;
    const size_t per = (size_t)(want / kQLanes) * kQBlockBytes<R>;
    const size_t total = 4 * ((per + 255) & ~(size_t)255);
    if (hipMalloc(&buf, total) != hipSuccess) {
      buf = nullptr;
      return set_error(DPM_E_NOMEM, "cannot allocate particle queues of " + std::to_string(total >> 20) + " MiB");
    }
    bytes = total;
    cap = want;
    unsigned char* p = (unsigned char*)buf;
    auto take = [&](size_t n) {
      unsigned char* r = p;
      p += (n + 255) & ~(size_t)255;
      return r;
    };
    for (int k = 0; k < 4; ++k) {
      QueueView<R>& v = q[k];
      v.base = take(per);
      v.cap = want;
      v.count = (unsigned long long*)ctl + k;
    }
    return DPM_OK;
  }
};

// device bytes for the dense beamlet rows of one group when the caller sets
// no budget (dpm_run_options.tally_budget)
constexpr size_t kBeamletBudget = (size_t)16 << 30;

// most transport rounds enqueued per host synchronisation (and per captured
// graph); rounds past the drain find empty queues and exit at once
constexpr int kRoundsPerSync = 8;

// round preparation as a kernel (not a memset: it is recorded into graphs):
// clear the next round's queue counts and both consumption heads
__global__ void k_round_prep(int* ctl, int nxt) {
  if (threadIdx.x == 0) {
    unsigned long long* cq = (unsigned long long*)ctl;
    cq[nxt] = 0;
    cq[2 + nxt] = 0;
    ctl[kCtlHeadG] = 0;
    ctl[kCtlHeadE] = 0;
  }
}

// histories per queue fill when the caller does not choose
constexpr int64_t kDefaultChunk = (int64_t)1 << 24;
// smallest chunk the autotuner times (below it the fixed cost per chunk --
// drain rounds, fold, kernel tails -- dominates)
constexpr int64_t kMinTunedChunk = (int64_t)1 << 21;

// A run's chunks go to kPipes pipes, each with its own stream, queues and
// control words, that run side by side: a pipe's kernel tails, drain rounds
// and folds overlap the other pipe's transport (a persistent kernel's
// blocks leave one by one as its lanes run dry, and the other stream's
// kernel takes their place).
constexpr int kPipes = 2;
constexpr int64_t kPipeChunk = (int64_t)1 << 23;

// histories per chunk and batch rows per chunk (per pipe)
struct ChunkPlan {
  int64_t chunk;
  int rows;
};
uint64_t batch_start(uint64_t b, uint64_t n, uint64_t nb);
// A caller's chunk is kept (at most kDefaultChunk), and so is the default
// kDefaultChunk on one pipe. On two pipes the run is cut into an even
// number of chunks of about kPipeChunk histories at most, of whole batches
// when there are batches, so that both pipes have work.
ChunkPlan plan_chunks(uint64_t nhist, int nbatch, int64_t opt_chunk, int npipes) {
  ChunkPlan c{1, 1};
  if (nhist == 0) return c;
  if (opt_chunk > 0 || npipes < 2) {
    if (opt_chunk <= 0) opt_chunk = kDefaultChunk;
    c.chunk = std::min<int64_t>(std::min<int64_t>(opt_chunk, kDefaultChunk), (int64_t)nhist);
    if (nbatch > 0) {
      const uint64_t per = nhist / nbatch + 1;
      c.rows = (int)std::max<uint64_t>(1, std::min<uint64_t>((uint64_t)nbatch, (uint64_t)c.chunk / per));
    }
    return c;
  }
  uint64_t nch = std::max<uint64_t>(npipes, (nhist + kPipeChunk - 1) / kPipeChunk);
  nch = (nch + npipes - 1) / npipes * npipes;
  if (nbatch > 0) {
    const uint64_t bpc = std::max<uint64_t>(1, ((uint64_t)nbatch + nch - 1) / nch);
    c.rows = (int)bpc;
    uint64_t span = 0;   // the longest run of bpc consecutive batches
    for (uint64_t b = 0; b < (uint64_t)nbatch; b += bpc)
      span = std::max(span, batch_start(std::min<uint64_t>(b + bpc, nbatch), nhist, nbatch) -
                                batch_start(b, nhist, nbatch));
    c.chunk = (int64_t)std::min<uint64_t>(span, (uint64_t)kDefaultChunk);
  } else {
    c.chunk = (int64_t)((nhist + nch - 1) / nch);
  }
  c.chunk = std::max<int64_t>(1, c.chunk);
  return c;
}

// grow-only device buffer kept by a tables handle across runs
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  bool ensure(size_t want) {
    if (want <= bytes && p) return true;
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
    if (hipMalloc(&p, want) != hipSuccess) { p = nullptr; return false; }
    bytes = want;
    return true;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  double* d() const { return (double*)p; }
};

// tally fold: per-batch rows -> sum, sum of squares; rows cleared
__global__ void k_fold(double* rows, int nrows, int64_t nvox, double* sum, double* sum2) {
  for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < nvox; v += (int64_t)gridDim.x * blockDim.x) {
    double s = 0, s2 = 0;
    for (int r = 0; r < nrows; ++r) {
      const double x = rows[(size_t)r * nvox + v];
      if (x != 0.0) {
        s += x;
        s2 += x * x;
        rows[(size_t)r * nvox + v] = 0.0;
      }
    }
    if (s2 != 0.0) {
      sum[v] += s;
      sum2[v] += s2;
    }
  }
}

// FP64 atomic-add throughput probe: nops device-scope additions to cells
// spread pseudo-randomly over a buffer of ncells doubles (< 2^32), the access
// pattern of the voxel tally, without return values as the transport issues them
__global__ void k_atomic_probe(double* buf, uint64_t ncells, uint64_t nops, uint64_t seed) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < nops; i += stride) {
    const uint64_t h = mix64(i ^ seed);
    unsafeAtomicAdd(&buf[((h >> 32) * ncells) >> 32], 1.0);
  }
}

// sum of a tally into *out (block partial sums, one atomic per block)
__global__ void k_sum(const double* x, int64_t n, double* out) {
  __shared__ double part[4];
  double acc = 0.0;
  for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < n; v += (int64_t)gridDim.x * blockDim.x)
    acc += x[v];
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double b = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) b += part[w];
    unsafeAtomicAdd(out, b);
  }
}

// clears n doubles (a kernel, not a memset, so that a caller may record the
// no-allocation run into a graph)
__global__ void k_zero(double* x, int64_t n) {
  for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < n; v += (int64_t)gridDim.x * blockDim.x)
    x[v] = 0.0;
}

// Beamlet rows -> CSR in three passes over (row, segment) blocks, 1-D grid
// block = row * kCsrSeg + segment:
//   k_bl_stats  row maxima (of the non-negative part) and the sum of the rows
//   k_bl_count  entries above threshold per segment
//   k_bl_fill   entries written in voxel order at their segment's offset
constexpr int kCsrSeg = 64;

__device__ __forceinline__ void bl_segment(int64_t nvox, int64_t* v0, int64_t* v1) {
  const int64_t seg = blockIdx.x % kCsrSeg;
  const int64_t len = (nvox + kCsrSeg - 1) / kCsrSeg;
  *v0 = seg * len;
  *v1 = *v0 + len < nvox ? *v0 + len : nvox;
}

__global__ void k_bl_stats(const double* rows, int64_t nvox, unsigned long long* rmax_bits, double* sum) {
  __shared__ double pa[4], pm[4];
  const int row = blockIdx.x / kCsrSeg;
  int64_t v0, v1;
  bl_segment(nvox, &v0, &v1);
  const double* r = rows + (size_t)row * nvox;
  double acc = 0.0, m = 0.0;
  for (int64_t v = v0 + threadIdx.x; v < v1; v += blockDim.x) {
    const double x = r[v];
    acc += x;
    m = fmax(m, x);
  }
  for (int o = 32; o > 0; o >>= 1) {
    acc += __shfl_xor(acc, o);
    m = fmax(m, __shfl_xor(m, o));
  }
  if ((threadIdx.x & 63) == 0) { pa[threadIdx.x >> 6] = acc; pm[threadIdx.x >> 6] = m; }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = 0.0, mm = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) { a += pa[w]; mm = fmax(mm, pm[w]); }
    if (a != 0.0) unsafeAtomicAdd(sum, a);
    // non-negative doubles order like their bit patterns
    if (mm > 0.0) atomicMax(&rmax_bits[row], (unsigned long long)__double_as_longlong(mm));
  }
}

__global__ void k_bl_count(const double* rows, int64_t nvox, const unsigned long long* rmax_bits, double rel,
                           int64_t* segcnt) {
  __shared__ int pc[4];
  const int row = blockIdx.x / kCsrSeg;
  int64_t v0, v1;
  bl_segment(nvox, &v0, &v1);
  const double* r = rows + (size_t)row * nvox;
  const double thr = rel * __longlong_as_double((long long)rmax_bits[row]);
  int c = 0;
  for (int64_t v = v0 + threadIdx.x; v < v1; v += blockDim.x) {
    const double x = r[v];
    c += (x > thr && x > 0.0) ? 1 : 0;
  }
  for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
  if ((threadIdx.x & 63) == 0) pc[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    int64_t tsum = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) tsum += pc[w];
    segcnt[blockIdx.x] = tsum;
  }
}

__global__ void k_bl_fill(const double* rows, int64_t nvox, const unsigned long long* rmax_bits, double rel,
                          const int64_t* segoff, int32_t* col, double* val) {
  __shared__ int64_t base;
  __shared__ int wsum[4];
  const int row = blockIdx.x / kCsrSeg;
  int64_t v0, v1;
  bl_segment(nvox, &v0, &v1);
  const double* r = rows + (size_t)row * nvox;
  const double thr = rel * __longlong_as_double((long long)rmax_bits[row]);
  if (threadIdx.x == 0) base = segoff[blockIdx.x];
  __syncthreads();
  for (int64_t b0 = v0; b0 < v1; b0 += blockDim.x) {
    const int64_t v = b0 + threadIdx.x;
    const double x = v < v1 ? r[v] : 0.0;
    const bool keep = v < v1 && x > thr && x > 0.0;
    const uint64_t b = __ballot(keep);
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) wsum[wave] = __popcll(b);
    __syncthreads();
    int before = 0;
    for (int w = 0; w < wave; ++w) before += wsum[w];
    const int rank = before + (int)__popcll(b & ((1ull << (threadIdx.x & 63)) - 1ull));
    if (keep) {
      col[base + rank] = (int32_t)v;
      val[base + rank] = x;
    }
    __syncthreads();
    if (threadIdx.x == 0) base += wsum[0] + wsum[1] + wsum[2] + wsum[3];
    __syncthreads();
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// handles
// ---------------------------------------------------------------------------
struct dpm_tables {
  int device = 0;
  unsigned char* scratch_dev = nullptr;   // this device's copy of `scratch` (dpm_simulate_noalloc)
  hipStream_t stream = nullptr;
  int num_cu = 0;
  int nmat = 0;
  double ecut = 0, gcut = 0;
  std::vector<double> rho;
  Tables<float> tf;
  Tables<double> td;
  void* dev_f = nullptr;
  void* dev_d = nullptr;
  hipStream_t pipe_stream = nullptr;   // the second pipe's stream (the first pipe uses `stream`)
  Workspace<float> wsf[kPipes];
  Workspace<double> wsd[kPipes];
  int* ctl = nullptr;        // kPipes x kCtlInts: [0..7] packed queue counts, heads, overflow
  int* ctl_host = nullptr;   // pinned copy of ctl, read after each block of rounds
  StatsDev* stats = nullptr;
  hipEvent_t ev[4] = {};     // run start, run end, pipe join, fold order
  // per pipe: [4k..4k+3] kernel bounds of round k of a sync block, then
  // block done, fold start, fold end
  hipEvent_t pev[kPipes][4 * kRoundsPerSync + 3] = {};
  // tally buffers reused across runs
  DevBuf buf_rows, buf_steps, buf_sum, buf_sum2, buf_he;
  // leading bytes of buf_rows known to be zero (left so by the folds of a
  // completed batch run), valid while buf_rows still holds rows_zero_ptr
  size_t rows_zero = 0;
  const void* rows_zero_ptr = nullptr;
  DevBuf buf_src;            // source description of the current run
  DevBuf buf_blaux;          // beamlet row maxima, segment counts and offsets
  // result of the last dpm_simulate, kept on the device for dpm_tally_fetch
  const double* last_tally = nullptr;
  const double* last_second = nullptr;
  int64_t last_n = 0;
  // CSR of the last beamlet run
  int32_t* csr_col = nullptr;
  double* csr_val = nullptr;
  int64_t csr_nnz = 0;
};

struct dpm_phantom {
  const dpm_tables* owner = nullptr;
  int geom = GEOM_SLAB;
  double lbox = 1.0;
  int geom_index = 0;
  int nx = 0, ny = 0, nz = 0;
  uint8_t* d_mat = nullptr;
  float* d_rho = nullptr;
  double maj_scale = 1.0;
  double maj_min = 1.0;     // smallest valid scale: max(rho_voxel / rho_nominal), >= 1
  int nbins = 0;
  int64_t nvox = 0;
};

namespace {

template <typename R>
void fill_geometry(const dpm_tables* t, const dpm_phantom* p, Geometry<R>& G) {
  std::memset(&G, 0, sizeof(G));
  G.L = (R)p->lbox;
  G.invL = (R)(1. / p->lbox);
  G.hL = (R)(0.5 * p->lbox);
  G.geomIndex = p->geom_index;
  G.nx = p->nx; G.ny = p->ny; G.nz = p->nz;
  G.org[0] = (R)(-0.5 * p->nx * p->lbox);
  G.org[1] = (R)(-0.5 * p->ny * p->lbox);
  G.org[2] = (R)(-0.5 * p->nz * p->lbox);
  G.vmat = p->d_mat;
  G.vrho = p->d_rho;
  for (int im = 0; im < t->nmat; ++im) G.rho[im] = (R)t->rho[im];
  G.guard = &t->stats->guard_push;
}

// device-time of a pair of events, 0 when unavailable
float elapsed(hipEvent_t a, hipEvent_t b) {
  float ms = 0.f;
  if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return 0.f;
  return ms;
}

template <typename R>
struct Launch;
template <>
struct Launch<float> {
  static hipError_t photons(const KParams<float>& P, const SourceParams* S, const PrimRange& r, int g, int b,
                            const QueueView<float>& in, int* h, const QueueView<float>& qe,
                            const QueueView<float>& qg, hipStream_t st) {
    return launch_photons_f32(P, S, r, g, b, in, h, qe, qg, st);
  }
  static hipError_t electrons(const KParams<float>& P, const SourceParams* S, const PrimRange& r, int g, int b,
                              const QueueView<float>& in, int* h, const QueueView<float>& qe,
                              const QueueView<float>& qg, hipStream_t st) {
    return launch_electrons_f32(P, S, r, g, b, in, h, qe, qg, st);
  }
};
template <>
struct Launch<double> {
  static hipError_t photons(const KParams<double>& P, const SourceParams* S, const PrimRange& r, int g, int b,
                            const QueueView<double>& in, int* h, const QueueView<double>& qe,
                            const QueueView<double>& qg, hipStream_t st) {
    return launch_photons_f64(P, S, r, g, b, in, h, qe, qg, st);
  }
  static hipError_t electrons(const KParams<double>& P, const SourceParams* S, const PrimRange& r, int g, int b,
                              const QueueView<double>& in, int* h, const QueueView<double>& qe,
                              const QueueView<double>& qg, hipStream_t st) {
    return launch_electrons_f64(P, S, r, g, b, in, h, qe, qg, st);
  }
};

// Source of one call as the host describes it: entries (plain source: one;
// beamlets: one per beamlet) split by particle kind, in call order
struct HostEntry {
  uint64_t first;   // global id of the entry's first history
  uint64_t n;       // histories
  double x, y, e;
  int particle;
  int row;          // tally row, or -1 (batch row / row 0)
  double angle_deg = 0.0;   // gantry fields
  double z = 0.0, u = 0.0, v = 0.0, w = 1.0;   // phase-space records: entry z and direction
};
struct HostSource {
  int kind = DPM_SRC_PENCIL;
  double emin = 0, emax = 0, fx = 0, fy = 0;
  int spectrum = DPM_SPEC_FLAT;
  double emode = 0;
  int phsp_project = 0;             // phase-space records outside a voxel phantom move to its surface
  double sad = 0;                   // gantry-field collimation (dpm_source.sad, jaws, jaw_*)
  int jaws = 0;
  double jx1 = 0, jx2 = 0, jy1 = 0, jy2 = 0, jtau = 0;
  std::vector<double> spec_edges, spec_cdf;   // DPM_SPEC_TABLE
  std::vector<HostEntry> list[2];   // [0] photons, [1] e-/e+
  void add(const HostEntry& h) { list[h.particle == DPM_PHOTON ? 0 : 1].push_back(h); }
  uint64_t count(int k) const {
    uint64_t c = 0;
    for (const HostEntry& h : list[k]) c += h.n;
    return c;
  }
};

// Uploads the source into one device block: SourceParams followed by the
// per-list arrays (cum, first, x, y, e, particle, row); returns the device
// address of the SourceParams.
// cos and sin of a gantry angle in degrees, exact at multiples of 90
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

// Fixed device memory a run uses instead of allocating (the no-allocation
// entry point): the source description goes to [src, src + src_bytes).
struct FixedMem {
  unsigned char* src = nullptr;
  size_t src_bytes = 0;
};

int upload_source(dpm_tables* t, const HostSource& hs, double z0, int nbatch, uint64_t hist_begin, uint64_t nhist,
                  SourceParams** out, const FixedMem* fm = nullptr) {
  size_t bytes = (sizeof(SourceParams) + 255) & ~(size_t)255;
  const size_t nsp = hs.spec_edges.size();   // table spectrum: edges and cumulative weights
  const size_t off_edges = bytes;
  bytes += (8 * nsp + 255) & ~(size_t)255;
  const size_t off_cdf = bytes;
  bytes += (8 * nsp + 255) & ~(size_t)255;
  constexpr int kArrays = 13;
  size_t off[2][kArrays];
  for (int k = 0; k < 2; ++k) {
    const size_t n = hs.list[k].size();
    const size_t sz[kArrays] = {8 * (n + 1), 8 * n, 8 * n, 8 * n, 8 * n, 4 * n, 4 * n, 8 * n, 8 * n,
                                8 * n, 8 * n, 8 * n, 8 * n};
    for (int a = 0; a < kArrays; ++a) {
      off[k][a] = bytes;
      bytes += (sz[a] + 255) & ~(size_t)255;
    }
  }
  if (fm) {
    if (bytes > fm->src_bytes)
      return set_error(DPM_E_INVALID, "source description of " + std::to_string(bytes) + " bytes exceeds the " +
                                          std::to_string(fm->src_bytes) + " of the fixed workspace");
  } else if (!t->buf_src.ensure(bytes)) {
    return set_error(DPM_E_NOMEM, "cannot allocate the source description");
  }
  unsigned char* dst = fm ? fm->src : (unsigned char*)t->buf_src.p;
  std::vector<unsigned char> h(bytes, 0);
  char* base = (char*)dst;
  SourceParams S;
  std::memset(&S, 0, sizeof(S));
  S.kind = hs.kind;
  S.phsp_project = hs.phsp_project;
  S.nbatch = nbatch;
  S.emin = hs.emin;
  S.emax = hs.emax;
  S.spectrum = hs.spectrum;
  S.emode = hs.emode;
  if (nsp > 1) {
    S.spec_n = (int)nsp - 1;
    std::memcpy(h.data() + off_edges, hs.spec_edges.data(), 8 * nsp);
    std::memcpy(h.data() + off_cdf, hs.spec_cdf.data(), 8 * nsp);
    S.spec_edges = (const double*)(base + off_edges);
    S.spec_cdf = (const double*)(base + off_cdf);
  }
  S.fx = hs.fx;
  S.fy = hs.fy;
  S.z0 = z0;
  S.sad = hs.sad;
  S.jaws = hs.jaws;
  S.jx1 = hs.jx1; S.jx2 = hs.jx2; S.jy1 = hs.jy1; S.jy2 = hs.jy2; S.jtau = hs.jtau;
  S.hist_begin = hist_begin;
  S.nhist_total = nhist;
  for (int k = 0; k < 2; ++k) {
    S.count[k] = hs.count(k);
    // several entries (beamlets): interleave their histories with a
    // multiplicative bijection of the index (the prime must not divide the count)
    S.perm[k] = 0;
    bool rows = false;
    for (const HostEntry& h : hs.list[k]) rows = rows || h.row >= 0;
    if (rows && hs.list[k].size() > 1 && S.count[k] > 1)
      for (uint64_t prime : {2654435761ull, 4294967291ull, 2246822519ull})
        if (S.count[k] % prime != 0) { S.perm[k] = prime; break; }
  }
  for (int k = 0; k < 2; ++k) {
    const std::vector<HostEntry>& L = hs.list[k];
    const int n = (int)L.size();
    uint64_t* cum = (uint64_t*)(h.data() + off[k][0]);
    uint64_t* first = (uint64_t*)(h.data() + off[k][1]);
    double* x = (double*)(h.data() + off[k][2]);
    double* y = (double*)(h.data() + off[k][3]);
    double* e = (double*)(h.data() + off[k][4]);
    int* pt = (int*)(h.data() + off[k][5]);
    int* row = (int*)(h.data() + off[k][6]);
    double* ca = (double*)(h.data() + off[k][7]);
    double* sa = (double*)(h.data() + off[k][8]);
    double* pz = (double*)(h.data() + off[k][9]);
    double* du = (double*)(h.data() + off[k][10]);
    double* dv = (double*)(h.data() + off[k][11]);
    double* dw = (double*)(h.data() + off[k][12]);
    cum[0] = 0;
    for (int i = 0; i < n; ++i) {
      cum[i + 1] = cum[i] + L[i].n;
      first[i] = L[i].first;
      x[i] = L[i].x; y[i] = L[i].y; e[i] = L[i].e;
      pt[i] = L[i].particle;
      row[i] = L[i].row;
      gantry_cos_sin(L[i].angle_deg, &ca[i], &sa[i]);
      pz[i] = L[i].z; du[i] = L[i].u; dv[i] = L[i].v; dw[i] = L[i].w;
    }
    PrimList& D = S.list[k];
    D.n = n;
    D.cum = (const uint64_t*)(base + off[k][0]);
    D.first = (const uint64_t*)(base + off[k][1]);
    D.x = (const double*)(base + off[k][2]);
    D.y = (const double*)(base + off[k][3]);
    D.e = (const double*)(base + off[k][4]);
    D.particle = (const int*)(base + off[k][5]);
    D.row = (const int*)(base + off[k][6]);
    D.ca = (const double*)(base + off[k][7]);
    D.sa = (const double*)(base + off[k][8]);
    D.z = (const double*)(base + off[k][9]);
    D.u = (const double*)(base + off[k][10]);
    D.v = (const double*)(base + off[k][11]);
    D.w = (const double*)(base + off[k][12]);
  }
  std::memcpy(h.data(), &S, sizeof(S));
  // ordered on the run stream, completed before the staging vector goes out
  // of scope (the previous run on the stream may still read a fixed destination)
  HIP_TRY(hipMemcpyAsync(dst, h.data(), bytes, hipMemcpyHostToDevice, t->stream));
  HIP_TRY(hipStreamSynchronize(t->stream));
  *out = (SourceParams*)dst;
  return DPM_OK;
}

// boundaries of batch b: first history ceil(b*N/nb)
uint64_t batch_start(uint64_t b, uint64_t n, uint64_t nb) {
  const unsigned __int128 num = (unsigned __int128)b * n;
  return (uint64_t)((num + nb - 1) / nb);
}

// Runs every history of the source and leaves the tally on the device in
// `rows` (rows x nbins|nvox doubles). Histories are taken in chunks over the
// concatenation [photon list | e-/e+ list]; the first round of a chunk
// generates its primaries inside the transport kernels.
// pipes of a run (dpm_run_options.pipes, checked by the entry points)
int pipe_count(const dpm_run_options* opt) { return opt->pipes == 2 ? 2 : 1; }

template <typename R>
int run_transport(dpm_tables* t, const dpm_phantom* p, const HostSource& hs, const dpm_run_options* opt,
                  Workspace<R>* ws, Tables<R>& T, const ChunkPlan& plan, double* d_rows, double* d_steps,
                  double* d_sum, double* d_sum2, double* d_hist_edep, dpm_stats* st, const FixedMem* fm = nullptr) {
  const hipStream_t s = t->stream;
  KParams<R> P;
  std::memset(&P, 0, sizeof(P));
  P.T = T;
  fill_geometry(t, p, P.G);
  P.k0 = (uint32_t)opt->seed;
  P.k1 = (uint32_t)(opt->seed >> 32);
  P.majScale = (R)p->maj_scale;
  P.tally = d_rows;
  P.steps = d_steps;
  P.nvox = p->nvox;
  P.nbins = p->nbins;
  P.hist_edep = d_hist_edep;
  P.hist_base = opt->hist_begin;
  P.stats = t->stats;
  P.refill_min = std::min(64, std::max(1, opt->refill_min));
  P.tally_mode = (opt->tally_mode == 1 || opt->tally_mode == 2) ? opt->tally_mode : 0;
  // FP32: the electron kernel's spline tables go to LDS when they fit
  P.lds_tables = (sizeof(R) == 4 && opt->table_memory == DPM_TABLES_AUTO &&
                  3 * T.nmat * T.itr1.n + T.moller.n + T.k1.n <= kLdsSplineNodes &&
                  T.itr1.n == T.dedx.n && T.itr1.n == T.brem.n) ? 1 : 0;
  {
    const int need = (2 + 3 * T.nmat) * T.ptot.n;
    const bool ok = sizeof(R) == 4 && opt->table_memory == DPM_TABLES_AUTO && T.lin_shared_grid;
    P.lds_photon = !ok ? 0 : (need <= kLdsLinearFloats ? 1 : (need <= kLdsLinearFloatsLarge ? 2 : 0));
  }

  const uint64_t n_ph = hs.count(0), n_el = hs.count(1);
  const uint64_t nhist = n_ph + n_el;
  // batch rows follow history ids, so the ids must run through the entries
  // in the order the chunks take them ([photon list | e-/e+ list], each in
  // list order, consecutive from hist_begin), and no entry has its own row
  bool ordered = true;
  {
    uint64_t next = opt->hist_begin;
    for (int k = 0; k < 2; ++k)
      for (const HostEntry& h : hs.list[k]) {
        ordered = ordered && h.row < 0 && h.first == next;
        next += h.n;
      }
  }
  const int nbatch = (ordered && p->geom == GEOM_VOXEL && opt->nbatch > 1) ? opt->nbatch : 0;
  const int64_t chunk = std::max<int64_t>(1, std::min<int64_t>(plan.chunk, kDefaultChunk));
  const int nrows_per_chunk = std::max(1, plan.rows);
  const int npipes = pipe_count(opt);
  const int wpc = opt->waves_per_cu > 0 ? opt->waves_per_cu : 16;
  const int blocks = std::max(1, t->num_cu * std::max(1, wpc / 4));
  SourceParams* dS = nullptr;
  int rc = upload_source(t, hs, p->geom == GEOM_SLAB ? -0.5 * p->lbox : -0.5 * p->nz * p->lbox, nbatch,
                         opt->hist_begin, nhist, &dS, fm);
  if (rc != DPM_OK) return rc;

  // the chunks in history order: whole batches per chunk (at least one, as
  // many as fit in `chunk` histories and the pipe's batch rows)
  struct Chunk {
    PrimRange r_ph, r_el;
    uint32_t bfirst;
    int nrows;
  };
  std::vector<Chunk> chunks;
  {
    uint64_t done = 0, next_batch = 0;
    while (done < nhist) {
      uint64_t n = std::min<uint64_t>((uint64_t)chunk, nhist - done);
      uint32_t bfirst = 0;
      int nrows = 1;
      if (nbatch > 0) {
        const uint64_t b0 = next_batch;
        uint64_t b1 = b0 + 1;
        while (b1 < (uint64_t)nbatch && (int)(b1 - b0) < nrows_per_chunk &&
               batch_start(b1 + 1, nhist, nbatch) - batch_start(b0, nhist, nbatch) <= (uint64_t)chunk)
          ++b1;
        n = batch_start(b1, nhist, nbatch) - batch_start(b0, nhist, nbatch);
        bfirst = (uint32_t)b0;
        nrows = (int)(b1 - b0);
        next_batch = b1;
      }
      // this chunk's primaries in each kind's index space
      const uint64_t a = done, b = done + n;
      chunks.push_back(Chunk{PrimRange{std::min(a, n_ph), std::min(b, n_ph), bfirst},
                             PrimRange{std::max(a, n_ph) - n_ph, std::max(b, n_ph) - n_ph, bfirst}, bfirst, nrows});
      done += n;
    }
  }
  // queue capacity: secondaries of the largest chunk (a chunk of whole
  // batches can exceed `chunk` when one batch does), plus the unfilled tails
  // of the per-wave append blocks (at most one block per wave and run of an
  // output queue -- the electron queues have two runs -- and kernel launch)
  int64_t largest = 1;
  for (const Chunk& c : chunks)
    largest = std::max<int64_t>(largest, (int64_t)(c.r_ph.end - c.r_ph.beg) + (int64_t)(c.r_el.end - c.r_el.beg));
  const int64_t holes = (int64_t)3 * blocks * (256 / 64) * 64;
  const int cap = (int)std::min<int64_t>((int64_t)16 * largest + holes + 65536, (int64_t)1 << 30);
  for (int q = 0; q < npipes; ++q) {
    const int rc = ws[q].ensure(cap, t->ctl + q * kCtlInts);
    if (rc != DPM_OK) return rc;
  }

  int rounds = 0;
  float ms_ph = 0, ms_el = 0, ms_fold = 0;
  int c_ph = 0, c_el = 0, c_fold = 0;
  const bool use_graph = opt->graph != 0;
  if (opt->rounds_per_sync != 0 && (opt->rounds_per_sync < 2 || opt->rounds_per_sync > kRoundsPerSync ||
                                    (opt->rounds_per_sync & 1)))
    return set_error(DPM_E_INVALID, "rounds_per_sync must be even, 2.." + std::to_string(kRoundsPerSync));
  const int rps = opt->rounds_per_sync ? opt->rounds_per_sync : kRoundsPerSync;
  int n_graphs = 0;
  constexpr int kDone = 4 * kRoundsPerSync, kFold0 = kDone + 1, kFold1 = kDone + 2;

  // one pipe: its stream, queues, control words and the chunk it runs
  struct PipeRun {
    hipStream_t s = nullptr;
    Workspace<R>* ws = nullptr;
    int* ctl = nullptr;
    const int* host = nullptr;
    hipEvent_t* ev = nullptr;
    KParams<R> P;
    hipGraphExec_t g_first = nullptr, g_rest = nullptr;
    size_t chunk = 0;
    bool first = false, fold_timed = false;
  };
  PipeRun pr[kPipes];
  const int64_t nb = p->geom == GEOM_SLAB ? p->nbins : p->nvox;
  for (int q = 0; q < kPipes; ++q) {
    PipeRun& u = pr[q];
    u.s = q == 0 ? s : t->pipe_stream;
    u.ws = &ws[q];
    u.ctl = t->ctl + q * kCtlInts;
    u.host = t->ctl_host + q * kCtlInts;
    u.ev = t->pev[q];
    u.P = P;
    u.P.overflow = u.ctl + kCtlOverflow;
    // batch rows: each pipe its own (folded after each of its chunks);
    // otherwise both pipes add into the same rows
    if (nbatch > 0) u.P.tally = d_rows + (size_t)q * nrows_per_chunk * nb;
  }
  // on every exit: let both pipes finish, then free their graphs
  struct Guard {
    PipeRun* pr;
    ~Guard() {
      for (int q = 0; q < kPipes; ++q) (void)hipStreamSynchronize(pr[q].s);
      for (int q = 0; q < kPipes; ++q) {
        if (pr[q].g_first) (void)hipGraphExecDestroy(pr[q].g_first);
        if (pr[q].g_rest) (void)hipGraphExecDestroy(pr[q].g_rest);
      }
    }
  } guard{pr};

  HIP_TRY(hipEventRecord(t->ev[0], s));
  // the second pipe starts after what the caller enqueued on the run stream
  HIP_TRY(hipEventRecord(t->ev[2], s));
  HIP_TRY(hipStreamWaitEvent(t->pipe_stream, t->ev[2], 0));

  // rps rounds of photons then electrons; the first generates the chunk's
  // primaries. Round r reads queues `cur` and appends photons to `nxt`;
  // electrons of both kernels go to the electron queue the electron kernel
  // of the round (photons) or the next round (electrons) reads.
  auto enqueue = [&](PipeRun& u, bool with_primaries, bool timed) -> hipError_t {
    const Chunk& ch = chunks[u.chunk];
    const PrimRange none{0, 0, ch.bfirst};
    QueueView<R>* q = u.ws->q;
    int c = 0;
    for (int k = 0; k < rps; ++k) {
      const int nx = c ^ 1;
      hipLaunchKernelGGL(k_round_prep, dim3(1), dim3(64), 0, u.s, u.ctl, nx);
      if (timed) (void)hipEventRecord(u.ev[4 * k], u.s);
      u.P.kind = 0;
      hipError_t e = Launch<R>::photons(u.P, dS, with_primaries && k == 0 ? ch.r_ph : none, p->geom, blocks, q[0 + c],
                                        u.ctl + kCtlHeadG, q[2 + c], q[0 + nx], u.s);
      if (e != hipSuccess) return e;
      if (timed) { (void)hipEventRecord(u.ev[4 * k + 1], u.s); (void)hipEventRecord(u.ev[4 * k + 2], u.s); }
      u.P.kind = 1;
      e = Launch<R>::electrons(u.P, dS, with_primaries && k == 0 ? ch.r_el : none, p->geom, blocks, q[2 + c],
                               u.ctl + kCtlHeadE, q[2 + nx], q[0 + nx], u.s);
      if (e != hipSuccess) return e;
      if (timed) (void)hipEventRecord(u.ev[4 * k + 3], u.s);
      c = nx;
    }
    return hipSuccess;
  };
  std::deque<int> inflight;   // pipes with a block of rounds enqueued, oldest first
  // one block of rps rounds on pipe q, then the copy of its control words
  auto launch = [&](int q) -> int {
    PipeRun& u = pr[q];
    if (use_graph) {
      // the first block of a chunk generates its primaries (a graph per
      // chunk); later blocks replay one graph per pipe
      hipGraphExec_t& ge = u.first ? u.g_first : u.g_rest;
      if (u.first && u.g_first) { (void)hipGraphExecDestroy(u.g_first); u.g_first = nullptr; }
      if (!ge) {
        hipGraph_t g = nullptr;
        HIP_TRY(hipStreamBeginCapture(u.s, hipStreamCaptureModeThreadLocal));
        const hipError_t e = enqueue(u, u.first, false);
        const hipError_t e2 = hipStreamEndCapture(u.s, &g);
        if (e != hipSuccess || e2 != hipSuccess) {
          if (g) (void)hipGraphDestroy(g);
          return set_error(DPM_E_DEVICE, "graph capture of the transport rounds failed");
        }
        const hipError_t e3 = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e3 != hipSuccess) { ge = nullptr; return set_error(DPM_E_DEVICE, "graph instantiation failed"); }
        ++n_graphs;
      }
      HIP_TRY(hipGraphLaunch(ge, u.s));
    } else {
      HIP_TRY(enqueue(u, u.first, true));
    }
    HIP_TRY(hipMemcpyAsync((void*)u.host, u.ctl, kCtlInts * sizeof(int), hipMemcpyDeviceToHost, u.s));
    HIP_TRY(hipEventRecord(u.ev[kDone], u.s));
    inflight.push_back(q);
    return DPM_OK;
  };
  size_t next_chunk = 0;
  auto start = [&](int q) -> int {
    PipeRun& u = pr[q];
    u.chunk = next_chunk++;
    u.first = true;
    HIP_TRY(hipMemsetAsync(u.ctl, 0, kCtlInts * sizeof(int), u.s));
    return launch(q);
  };
  for (int q = 0; q < npipes && next_chunk < chunks.size(); ++q)
    if ((rc = start(q)) != DPM_OK) return rc;
  bool fold_enqueued = false;
  while (!inflight.empty()) {
    const int q = inflight.front();
    inflight.pop_front();
    PipeRun& u = pr[q];
    HIP_TRY(hipEventSynchronize(u.ev[kDone]));
    u.first = false;
    if (!use_graph)
      for (int k = 0; k < rps; ++k) {
        ms_ph += elapsed(u.ev[4 * k], u.ev[4 * k + 1]);
        ms_el += elapsed(u.ev[4 * k + 2], u.ev[4 * k + 3]);
        ++c_ph; ++c_el;
      }
    if (u.fold_timed) {   // the pipe's previous fold preceded this block on its stream
      ms_fold += elapsed(u.ev[kFold0], u.ev[kFold1]);
      u.fold_timed = false;
    }
    rounds += rps;
    if (u.host[kCtlOverflow] != 0)
      return set_error(DPM_E_OVERFLOW, "particle queue overflow (capacity " + std::to_string(cap) +
                                           "); use a smaller chunk");
    if (rounds > 100000) return set_error(DPM_E_DEVICE, "transport did not drain");
    // an even number of rounds leaves the next inputs in queues 0 and 2
    // (packed counters: ints 0-1 and 4-5)
    if ((u.host[0] | u.host[1] | u.host[4] | u.host[5]) != 0) {
      if ((rc = launch(q)) != DPM_OK) return rc;
      continue;
    }
    if (nbatch > 0) {
      // folds add into the shared sum and sum of squares: one at a time
      if (fold_enqueued) HIP_TRY(hipStreamWaitEvent(u.s, t->ev[3], 0));
      HIP_TRY(hipEventRecord(u.ev[kFold0], u.s));
      hipLaunchKernelGGL(k_fold, dim3(std::max(1, t->num_cu * 8)), dim3(256), 0, u.s, u.P.tally,
                         chunks[u.chunk].nrows, p->nvox, d_sum, d_sum2);
      HIP_TRY(hipGetLastError());
      HIP_TRY(hipEventRecord(u.ev[kFold1], u.s));
      HIP_TRY(hipEventRecord(t->ev[3], u.s));
      fold_enqueued = true;
      u.fold_timed = true;
      ++c_fold;
    }
    if (next_chunk < chunks.size() && (rc = start(q)) != DPM_OK) return rc;
  }
  // the run stream waits for the second pipe
  HIP_TRY(hipEventRecord(t->ev[2], t->pipe_stream));
  HIP_TRY(hipStreamWaitEvent(s, t->ev[2], 0));
  HIP_TRY(hipEventRecord(t->ev[1], s));
  HIP_TRY(hipEventSynchronize(t->ev[1]));
  for (int q = 0; q < kPipes; ++q)
    if (pr[q].fold_timed) ms_fold += elapsed(pr[q].ev[kFold0], pr[q].ev[kFold1]);
  if (st) {
    st->seconds = elapsed(t->ev[0], t->ev[1]) * 1e-3;
    st->rounds = rounds;
    // slot 0 (source) stays 0: primaries are generated inside the transport kernels
    st->kernel_ms[0] = 0; st->kernel_ms[1] = ms_ph; st->kernel_ms[2] = ms_el; st->kernel_ms[3] = ms_fold;
    st->kernel_calls[0] = 0; st->kernel_calls[1] = c_ph; st->kernel_calls[2] = c_el; st->kernel_calls[3] = c_fold;
    st->graphs = n_graphs;
    st->lds_tables = P.lds_tables | ((P.lds_photon != 0) << 1);
  }
  return DPM_OK;
}

int finish_stats(dpm_tables* t, dpm_stats* st) {
  StatsDev h;
  HIP_TRY(hipMemcpyAsync(&h, t->stats, sizeof(h), hipMemcpyDeviceToHost, t->stream));
  HIP_TRY(hipStreamSynchronize(t->stream));
  if (st) {
    for (int i = 0; i < DPM_NCNT; ++i) {
      st->counters_photon[i] = h.cnt[0][i];
      st->counters_electron[i] = h.cnt[1][i];
      st->counters[i] = h.cnt[0][i] + h.cnt[1][i];
    }
    st->n_pair = h.cnt[0][kCntPair] + h.cnt[1][kCntPair];
    st->n_annih = h.cnt[0][kCntAnnih] + h.cnt[1][kCntAnnih];
    st->lane_slots[0] = h.cnt[0][kCntLaneSlots];
    st->lane_slots[1] = h.cnt[1][kCntLaneSlots];
    st->e_escaped = h.escaped;
    st->e_primary = h.e_primary;
    st->e_scored = h.scored;
    st->hist_size = std::max(20, h.maxiz + 10);
    st->guard_exits = h.cnt[0][kCntGuard] + h.cnt[1][kCntGuard] + h.guard_push;
  }
  return DPM_OK;
}

// the plain source: one entry of nhist histories from id hist_begin
HostSource source_from(const dpm_source* src, uint64_t nhist, uint64_t hist_begin) {
  HostSource hs;
  hs.kind = src->kind;
  hs.emin = src->emin;
  hs.emax = src->emax;
  hs.fx = src->field_x;
  hs.fy = src->field_y;
  hs.spectrum = src->spectrum;
  hs.emode = src->emode;
  if (src->kind == DPM_SRC_GANTRY) {
    hs.sad = src->sad;
    hs.jaws = src->jaws ? 1 : 0;
    if (hs.jaws) {
      hs.jx1 = src->jaw_x1; hs.jx2 = src->jaw_x2; hs.jy1 = src->jaw_y1; hs.jy2 = src->jaw_y2;
      hs.jtau = src->jaw_tau;
    }
  }
  if (src->spectrum == DPM_SPEC_TABLE) {
    // cumulative weights, normalised by their total; the last is 1 exactly
    const int n = src->spec_n;
    hs.spec_edges.assign(src->spec_edges, src->spec_edges + n + 1);
    hs.spec_cdf.assign(n + 1, 0.0);
    for (int i = 0; i < n; ++i) hs.spec_cdf[i + 1] = hs.spec_cdf[i] + src->spec_weights[i];
    const double total = hs.spec_cdf[n];
    for (int i = 1; i < n; ++i) hs.spec_cdf[i] /= total;
    hs.spec_cdf[n] = 1.0;
  }
  HostEntry h{hist_begin, nhist, src->x0, src->y0, src->energy, src->particle, -1};
  h.angle_deg = src->angle_deg;
  hs.add(h);
  return hs;
}

int check_source(const dpm_tables* t, int particle, double e, double emin, double emax);

// the source description of dpm_simulate / dpm_simulate_fields
int validate_source(const dpm_tables* t, const dpm_phantom* p, const dpm_source* src) {
  if (src->kind != DPM_SRC_PENCIL && src->kind != DPM_SRC_FIELD && src->kind != DPM_SRC_GANTRY)
    return set_error(DPM_E_INVALID, "unknown source kind");
  if (src->kind == DPM_SRC_GANTRY && p->geom != GEOM_VOXEL)
    return set_error(DPM_E_INVALID, "gantry fields need a voxel phantom");
  if (!std::isfinite(src->angle_deg)) return set_error(DPM_E_INVALID, "gantry angle is not finite");
  if (src->kind != DPM_SRC_GANTRY && (src->sad != 0.0 || src->jaws != 0))
    return set_error(DPM_E_INVALID, "sad and jaws collimate gantry fields (DPM_SRC_GANTRY) only");
  if (src->kind == DPM_SRC_GANTRY) {
    if (!std::isfinite(src->sad) || src->sad < 0.0) return set_error(DPM_E_INVALID, "sad must be finite and >= 0");
    // the point source must lie outside the phantom: beyond its circumscribed sphere
    const double hx = 0.5 * p->nx * p->lbox, hy = 0.5 * p->ny * p->lbox, hz = 0.5 * p->nz * p->lbox;
    if (src->sad > 0.0 && !(src->sad > std::sqrt(hx * hx + hy * hy + hz * hz)))
      return set_error(DPM_E_INVALID, "sad must put the source outside the phantom (beyond half its diagonal)");
    if (src->jaws) {
      const double j[5] = {src->jaw_x1, src->jaw_x2, src->jaw_y1, src->jaw_y2, src->jaw_tau};
      for (double v : j)
        if (!std::isfinite(v)) return set_error(DPM_E_INVALID, "jaw positions and thickness must be finite");
      if (!(src->jaw_x1 < src->jaw_x2) || !(src->jaw_y1 < src->jaw_y2))
        return set_error(DPM_E_INVALID, "jaw apertures must have x1 < x2 and y1 < y2");
      if (src->jaw_tau < 0.0) return set_error(DPM_E_INVALID, "jaw optical thickness must be >= 0");
    }
  }
  if (src->spectrum != DPM_SPEC_FLAT && src->spectrum != DPM_SPEC_TRIANGULAR && src->spectrum != DPM_SPEC_TABLE)
    return set_error(DPM_E_INVALID, "unknown spectrum");
  if (src->spectrum == DPM_SPEC_TABLE) {
    if (src->spec_n < 1 || src->spec_n > (1 << 20) || !src->spec_edges || !src->spec_weights)
      return set_error(DPM_E_INVALID, "spectrum table needs 1 .. 2^20 bins with edges and weights");
    double total = 0.0;
    for (int i = 0; i < src->spec_n; ++i) {
      const double w = src->spec_weights[i];
      if (!std::isfinite(w) || w < 0.0) return set_error(DPM_E_INVALID, "spectrum weights must be finite and >= 0");
      if (!(src->spec_edges[i + 1] > src->spec_edges[i]) || !std::isfinite(src->spec_edges[i + 1]))
        return set_error(DPM_E_INVALID, "spectrum bin edges must increase");
      total += w;
    }
    if (!(total > 0.0) || !std::isfinite(total)) return set_error(DPM_E_INVALID, "spectrum weights sum to zero");
    return check_source(t, src->particle, src->energy, src->spec_edges[0], src->spec_edges[src->spec_n]);
  }
  if (src->spectrum == DPM_SPEC_TRIANGULAR && !(src->emax > src->emin && src->emode >= src->emin &&
                                                src->emode <= src->emax))
    return set_error(DPM_E_INVALID, "triangular spectrum needs emin <= emode <= emax, emin < emax");
  return check_source(t, src->particle, src->energy, src->emin, src->emax);
}

int check_source(const dpm_tables* t, int particle, double e, double emin, double emax) {
  if (particle != DPM_ELECTRON && particle != DPM_PHOTON && particle != DPM_POSITRON)
    return set_error(DPM_E_INVALID, "unknown primary particle");
  const double lo = emax > emin ? emin : e;
  const double hi = emax > emin ? emax : e;
  const double cut = particle == DPM_PHOTON ? t->gcut : t->ecut;
  if (!(lo > cut) || !(hi < 21.0) || !std::isfinite(hi))
    return set_error(DPM_E_INVALID, "primary energy must lie above the tracking cut and below the table maximum");
  return DPM_OK;
}

}  // namespace

extern "C" {

const char* dpm_last_error(void) { return g_last_error.c_str(); }
int dpm_abi_version(void) { return DPM_ABI_VERSION; }

int dpm_host_alloc(size_t bytes, void** out) {
  if (!out || bytes == 0) return set_error(DPM_E_INVALID, "null argument or zero size");
  *out = nullptr;
  if (hipHostMalloc(out, bytes, hipHostMallocDefault) != hipSuccess || !*out)
    return set_error(DPM_E_NOMEM, "cannot allocate " + std::to_string(bytes) + " bytes of pinned host memory");
  return DPM_OK;
}

int dpm_host_free(void* p) {
  if (p && hipHostFree(p) != hipSuccess) return set_error(DPM_E_DEVICE, "hipHostFree failed");
  return DPM_OK;
}

int dpm_probe_atomics(dpm_tables* t, int64_t ncells, int64_t nops, double* seconds, double* check) {
  if (!t || !seconds) return set_error(DPM_E_INVALID, "null argument");
  if (ncells < 1 || ncells >= ((int64_t)1 << 32) || nops < 1 || nops > ((int64_t)1 << 40))
    return set_error(DPM_E_INVALID, "probe needs 1 <= ncells < 2^32 and 1 <= nops <= 2^40");
  HIP_TRY(hipSetDevice(t->device));
  double* buf = nullptr;
  if (hipMalloc(&buf, sizeof(double) * (size_t)ncells) != hipSuccess)
    return set_error(DPM_E_NOMEM, "cannot allocate the probe buffer");
  struct Free {
    double* p;
    ~Free() { (void)hipFree(p); }
  } guard{buf};
  HIP_TRY(hipMemsetAsync(buf, 0, sizeof(double) * (size_t)ncells, t->stream));
  const int blocks = std::max(1, t->num_cu * 16);
  // one untimed pass brings the buffer's pages and the code in
  hipLaunchKernelGGL(k_atomic_probe, dim3(blocks), dim3(256), 0, t->stream, buf, (uint64_t)ncells,
                     (uint64_t)std::min<int64_t>(nops, (int64_t)blocks * 256), (uint64_t)0x51);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipEventRecord(t->ev[0], t->stream));
  hipLaunchKernelGGL(k_atomic_probe, dim3(blocks), dim3(256), 0, t->stream, buf, (uint64_t)ncells, (uint64_t)nops,
                     (uint64_t)0x1234);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipEventRecord(t->ev[1], t->stream));
  HIP_TRY(hipEventSynchronize(t->ev[1]));
  *seconds = elapsed(t->ev[0], t->ev[1]) * 1e-3;
  if (check) {
    // total added: warm-up pass + timed pass (every addition is 1.0)
    HIP_TRY(hipMemsetAsync(t->stats, 0, sizeof(StatsDev), t->stream));
    hipLaunchKernelGGL(k_sum, dim3(std::max(1, t->num_cu * 4)), dim3(256), 0, t->stream, buf, ncells,
                       &t->stats->scored);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(check, &t->stats->scored, sizeof(double), hipMemcpyDeviceToHost, t->stream));
    HIP_TRY(hipStreamSynchronize(t->stream));
  }
  return DPM_OK;
}

int dpm_tables_create(const dpm_table_desc* desc, int device, dpm_tables** out) {
  if (!out) return set_error(DPM_E_INVALID, "null output handle");
  *out = nullptr;
  int rc = validate(desc);
  if (rc != DPM_OK) return rc;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return set_error(DPM_E_DEVICE, "no HIP device available");
  if (device < 0 || device >= ndev) return set_error(DPM_E_INVALID, "device index out of range");
  HIP_TRY(hipSetDevice(device));
  dpm_tables* t = new dpm_tables();
  t->device = device;
  t->nmat = desc->nmat;
  t->ecut = desc->electron_cut;
  t->gcut = desc->gamma_cut;
  t->rho.assign(desc->density, desc->density + desc->nmat);
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) { delete t; return set_error(DPM_E_DEVICE, "device query failed"); }
  t->num_cu = prop.multiProcessorCount;
  if ((rc = build_pack<float>(desc, t->tf, &t->dev_f)) != DPM_OK || (rc = build_pack<double>(desc, t->td, &t->dev_d)) != DPM_OK) {
    dpm_tables_destroy(t);
    return rc;
  }
  // both streams at the default priority
  if (hipStreamCreateWithFlags(&t->stream, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&t->pipe_stream, hipStreamNonBlocking) != hipSuccess ||
      hipMalloc(&t->ctl, kPipes * kCtlInts * sizeof(int)) != hipSuccess ||
      hipHostMalloc(&t->ctl_host, kPipes * kCtlInts * sizeof(int), hipHostMallocDefault) != hipSuccess ||
      hipMalloc(&t->stats, sizeof(StatsDev)) != hipSuccess) {
    dpm_tables_destroy(t);
    return set_error(DPM_E_DEVICE, "cannot create the run streams or control block");
  }
  for (auto& e : t->ev)
    if (hipEventCreate(&e) != hipSuccess) { dpm_tables_destroy(t); return set_error(DPM_E_DEVICE, "event creation failed"); }
  // the device address of this GPU's scratch copy (the device is current)
  {
    void* sp = nullptr;
    const hipError_t e = hipGetSymbolAddress(&sp, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !sp) {
      dpm_tables_destroy(t);
      return set_error(DPM_E_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    }
    t->scratch_dev = (unsigned char*)sp;
  }
  for (auto& pe : t->pev)
    for (auto& e : pe)
      if (hipEventCreate(&e) != hipSuccess) { dpm_tables_destroy(t); return set_error(DPM_E_DEVICE, "event creation failed"); }
  *out = t;
  return DPM_OK;
}

int dpm_tables_load(const char* dir, int device, dpm_tables** out) {
  if (!dir || !out) return set_error(DPM_E_INVALID, "null argument");
  SimMaterialData mat;
  SimElectronData el;
  SimPhotonData ph;
  std::string err;
  if (!mat.Load(dir, err) || !el.Load(dir, err) || !ph.Load(dir, err)) return set_error(DPM_E_IO, err);
  if ((int)el.fITr1MFP.size() != mat.fNumMaterial || (int)el.fDEDX.size() != mat.fNumMaterial ||
      (int)el.fIMFPBrem.size() != mat.fNumMaterial || (int)ph.fIMFPTotal.size() != mat.fNumMaterial ||
      el.fSBTables.nmat != mat.fNumMaterial)
    return set_error(DPM_E_IO, "table files disagree on the number of materials");
  SimTableView view;
  FlattenTables(mat, el, ph, view);
  return dpm_tables_create(&view.desc, device, out);
}

void dpm_tables_destroy(dpm_tables* t) {
  if (!t) return;
  (void)hipSetDevice(t->device);
  if (t->stream) (void)hipStreamSynchronize(t->stream);
  if (t->pipe_stream) (void)hipStreamSynchronize(t->pipe_stream);
  for (int q = 0; q < kPipes; ++q) {
    t->wsf[q].release();
    t->wsd[q].release();
  }
  for (DevBuf* b : {&t->buf_rows, &t->buf_steps, &t->buf_sum, &t->buf_sum2, &t->buf_he, &t->buf_src, &t->buf_blaux}) b->release();
  if (t->csr_col) (void)hipFree(t->csr_col);
  if (t->csr_val) (void)hipFree(t->csr_val);
  if (t->dev_f) (void)hipFree(t->dev_f);
  if (t->dev_d) (void)hipFree(t->dev_d);
  if (t->ctl) (void)hipFree(t->ctl);
  if (t->ctl_host) (void)hipHostFree(t->ctl_host);
  if (t->stats) (void)hipFree(t->stats);
  for (auto& e : t->ev)
    if (e) (void)hipEventDestroy(e);
  for (auto& pe : t->pev)
    for (auto& e : pe)
      if (e) (void)hipEventDestroy(e);
  if (t->stream) (void)hipStreamDestroy(t->stream);
  if (t->pipe_stream) (void)hipStreamDestroy(t->pipe_stream);
  delete t;
}

int dpm_tables_num_materials(const dpm_tables* t) { return t ? t->nmat : 0; }

int dpm_phantom_create_slab(const dpm_tables* t, double lbox, int geom_index, dpm_phantom** out) {
  if (!t || !out) return set_error(DPM_E_INVALID, "null argument");
  if (!(lbox >= 0.05) || !(lbox <= 200.0)) return set_error(DPM_E_INVALID, "voxel size must be in [0.05, 200] mm");
  if (geom_index < 0) return set_error(DPM_E_INVALID, "negative geometry index");
  const int need = geom_index == 1 || geom_index == 2 ? 2 : (geom_index == 3 ? 3 : 1);
  if (t->nmat < need) return set_error(DPM_E_INVALID, "the slab geometry needs more materials than the tables hold");
  dpm_phantom* p = new dpm_phantom();
  p->owner = t;
  p->geom = GEOM_SLAB;
  p->lbox = lbox;
  p->geom_index = geom_index;
  p->nbins = (int)std::floor(100.0 / lbox + 0.5) + 2;
  p->nvox = p->nbins;
  *out = p;
  return DPM_OK;
}

int dpm_phantom_create_voxel(const dpm_tables* t, int nx, int ny, int nz, double lbox, const uint8_t* mat,
                             const float* rho, dpm_phantom** out) {
  if (!t || !out || !mat || !rho) return set_error(DPM_E_INVALID, "null argument");
  if (nx < 1 || ny < 1 || nz < 1 || (int64_t)nx * ny * nz > ((int64_t)1 << 31))
    return set_error(DPM_E_INVALID, "phantom dimensions out of range");
  if (!(lbox > 0.0)) return set_error(DPM_E_INVALID, "voxel size must be positive");
  const int64_t nv = (int64_t)nx * ny * nz;
  double scale = 1.0;
  for (int64_t k = 0; k < nv; ++k) {
    if (mat[k] >= t->nmat) return set_error(DPM_E_INVALID, "voxel material index out of range");
    if (!(rho[k] > 0.0f) || !std::isfinite(rho[k])) return set_error(DPM_E_INVALID, "voxel density must be positive");
    scale = std::max(scale, (double)rho[k] / t->rho[mat[k]]);
  }
  HIP_TRY(hipSetDevice(t->device));
  dpm_phantom* p = new dpm_phantom();
  p->owner = t;
  p->geom = GEOM_VOXEL;
  p->lbox = lbox;
  p->nx = nx; p->ny = ny; p->nz = nz;
  p->nvox = nv;
  p->maj_scale = scale;
  p->maj_min = scale;
  if (hipMalloc(&p->d_mat, nv) != hipSuccess || hipMalloc(&p->d_rho, nv * sizeof(float)) != hipSuccess ||
      hipMemcpy(p->d_mat, mat, nv, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(p->d_rho, rho, nv * sizeof(float), hipMemcpyHostToDevice) != hipSuccess) {
    dpm_phantom_destroy(p);
    return set_error(DPM_E_NOMEM, "cannot upload the phantom");
  }
  *out = p;
  return DPM_OK;
}

int dpm_phantom_set_majorant(dpm_phantom* p, double scale) {
  if (!p) return set_error(DPM_E_INVALID, "null argument");
  if (!std::isfinite(scale) || scale < p->maj_min)
    return set_error(DPM_E_INVALID, "majorant scale below max(rho / rho_nominal) = " + std::to_string(p->maj_min));
  p->maj_scale = scale;
  return DPM_OK;
}

double dpm_phantom_get_majorant(const dpm_phantom* p) { return p ? p->maj_scale : 0.0; }

void dpm_phantom_destroy(dpm_phantom* p) {
  if (!p) return;
  if (p->d_mat) (void)hipFree(p->d_mat);
  if (p->d_rho) (void)hipFree(p->d_rho);
  delete p;
}

int64_t dpm_phantom_num_bins(const dpm_phantom* p) { return p ? (p->geom == GEOM_SLAB ? p->nbins : p->nvox) : 0; }

static int simulate_source(dpm_tables* t, const dpm_phantom* p, const HostSource& hs, uint64_t nhist,
                           const dpm_run_options& opt, double* tally, double* steps, dpm_stats* st);

int dpm_simulate(dpm_tables* t, const dpm_phantom* p, const dpm_source* src, uint64_t nhist,
                 const dpm_run_options* opt_in, double* tally, double* steps, dpm_stats* st) {
  if (!t || !p || !src) return set_error(DPM_E_INVALID, "null argument");
  if (p->owner != t) return set_error(DPM_E_INVALID, "phantom belongs to another tables handle");
  if (nhist == 0) return set_error(DPM_E_INVALID, "zero histories");
  if (nhist > 0xFFFFFFFFull) return set_error(DPM_E_INVALID, "at most 2^32-1 histories per call");
  dpm_run_options opt;
  std::memset(&opt, 0, sizeof(opt));
  if (opt_in) opt = *opt_in;
  if (opt.hist_begin + nhist > 0x100000000ull) return set_error(DPM_E_INVALID, "history ids exceed 32 bits");
  if (opt.precision != DPM_FP32 && opt.precision != DPM_FP64) return set_error(DPM_E_INVALID, "unknown precision");
  if (opt.pipes < 0 || opt.pipes > kPipes) return set_error(DPM_E_INVALID, "pipes must be 0, 1 or 2");
  if (opt.nbatch > 1 && (uint64_t)opt.nbatch > nhist) return set_error(DPM_E_INVALID, "more batches than histories");
  const int rc = validate_source(t, p, src);
  if (rc != DPM_OK) return rc;
  return simulate_source(t, p, source_from(src, nhist, opt.hist_begin), nhist, opt, tally, steps, st);
}

int dpm_simulate_fields(dpm_tables* t, const dpm_phantom* p, const dpm_source* src, const dpm_field* fields, int nf,
                        const dpm_run_options* opt_in, double* tally, double* steps, dpm_stats* st) {
  if (!t || !p || !src || !fields || nf <= 0) return set_error(DPM_E_INVALID, "null argument");
  if (p->owner != t) return set_error(DPM_E_INVALID, "phantom belongs to another tables handle");
  if (nf > (1 << 20)) return set_error(DPM_E_INVALID, "too many fields");
  dpm_run_options opt;
  std::memset(&opt, 0, sizeof(opt));
  if (opt_in) opt = *opt_in;
  if (opt.precision != DPM_FP32 && opt.precision != DPM_FP64) return set_error(DPM_E_INVALID, "unknown precision");
  if (opt.pipes < 0 || opt.pipes > kPipes) return set_error(DPM_E_INVALID, "pipes must be 0, 1 or 2");
  int rc = validate_source(t, p, src);
  if (rc != DPM_OK) return rc;
  HostSource hs = source_from(src, 1, 0);
  hs.list[0].clear();
  hs.list[1].clear();
  uint64_t total = 0;
  for (int i = 0; i < nf; ++i) {
    if (fields[i].nhist == 0) return set_error(DPM_E_INVALID, "field without histories");
    if (!std::isfinite(fields[i].angle_deg)) return set_error(DPM_E_INVALID, "gantry angle is not finite");
    HostEntry h{opt.hist_begin + total, fields[i].nhist, fields[i].x0, fields[i].y0, src->energy, src->particle, -1};
    h.angle_deg = fields[i].angle_deg;
    hs.add(h);
    total += fields[i].nhist;
  }
  if (opt.hist_begin + total > 0x100000000ull) return set_error(DPM_E_INVALID, "history ids exceed 32 bits");
  if (opt.nbatch > 1 && (uint64_t)opt.nbatch > total) return set_error(DPM_E_INVALID, "more batches than histories");
  return simulate_source(t, p, hs, total, opt, tally, steps, st);
}

int dpm_simulate_phsp(dpm_tables* t, const dpm_phantom* p, const dpm_phsp_particle* rec, uint64_t n,
                      const dpm_run_options* opt_in, double* tally, double* steps, dpm_stats* st) {
  if (!t || !p || !rec) return set_error(DPM_E_INVALID, "null argument");
  if (p->owner != t) return set_error(DPM_E_INVALID, "phantom belongs to another tables handle");
  if (n == 0) return set_error(DPM_E_INVALID, "no phase-space records");
  if (n > ((uint64_t)1 << 30)) return set_error(DPM_E_INVALID, "at most 2^30 phase-space records per call");
  dpm_run_options opt;
  std::memset(&opt, 0, sizeof(opt));
  if (opt_in) opt = *opt_in;
  if (opt.hist_begin + n > 0x100000000ull) return set_error(DPM_E_INVALID, "history ids exceed 32 bits");
  if (opt.precision != DPM_FP32 && opt.precision != DPM_FP64) return set_error(DPM_E_INVALID, "unknown precision");
  if (opt.pipes < 0 || opt.pipes > kPipes) return set_error(DPM_E_INVALID, "pipes must be 0, 1 or 2");
  if (opt.nbatch > 1 && (uint64_t)opt.nbatch > n) return set_error(DPM_E_INVALID, "more batches than histories");
  HostSource hs;
  hs.kind = 3;
  hs.phsp_project = p->geom == GEOM_VOXEL;
  uint64_t nph = 0;
  for (uint64_t i = 0; i < n; ++i) {
    const dpm_phsp_particle& r = rec[i];
    int rc = check_source(t, r.particle, r.energy, 0, 0);
    if (rc != DPM_OK) return set_error(rc, "phase-space record " + std::to_string(i) + ": " + g_last_error);
    const double nn = r.u * r.u + r.v * r.v + r.w * r.w;
    if (!std::isfinite(r.x) || !std::isfinite(r.y) || !std::isfinite(r.z) || !(std::fabs(nn - 1.0) < 1e-6))
      return set_error(DPM_E_INVALID, "phase-space record " + std::to_string(i) +
                                          ": position not finite or direction not a unit vector");
    nph += r.particle == DPM_PHOTON;
  }
  // history ids: the photon records first, then the e-/e+ records, each in
  // the caller's order (the order the chunks run them)
  hs.list[0].reserve(nph);
  hs.list[1].reserve(n - nph);
  uint64_t iph = 0, iel = 0;
  for (uint64_t i = 0; i < n; ++i) {
    const dpm_phsp_particle& r = rec[i];
    const uint64_t id = opt.hist_begin + (r.particle == DPM_PHOTON ? iph++ : nph + iel++);
    HostEntry h{id, 1, r.x, r.y, r.energy, r.particle, -1};
    h.z = r.z; h.u = r.u; h.v = r.v; h.w = r.w;
    hs.add(h);
  }
  return simulate_source(t, p, hs, n, opt, tally, steps, st);
}

static int simulate_source(dpm_tables* t, const dpm_phantom* p, const HostSource& hs, uint64_t nhist,
                           const dpm_run_options& opt, double* tally, double* steps, dpm_stats* st) {
  int rc = DPM_OK;
  HIP_TRY(hipSetDevice(t->device));
  const hipStream_t s = t->stream;
  const bool slab = p->geom == GEOM_SLAB;
  const int nbatch = (!slab && opt.nbatch > 1) ? opt.nbatch : 0;
  const int64_t nb = slab ? p->nbins : p->nvox;
  // tally rows: one, or the batches of one chunk for each pipe
  const ChunkPlan plan = plan_chunks(nhist, nbatch, opt.chunk, pipe_count(&opt));
  const int rows = nbatch > 0 ? plan.rows * pipe_count(&opt) : 1;
  if (!t->buf_rows.ensure(sizeof(double) * nb * rows) || (slab && !t->buf_steps.ensure(sizeof(double) * nb)) ||
      (nbatch > 0 && (!t->buf_sum.ensure(sizeof(double) * nb) || !t->buf_sum2.ensure(sizeof(double) * nb))) ||
      (opt.hist_edep && !t->buf_he.ensure(sizeof(double) * nhist)))
    return set_error(DPM_E_NOMEM, "cannot allocate the tally");
  double* d_rows = t->buf_rows.d();
  double* d_steps = slab ? t->buf_steps.d() : nullptr;
  double* d_sum = nbatch > 0 ? t->buf_sum.d() : nullptr;
  double* d_sum2 = nbatch > 0 ? t->buf_sum2.d() : nullptr;
  double* d_he = opt.hist_edep ? t->buf_he.d() : nullptr;
  // a batch run's folds clear every row they use, so a following batch run
  // finds its rows zero and skips clearing them again
  const size_t rows_bytes = sizeof(double) * nb * rows;
  const bool rows_clean = nbatch > 0 && t->rows_zero_ptr == t->buf_rows.p && t->rows_zero >= rows_bytes;
  t->rows_zero = 0;
  if (!rows_clean) (void)hipMemsetAsync(d_rows, 0, rows_bytes, s);
  if (d_steps) (void)hipMemsetAsync(d_steps, 0, sizeof(double) * nb, s);
  if (d_sum) (void)hipMemsetAsync(d_sum, 0, sizeof(double) * nb, s);
  if (d_sum2) (void)hipMemsetAsync(d_sum2, 0, sizeof(double) * nb, s);
  if (d_he) (void)hipMemsetAsync(d_he, 0, sizeof(double) * nhist, s);
  (void)hipMemsetAsync(t->stats, 0, sizeof(StatsDev), s);
  dpm_stats local;
  std::memset(&local, 0, sizeof(local));
  rc = opt.precision == DPM_FP64
           ? run_transport<double>(t, p, hs, &opt, t->wsd, t->td, plan, d_rows, d_steps, d_sum, d_sum2, d_he, &local)
           : run_transport<float>(t, p, hs, &opt, t->wsf, t->tf, plan, d_rows, d_steps, d_sum, d_sum2, d_he, &local);
  if (rc == DPM_OK) {
    // scored energy: reduced on the device from the tally that is returned
    hipLaunchKernelGGL(k_sum, dim3(std::max(1, t->num_cu * 4)), dim3(256), 0, s, nbatch > 0 ? d_sum : d_rows, nb,
                       &t->stats->scored);
    if (hipGetLastError() != hipSuccess) rc = set_error(DPM_E_DEVICE, "tally sum launch failed");
  }
  if (rc == DPM_OK) rc = finish_stats(t, &local);
  if (rc == DPM_OK && nbatch > 0) {
    t->rows_zero = rows_bytes;
    t->rows_zero_ptr = t->buf_rows.p;
  }
  // the result stays on the device until the next run (dpm_tally_fetch)
  t->last_tally = rc == DPM_OK ? (nbatch > 0 ? d_sum : d_rows) : nullptr;
  t->last_second = rc == DPM_OK ? (slab ? d_steps : d_sum2) : nullptr;
  t->last_n = rc == DPM_OK ? nb : 0;
  if (rc == DPM_OK && tally) rc = dpm_tally_fetch(t, tally, steps);
  if (rc == DPM_OK && opt.hist_edep &&
      hipMemcpy(opt.hist_edep, d_he, sizeof(double) * nhist, hipMemcpyDeviceToHost) != hipSuccess)
    rc = set_error(DPM_E_DEVICE, "per-history copy failed");
  if (st) *st = local;
  return rc;
}

int dpm_tally_fetch(dpm_tables* t, double* tally, double* second) {
  if (!t || !tally) return set_error(DPM_E_INVALID, "null argument");
  if (!t->last_tally) return set_error(DPM_E_INVALID, "no tally on the device (run dpm_simulate first)");
  HIP_TRY(hipSetDevice(t->device));
  const size_t bytes = sizeof(double) * (size_t)t->last_n;
  HIP_TRY(hipMemcpyAsync(tally, t->last_tally, bytes, hipMemcpyDeviceToHost, t->stream));
  if (second) {
    if (t->last_second) HIP_TRY(hipMemcpyAsync(second, t->last_second, bytes, hipMemcpyDeviceToHost, t->stream));
    else std::memset(second, 0, bytes);
  }
  HIP_TRY(hipStreamSynchronize(t->stream));
  return DPM_OK;
}

// Histories per chunk that the scratch queues hold for a grid of `blocks`
// blocks: the run's queue capacity is 16 records per history of a chunk plus
// the unfilled tails of the per-wave append blocks plus slack (run_transport).
static int64_t noalloc_chunk(int blocks, bool fp64) {
  const size_t qbytes = kScratchBytes - kScratchSrcBytes;
  const int cap = fp64 ? Workspace<double>::capacity_of(qbytes) : Workspace<float>::capacity_of(qbytes);
  const int64_t holes = (int64_t)3 * blocks * (256 / 64) * 64;
  return ((int64_t)cap - holes - 65536) / 16;
}

int dpm_simulate_noalloc(dpm_tables* t, const dpm_phantom* p, const dpm_source* src, uint64_t nhist,
                         const dpm_run_options* opt_in, double* d_tally, double* d_steps, dpm_stats* st) {
  if (!t || !p || !src || !d_tally) return set_error(DPM_E_INVALID, "null argument");
  if (p->owner != t) return set_error(DPM_E_INVALID, "phantom belongs to another tables handle");
  if (!t->scratch_dev) return set_error(DPM_E_DEVICE, "no device scratch on this tables handle");
  if (nhist == 0) return set_error(DPM_E_INVALID, "zero histories");
  if (nhist > 0xFFFFFFFFull) return set_error(DPM_E_INVALID, "at most 2^32-1 histories per call");
  dpm_run_options opt;
  std::memset(&opt, 0, sizeof(opt));
  if (opt_in) opt = *opt_in;
  if (opt.hist_begin + nhist > 0x100000000ull) return set_error(DPM_E_INVALID, "history ids exceed 32 bits");
  if (opt.precision != DPM_FP32 && opt.precision != DPM_FP64) return set_error(DPM_E_INVALID, "unknown precision");
  if (opt.nbatch > 1) return set_error(DPM_E_INVALID, "no batch statistics without allocation (nbatch must be 0 or 1)");
  if (opt.hist_edep) return set_error(DPM_E_INVALID, "no per-history tally without allocation");
  if (opt.pipes > 1) return set_error(DPM_E_INVALID, "one pipe without allocation");
  if (opt.graph) return set_error(DPM_E_INVALID, "graph replay instantiates graphs: not without allocation");
  const int rc0 = validate_source(t, p, src);
  if (rc0 != DPM_OK) return rc0;
  opt.nbatch = 0;
  opt.pipes = 1;
  // the chunk that the scratch queues hold at this grid (default: one block
  // of 4 waves per CU, whose append-block tails leave room for the most histories)
  const int wpc = opt.waves_per_cu > 0 ? opt.waves_per_cu : 4;
  opt.waves_per_cu = wpc;
  const int blocks = std::max(1, t->num_cu * std::max(1, wpc / 4));
  const int64_t cmax = noalloc_chunk(blocks, opt.precision == DPM_FP64);
  if (cmax < 64) return set_error(DPM_E_INVALID, "waves_per_cu too large for the device scratch");
  if (opt.chunk > cmax)
    return set_error(DPM_E_INVALID, "chunk of " + std::to_string(opt.chunk) + " histories exceeds the " +
                                        std::to_string(cmax) + " the device scratch holds at this grid");
  const ChunkPlan plan{opt.chunk > 0 ? opt.chunk : cmax, 1};
  HIP_TRY(hipSetDevice(t->device));
  const hipStream_t s = t->stream;
  const bool slab = p->geom == GEOM_SLAB;
  const int64_t nb = slab ? p->nbins : p->nvox;
  if (slab && !d_steps) return set_error(DPM_E_INVALID, "a slab run needs the steps array");
  static_assert(sizeof(StatsDev) % sizeof(double) == 0, "the statistics block is cleared as doubles");
  // the caller's outputs start at zero (they accumulate atomics)
  const unsigned zb = (unsigned)std::max<int64_t>(1, std::min<int64_t>((nb + 255) / 256, (int64_t)t->num_cu * 8));
  hipLaunchKernelGGL(k_zero, dim3(zb), dim3(256), 0, s, d_tally, nb);
  if (slab && d_steps) hipLaunchKernelGGL(k_zero, dim3(zb), dim3(256), 0, s, d_steps, nb);
  hipLaunchKernelGGL(k_zero, dim3(1), dim3(256), 0, s, (double*)t->stats, (int64_t)(sizeof(StatsDev) / sizeof(double)));
  HIP_TRY(hipGetLastError());
  // source description and queues in this GPU's scratch
  FixedMem fm;
  fm.src = t->scratch_dev;
  fm.src_bytes = kScratchSrcBytes;
  unsigned char* qbase = t->scratch_dev + kScratchSrcBytes;
  const size_t qbytes = kScratchBytes - kScratchSrcBytes;
  dpm_stats local;
  std::memset(&local, 0, sizeof(local));
  int rc;
  const HostSource hs = source_from(src, nhist, opt.hist_begin);
  if (opt.precision == DPM_FP64) {
    Workspace<double> ws[kPipes];
    ws[0].view(qbase, qbytes);
    rc = run_transport<double>(t, p, hs, &opt, ws, t->td, plan, d_tally, slab ? d_steps : nullptr, nullptr, nullptr,
                               nullptr, &local, &fm);
  } else {
    Workspace<float> ws[kPipes];
    ws[0].view(qbase, qbytes);
    rc = run_transport<float>(t, p, hs, &opt, ws, t->tf, plan, d_tally, slab ? d_steps : nullptr, nullptr, nullptr,
                              nullptr, &local, &fm);
  }
  if (rc == DPM_OK) {
    hipLaunchKernelGGL(k_sum, dim3(std::max(1, t->num_cu * 4)), dim3(256), 0, s, d_tally, nb, &t->stats->scored);
    if (hipGetLastError() != hipSuccess) rc = set_error(DPM_E_DEVICE, "tally sum launch failed");
  }
  if (rc == DPM_OK) rc = finish_stats(t, &local);
  // the result is the caller's: nothing stays behind for dpm_tally_fetch
  t->last_tally = t->last_second = nullptr;
  t->last_n = 0;
  if (st) *st = local;
  return rc;
}

int dpm_simulate_beamlets(dpm_tables* t, const dpm_phantom* p, const dpm_beamlet* b, int nb,
                          const dpm_run_options* opt_in, double rel_threshold, int64_t* row_ptr, dpm_stats* st) {
  if (!t || !p || !b || nb <= 0 || !row_ptr) return set_error(DPM_E_INVALID, "null argument");
  if (p->owner != t) return set_error(DPM_E_INVALID, "phantom belongs to another tables handle");
  if (p->geom != GEOM_VOXEL) return set_error(DPM_E_INVALID, "beamlets need a voxel phantom");
  if (nb > (1 << 20)) return set_error(DPM_E_INVALID, "too many beamlets");
  dpm_run_options opt;
  std::memset(&opt, 0, sizeof(opt));
  if (opt_in) opt = *opt_in;
  if (opt.precision != DPM_FP32 && opt.precision != DPM_FP64) return set_error(DPM_E_INVALID, "unknown precision");
  if (opt.pipes < 0 || opt.pipes > kPipes) return set_error(DPM_E_INVALID, "pipes must be 0, 1 or 2");
  if (opt.tally_budget < 0) return set_error(DPM_E_INVALID, "negative tally budget");
  // each beamlet's histories keep the ids they have when the beamlets run one
  // after the other from hist_begin, whatever group it runs in
  std::vector<uint64_t> first(nb + 1, 0);
  for (int i = 0; i < nb; ++i) {
    int rc = check_source(t, b[i].particle, b[i].energy, 0, 0);
    if (rc != DPM_OK) return rc;
    if (b[i].nhist == 0) return set_error(DPM_E_INVALID, "beamlet without histories");
    first[i + 1] = first[i] + b[i].nhist;
  }
  if (opt.hist_begin + first[nb] > 0x100000000ull) return set_error(DPM_E_INVALID, "history ids exceed 32 bits");
  // dense FP64 rows of a group of beamlets within the budget
  const size_t row_bytes = sizeof(double) * (size_t)p->nvox;
  const size_t budget = opt.tally_budget > 0 ? (size_t)opt.tally_budget : kBeamletBudget;
  if (row_bytes > budget)
    return set_error(DPM_E_INVALID, "one beamlet's dense row (" + std::to_string(row_bytes >> 20) +
                                        " MiB) exceeds the tally budget");
  const int group = (int)std::min<size_t>((size_t)nb, budget / row_bytes);
  HIP_TRY(hipSetDevice(t->device));
  const hipStream_t s = t->stream;
  if (t->csr_col) (void)hipFree(t->csr_col);
  if (t->csr_val) (void)hipFree(t->csr_val);
  t->csr_col = nullptr;
  t->csr_val = nullptr;
  t->csr_nnz = 0;
  t->last_tally = t->last_second = nullptr;
  t->last_n = 0;
  const size_t nseg = (size_t)group * kCsrSeg;
  t->rows_zero = 0;   // the beamlet rows are cleared by this run, not left zero
  if (!t->buf_rows.ensure(row_bytes * group) ||
      !t->buf_blaux.ensure(sizeof(double) * nb + 2 * sizeof(int64_t) * nseg + 1024))
    return set_error(DPM_E_NOMEM, "cannot allocate the beamlet tally (" + std::to_string((row_bytes * group) >> 20) +
                                      " MiB)");
  double* d_rows = t->buf_rows.d();
  unsigned char* q = (unsigned char*)t->buf_blaux.p;
  unsigned long long* d_rmax = (unsigned long long*)q; q += ((sizeof(double) * nb + 255) & ~(size_t)255);
  int64_t* d_segcnt = (int64_t*)q; q += ((sizeof(int64_t) * nseg + 255) & ~(size_t)255);
  int64_t* d_segoff = (int64_t*)q;
  (void)hipMemsetAsync(d_rmax, 0, sizeof(double) * nb, s);
  (void)hipMemsetAsync(t->stats, 0, sizeof(StatsDev), s);
  dpm_stats total;
  std::memset(&total, 0, sizeof(total));
  opt.nbatch = 0;
  // the CSR of each group, joined at the end
  struct Part {
    int32_t* col = nullptr;
    double* val = nullptr;
    int64_t nnz = 0;
  };
  std::vector<Part> parts;
  struct PartsGuard {
    std::vector<Part>& v;
    ~PartsGuard() {
      for (Part& x : v) {
        if (x.col) (void)hipFree(x.col);
        if (x.val) (void)hipFree(x.val);
      }
    }
  } pguard{parts};
  int rc = DPM_OK;
  int64_t acc = 0;
  row_ptr[0] = 0;
  std::vector<int64_t> seg(nseg);
  for (int b0 = 0; b0 < nb && rc == DPM_OK; b0 += group) {
    const int b1 = std::min(nb, b0 + group), ng = b1 - b0;
    HostSource hs;
    for (int i = b0; i < b1; ++i)
      hs.add(HostEntry{opt.hist_begin + first[i], b[i].nhist, b[i].x0, b[i].y0, b[i].energy, b[i].particle, i - b0});
    const uint64_t nh = first[b1] - first[b0];
    (void)hipMemsetAsync(d_rows, 0, row_bytes * ng, s);
    const ChunkPlan plan = plan_chunks(nh, 0, opt.chunk, pipe_count(&opt));
    dpm_stats local;
    std::memset(&local, 0, sizeof(local));
    rc = opt.precision == DPM_FP64
             ? run_transport<double>(t, p, hs, &opt, t->wsd, t->td, plan, d_rows, nullptr, nullptr, nullptr, nullptr,
                                     &local)
             : run_transport<float>(t, p, hs, &opt, t->wsf, t->tf, plan, d_rows, nullptr, nullptr, nullptr, nullptr,
                                    &local);
    if (rc != DPM_OK) break;
    total.seconds += local.seconds;
    total.rounds += local.rounds;
    total.graphs += local.graphs;
    total.lds_tables = local.lds_tables;
    for (int k = 0; k < 4; ++k) {
      total.kernel_ms[k] += local.kernel_ms[k];
      total.kernel_calls[k] += local.kernel_calls[k];
    }
    const unsigned ngs = (unsigned)((size_t)ng * kCsrSeg);
    hipLaunchKernelGGL(k_bl_stats, dim3(ngs), dim3(256), 0, s, d_rows, p->nvox, d_rmax + b0, &t->stats->scored);
    hipLaunchKernelGGL(k_bl_count, dim3(ngs), dim3(256), 0, s, d_rows, p->nvox, d_rmax + b0, rel_threshold,
                       d_segcnt);
    if (hipGetLastError() != hipSuccess) { rc = set_error(DPM_E_DEVICE, "beamlet CSR count launch failed"); break; }
    // the run stream is non-blocking: copies are ordered on it
    if (hipMemcpyAsync(seg.data(), d_segcnt, sizeof(int64_t) * ngs, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
      rc = set_error(DPM_E_DEVICE, "CSR count failed");
      break;
    }
    // segment counts -> offsets within the group's arrays, and row pointers
    int64_t g = 0;
    for (int i = 0; i < ng; ++i) {
      for (int k = 0; k < kCsrSeg; ++k) {
        const int64_t c = seg[(size_t)i * kCsrSeg + k];
        seg[(size_t)i * kCsrSeg + k] = g;
        g += c;
      }
      row_ptr[b0 + i + 1] = acc + g;
    }
    acc += g;
    Part part;
    part.nnz = g;
    if (g > 0) {
      if (hipMalloc(&part.col, sizeof(int32_t) * g) != hipSuccess || hipMalloc(&part.val, sizeof(double) * g) != hipSuccess) {
        parts.push_back(part);
        rc = set_error(DPM_E_NOMEM, "cannot allocate the CSR output");
        break;
      }
      if (hipMemcpyAsync(d_segoff, seg.data(), sizeof(int64_t) * ngs, hipMemcpyHostToDevice, s) != hipSuccess) {
        parts.push_back(part);
        rc = set_error(DPM_E_DEVICE, "CSR offset copy failed");
        break;
      }
      hipLaunchKernelGGL(k_bl_fill, dim3(ngs), dim3(256), 0, s, d_rows, p->nvox, d_rmax + b0, rel_threshold, d_segoff,
                         part.col, part.val);
      // the staging vector is reused by the next group: complete the copy first
      if (hipStreamSynchronize(s) != hipSuccess) {
        parts.push_back(part);
        rc = set_error(DPM_E_DEVICE, "CSR build failed");
        break;
      }
    }
    parts.push_back(part);
  }
  if (rc == DPM_OK) rc = finish_stats(t, &total);
  if (rc == DPM_OK && acc > 0) {
    if (parts.size() == 1) {
      // one group: its arrays are the result
      t->csr_col = parts[0].col;
      t->csr_val = parts[0].val;
      parts[0].col = nullptr;
      parts[0].val = nullptr;
      t->csr_nnz = acc;
    } else if (hipMalloc(&t->csr_col, sizeof(int32_t) * acc) != hipSuccess ||
               hipMalloc(&t->csr_val, sizeof(double) * acc) != hipSuccess) {
      rc = set_error(DPM_E_NOMEM, "cannot allocate the CSR output");
    } else {
      int64_t o = 0;
      for (const Part& x : parts) {
        if (x.nnz == 0) continue;
        if (hipMemcpyAsync(t->csr_col + o, x.col, sizeof(int32_t) * x.nnz, hipMemcpyDeviceToDevice, s) != hipSuccess ||
            hipMemcpyAsync(t->csr_val + o, x.val, sizeof(double) * x.nnz, hipMemcpyDeviceToDevice, s) != hipSuccess) {
          rc = set_error(DPM_E_DEVICE, "CSR join failed");
          break;
        }
        o += x.nnz;
      }
      if (rc == DPM_OK && hipStreamSynchronize(s) != hipSuccess) rc = set_error(DPM_E_DEVICE, "CSR join failed");
      if (rc == DPM_OK) t->csr_nnz = acc;
    }
  }
  if (rc != DPM_OK) {
    if (t->csr_col) (void)hipFree(t->csr_col);
    if (t->csr_val) (void)hipFree(t->csr_val);
    t->csr_col = nullptr;
    t->csr_val = nullptr;
    t->csr_nnz = 0;
  }
  total.beamlet_groups = (int)((nb + group - 1) / group);
  if (st) *st = total;
  return rc;
}

}  // extern "C"

namespace {

// Times every candidate schedule on the sample source `hs` (device time of
// run_transport, best of two runs) and writes the fastest to *out.
template <typename R>
int autotune_impl(dpm_tables* t, const dpm_phantom* p, const HostSource& hs, int rows, int nbatch, Workspace<R>* ws,
                  Tables<R>& T, dpm_schedule* out) {
  const uint64_t n = hs.count(0) + hs.count(1);
  const int64_t nb = p->geom == GEOM_SLAB ? p->nbins : p->nvox;
  // batch statistics as the real run keeps them: the batch rows of a chunk
  // (their footprint sets the cost of the tally atomics) and the fold
  if (p->geom == GEOM_SLAB || nbatch < 2 || (uint64_t)nbatch > n) nbatch = 0;
  // the chunks of a real run of the sample (the automatic plan of each pipe count)
  ChunkPlan plan[kPipes + 1];
  int rows_max = rows;
  for (int np = 1; np <= kPipes; ++np) {
    plan[np] = plan_chunks(n, nbatch, 0, np);
    rows_max = std::max(rows_max, nbatch > 0 ? plan[np].rows * np : rows);
  }
  t->rows_zero = 0;   // candidates leave their tallies in the rows
  if (!t->buf_rows.ensure(sizeof(double) * nb * rows_max) ||
      (p->geom == GEOM_SLAB && !t->buf_steps.ensure(sizeof(double) * nb)) ||
      (nbatch > 0 && (!t->buf_sum.ensure(sizeof(double) * nb) || !t->buf_sum2.ensure(sizeof(double) * nb))))
    return set_error(DPM_E_NOMEM, "cannot allocate the tally");
  static const int kWaves[] = {8, 12, 16};
  static const int kRefill[] = {1, 4, 16};
  static const int kTally[] = {0, 1, 2};
  // device time of one schedule (best of two runs of the sample)
  auto time_schedule = [&](const dpm_run_options& opt, const ChunkPlan& cp, double* sec) -> int {
    const int np = pipe_count(&opt);
    const int rows_run = nbatch > 0 ? cp.rows * np : rows;
    if (!t->buf_rows.ensure(sizeof(double) * nb * rows_run)) return set_error(DPM_E_NOMEM, "cannot allocate the tally");
    *sec = 1e300;
    for (int rep = 0; rep < 2; ++rep) {
      (void)hipMemsetAsync(t->buf_rows.p, 0, sizeof(double) * nb * rows_run, t->stream);
      if (p->geom == GEOM_SLAB) (void)hipMemsetAsync(t->buf_steps.p, 0, sizeof(double) * nb, t->stream);
      (void)hipMemsetAsync(t->stats, 0, sizeof(StatsDev), t->stream);
      dpm_stats st;
      std::memset(&st, 0, sizeof(st));
      const int rc = run_transport<R>(t, p, hs, &opt, ws, T, cp, t->buf_rows.d(),
                                      p->geom == GEOM_SLAB ? t->buf_steps.d() : nullptr,
                                      nbatch > 0 ? t->buf_sum.d() : nullptr, nbatch > 0 ? t->buf_sum2.d() : nullptr,
                                      nullptr, &st);
      if (rc != DPM_OK) return rc;
      *sec = std::min(*sec, st.seconds);
    }
    return DPM_OK;
  };
  dpm_schedule best;
  std::memset(&best, 0, sizeof(best));
  double best_s = 1e300, default_s = 0.0;
  int cand = 0;
  dpm_run_options opt;
  std::memset(&opt, 0, sizeof(opt));
  opt.seed = 0x5eed;
  opt.hist_begin = 0;
  opt.nbatch = nbatch;
  // stage 1: pipes x waves per CU x refill threshold x tally mode, each with
  // the automatic chunk plan of its pipe count
  for (int np = 1; np <= kPipes; ++np)
    for (int w : kWaves)
      for (int r : kRefill)
        for (int tm : kTally) {
          if (tm != 0 && p->geom == GEOM_SLAB) continue;  // the slab tally is privatised in LDS either way
          opt.chunk = 0;
          opt.waves_per_cu = w;
          opt.refill_min = r;
          opt.tally_mode = tm;
          opt.pipes = np;
          double sec;
          const int rc = time_schedule(opt, plan[np], &sec);
          if (rc != DPM_OK) return rc;
          ++cand;
          if (np == 1 && w == 16 && r == 1 && tm == 0) default_s = sec;
          if (sec < best_s) {
            best_s = sec;
            best.waves_per_cu = w;
            best.refill_min = r;
            best.tally_mode = tm;
            best.pipes = np;
          }
        }
  // stage 2: the chunk (histories per queue fill) at the best schedule --
  // caps of 2^21..2^24 histories that split the sample into two or more
  // chunks, each holding at least one whole batch; 0 keeps the automatic plan
  best.chunk = 0;
  opt.waves_per_cu = best.waves_per_cu;
  opt.refill_min = best.refill_min;
  opt.tally_mode = best.tally_mode;
  opt.pipes = best.pipes;
  for (int64_t c = kMinTunedChunk; c <= kDefaultChunk; c <<= 1) {
    if ((uint64_t)c >= n || c == plan[best.pipes].chunk) continue;
    if (nbatch > 0 && (uint64_t)c < (n + (uint64_t)nbatch - 1) / (uint64_t)nbatch) continue;
    opt.chunk = c;
    double sec;
    const int rc = time_schedule(opt, plan_chunks(n, nbatch, c, best.pipes), &sec);
    if (rc != DPM_OK) return rc;
    ++cand;
    if (sec < 0.98 * best_s) {   // a fixed chunk must win by more than the timing noise
      best_s = sec;
      best.chunk = c;
    }
  }
  best.candidates = cand;
  best.histories_per_s = (double)n / best_s;
  best.default_histories_per_s = default_s > 0 ? (double)n / default_s : 0.0;
  *out = best;
  return DPM_OK;
}

}  // namespace

extern "C" {

int dpm_autotune(dpm_tables* t, const dpm_phantom* p, const dpm_source* src, uint64_t nsample, int precision,
                 int nbatch, dpm_schedule* best) {
  if (!t || !p || !src || !best) return set_error(DPM_E_INVALID, "null argument");
  if (p->owner != t) return set_error(DPM_E_INVALID, "phantom belongs to another tables handle");
  if (nsample < 1000 || nsample > ((uint64_t)1 << 24)) return set_error(DPM_E_INVALID, "sample of 1e3 .. 2^24 histories");
  if (precision != DPM_FP32 && precision != DPM_FP64) return set_error(DPM_E_INVALID, "unknown precision");
  int rc = validate_source(t, p, src);
  if (rc != DPM_OK) return rc;
  HIP_TRY(hipSetDevice(t->device));
  const HostSource hs = source_from(src, nsample, 0);
  return precision == DPM_FP64 ? autotune_impl<double>(t, p, hs, 1, nbatch, t->wsd, t->td, best)
                               : autotune_impl<float>(t, p, hs, 1, nbatch, t->wsf, t->tf, best);
}

int dpm_autotune_beamlets(dpm_tables* t, const dpm_phantom* p, const dpm_beamlet* b, int nb, uint64_t nsample,
                          int precision, dpm_schedule* best) {
  if (!t || !p || !b || nb <= 0 || !best) return set_error(DPM_E_INVALID, "null argument");
  if (p->owner != t) return set_error(DPM_E_INVALID, "phantom belongs to another tables handle");
  if (p->geom != GEOM_VOXEL) return set_error(DPM_E_INVALID, "beamlets need a voxel phantom");
  if (nb > (1 << 20)) return set_error(DPM_E_INVALID, "too many beamlets");
  if (nsample < 1000 || nsample > ((uint64_t)1 << 24)) return set_error(DPM_E_INVALID, "sample of 1e3 .. 2^24 histories");
  if (precision != DPM_FP32 && precision != DPM_FP64) return set_error(DPM_E_INVALID, "unknown precision");
  // the beamlets of one group of a real run (dense rows within the default
  // budget), each with its history count scaled so that the total is ~nsample
  nb = (int)std::min<size_t>((size_t)nb, std::max<size_t>(1, kBeamletBudget / (sizeof(double) * (size_t)p->nvox)));
  uint64_t total = 0;
  for (int i = 0; i < nb; ++i) {
    int rc = check_source(t, b[i].particle, b[i].energy, 0, 0);
    if (rc != DPM_OK) return rc;
    total += b[i].nhist;
  }
  if (total == 0) return set_error(DPM_E_INVALID, "beamlets without histories");
  HostSource hs;
  uint64_t first = 0;
  for (int i = 0; i < nb; ++i) {
    const uint64_t n = std::max<uint64_t>(1, (uint64_t)((double)b[i].nhist * (double)nsample / (double)total));
    hs.add(HostEntry{first, n, b[i].x0, b[i].y0, b[i].energy, b[i].particle, i});
    first += n;
  }
  HIP_TRY(hipSetDevice(t->device));
  return precision == DPM_FP64 ? autotune_impl<double>(t, p, hs, nb, 0, t->wsd, t->td, best)
                               : autotune_impl<float>(t, p, hs, nb, 0, t->wsf, t->tf, best);
}

int dpm_beamlets_fetch(dpm_tables* t, int32_t* col, double* val) {
  if (!t || !col || !val) return set_error(DPM_E_INVALID, "null argument");
  if (t->csr_nnz == 0) return DPM_OK;
  HIP_TRY(hipSetDevice(t->device));
  HIP_TRY(hipMemcpy(col, t->csr_col, sizeof(int32_t) * t->csr_nnz, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(val, t->csr_val, sizeof(double) * t->csr_nnz, hipMemcpyDeviceToHost));
  return DPM_OK;
}

int dpm_sample(dpm_tables* t, int kind, int precision, int n, const double* energy, const int* imat,
               const uint32_t* r1, const uint32_t* r2, const uint32_t* r3, double* out, int32_t* ebin, int32_t* abin) {
  if (!t || !energy || !r1 || !r2 || !out || !ebin || !abin || n < 0) return set_error(DPM_E_INVALID, "null argument");
  if (kind < 0 || kind > 4) return set_error(DPM_E_INVALID, "unknown sampler");
  if (kind <= 2 && !r3) return set_error(DPM_E_INVALID, "sampler needs three uniforms");
  if (kind == 0 && !imat) return set_error(DPM_E_INVALID, "SB sampler needs material indices");
  if (n == 0) return DPM_OK;
  for (int i = 0; i < n; ++i) {
    if ((r1[i] >> 24) || (r2[i] >> 24) || (r3 && (r3[i] >> 24))) return set_error(DPM_E_INVALID, "uniforms must be 24-bit");
    if (kind == 0 && (imat[i] < 0 || imat[i] >= t->nmat)) return set_error(DPM_E_INVALID, "material index out of range");
    if (!(energy[i] > 0.0) || !(energy[i] < 21.0)) return set_error(DPM_E_INVALID, "energy outside the tables");
  }
  HIP_TRY(hipSetDevice(t->device));
  const hipStream_t s = t->stream;
  void* buf = nullptr;
  const size_t bytes = (size_t)n * (8 + 4 + 4 * 3 + 8 + 4 + 4) + 8 * 256;
  HIP_TRY(hipMalloc(&buf, bytes));
  unsigned char* q = (unsigned char*)buf;
  auto take = [&](size_t nbytes) { unsigned char* r = q; q += (nbytes + 255) & ~(size_t)255; return r; };
  double* de = (double*)take(8 * (size_t)n);
  int* dm = (int*)take(4 * (size_t)n);
  uint32_t* d1 = (uint32_t*)take(4 * (size_t)n);
  uint32_t* d2 = (uint32_t*)take(4 * (size_t)n);
  uint32_t* d3 = (uint32_t*)take(4 * (size_t)n);
  double* dout = (double*)take(8 * (size_t)n);
  int* deb = (int*)take(4 * (size_t)n);
  int* dab = (int*)take(4 * (size_t)n);
  int rc = DPM_OK;
  hipError_t e = hipMemcpy(de, energy, 8 * (size_t)n, hipMemcpyHostToDevice);
  // memsets on the run stream: a null-stream memset is not ordered before a
  // launch on the non-blocking run stream
  if (e == hipSuccess) e = imat ? hipMemcpy(dm, imat, 4 * (size_t)n, hipMemcpyHostToDevice) : hipMemsetAsync(dm, 0, 4 * (size_t)n, s);
  if (e == hipSuccess) e = hipMemcpy(d1, r1, 4 * (size_t)n, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(d2, r2, 4 * (size_t)n, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = r3 ? hipMemcpy(d3, r3, 4 * (size_t)n, hipMemcpyHostToDevice) : hipMemsetAsync(d3, 0, 4 * (size_t)n, s);
  if (e == hipSuccess)
    e = precision == DPM_FP64 ? launch_sample_f64(t->td, kind, n, de, dm, d1, d2, d3, dout, deb, dab, s)
                              : launch_sample_f32(t->tf, kind, n, de, dm, d1, d2, d3, dout, deb, dab, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e == hipSuccess) e = hipMemcpy(out, dout, 8 * (size_t)n, hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(ebin, deb, 4 * (size_t)n, hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(abin, dab, 4 * (size_t)n, hipMemcpyDeviceToHost);
  if (e != hipSuccess) rc = set_error(DPM_E_DEVICE, std::string("dpm_sample: ") + hipGetErrorString(e));
  (void)hipFree(buf);
  return rc;
}

int dpm_perform(dpm_tables* t, int kind, int precision, int n, double energy, int imat, uint64_t seed,
                double* out_value, double* out_cos) {
  if (!t || !out_value || !out_cos || n < 0) return set_error(DPM_E_INVALID, "null argument");
  if (kind < 0 || kind > 3) return set_error(DPM_E_INVALID, "unknown interaction");
  if (precision != DPM_FP32 && precision != DPM_FP64) return set_error(DPM_E_INVALID, "unknown precision");
  if (kind == 0 && (imat < 0 || imat >= t->nmat)) return set_error(DPM_E_INVALID, "material index out of range");
  // the interaction must be possible at this energy (SimDPMLike.cc:174, 206)
  const double lo = kind == 0 ? t->gcut : (kind == 1 ? 2.0 * t->ecut : (kind == 2 ? t->ecut : t->gcut));
  if (!(energy > lo) || !(energy < 21.0)) return set_error(DPM_E_INVALID, "energy outside the interaction's range");
  if (n == 0) return DPM_OK;
  HIP_TRY(hipSetDevice(t->device));
  const hipStream_t s = t->stream;
  void* buf = nullptr;
  const size_t half = ((size_t)n * 8 + 255) & ~(size_t)255;
  HIP_TRY(hipMalloc(&buf, 2 * half));
  double* dv = (double*)buf;
  double* dc = (double*)((unsigned char*)buf + half);
  hipError_t e = precision == DPM_FP64 ? launch_perform_f64(t->td, kind, n, energy, imat, seed, dv, dc, s)
                                       : launch_perform_f32(t->tf, kind, n, energy, imat, seed, dv, dc, s);
  if (e == hipSuccess) e = hipMemcpyAsync(out_value, dv, 8 * (size_t)n, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipMemcpyAsync(out_cos, dc, 8 * (size_t)n, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  (void)hipFree(buf);
  if (e != hipSuccess) return set_error(DPM_E_DEVICE, std::string("dpm_perform: ") + hipGetErrorString(e));
  return DPM_OK;
}

int dpm_eval(dpm_tables* t, int which, int precision, int n, const double* energy, const int* imat, double* out) {
  if (!t || !energy || !out || n < 0) return set_error(DPM_E_INVALID, "null argument");
  if (which < 0 || which > 8) return set_error(DPM_E_INVALID, "unknown table");
  if (n == 0) return DPM_OK;
  for (int i = 0; i < n; ++i) {
    if (imat && (imat[i] < 0 || imat[i] >= t->nmat)) return set_error(DPM_E_INVALID, "material index out of range");
    if (!(energy[i] > 0.0) || !std::isfinite(energy[i])) return set_error(DPM_E_INVALID, "energy must be positive");
  }
  HIP_TRY(hipSetDevice(t->device));
  const hipStream_t s = t->stream;
  void* buf = nullptr;
  HIP_TRY(hipMalloc(&buf, (size_t)n * 20 + 1024));
  double* de = (double*)buf;
  int* dm = (int*)((unsigned char*)buf + (((size_t)n * 8 + 255) & ~(size_t)255));
  double* dout = (double*)((unsigned char*)dm + (((size_t)n * 4 + 255) & ~(size_t)255));
  int rc = DPM_OK;
  hipError_t e = hipMemcpy(de, energy, 8 * (size_t)n, hipMemcpyHostToDevice);
  // memsets on the run stream: a null-stream memset is not ordered before a
  // launch on the non-blocking run stream
  if (e == hipSuccess) e = imat ? hipMemcpy(dm, imat, 4 * (size_t)n, hipMemcpyHostToDevice) : hipMemsetAsync(dm, 0, 4 * (size_t)n, s);
  if (e == hipSuccess)
    e = precision == DPM_FP64 ? launch_eval_f64(t->td, which, n, de, dm, dout, s) : launch_eval_f32(t->tf, which, n, de, dm, dout, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e == hipSuccess) e = hipMemcpy(out, dout, 8 * (size_t)n, hipMemcpyDeviceToHost);
  if (e != hipSuccess) rc = set_error(DPM_E_DEVICE, std::string("dpm_eval: ") + hipGetErrorString(e));
  (void)hipFree(buf);
  return rc;
}

}  // extern "C"
