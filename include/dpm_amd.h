/*
 * dpm_amd.h -- C ABI of the MI355X DPM transport engine (libdpm_amd.so).
 *
 * The engine re-implements, as gfx950 HIP kernels, the per-history coupled
 * electron-photon transport of the reference's dpm_Simulate:
 *   Simulate(...)                   Simulation/inc/SimDPMLike.hh:46, SimDPMLike.cc:40-251
 * called from the reference CLI at dpm_Simulate.cc:82 after the three data
 * objects were loaded (dpm_Simulate.cc:71-78). Everything crossing this ABI is
 * plain C: pointers, sizes and status codes. No C++ exception crosses it.
 *
 * Conventions
 *  - Every function returns DPM_OK (0) or a negative DPM_E* code; the text of
 *    the last error of the calling thread is dpm_last_error().
 *  - Handles are thread-compatible, not thread-safe: one thread at a time per
 *    handle. Each tables handle owns one HIP stream and the scratch buffers
 *    of the runs made with it.
 *  - Lengths are in mm, energies in MeV, densities in g/cm3 (as the reference).
 */
#ifndef DPM_AMD_H
#define DPM_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DPM_ABI_VERSION 4

enum {
  DPM_OK = 0,
  DPM_E_INVALID = -1,   /* bad argument */
  DPM_E_IO = -2,        /* table file missing or malformed */
  DPM_E_DEVICE = -3,    /* HIP runtime error / no device */
  DPM_E_OVERFLOW = -4,  /* a particle queue overflowed (reduce histories per chunk) */
  DPM_E_NOMEM = -5
};

enum { DPM_FP32 = 0, DPM_FP64 = 1 };              /* transport arithmetic */
enum { DPM_ELECTRON = -1, DPM_PHOTON = 0, DPM_POSITRON = 1 };

/* ---------------------------------------------------------------------------
 * Tables. The descriptor carries the raw node data of the reference's
 * run-time table classes; the library derives the interpolation coefficients
 * (as SimDataSpline::SetUp does, SimDataSpline.cc:33-67) and the device packs.
 * ------------------------------------------------------------------------- */
typedef struct dpm_spline_desc {   /* one log-grid spline: SimDataSpline (SimDataSpline.hh:38-50) */
  int n;                           /* nodes */
  const double* x;                 /* [n] energies, log-spaced */
  const double* y;                 /* [n] values */
} dpm_spline_desc;

typedef struct dpm_linear_desc {   /* one linear-grid table: SimDataLinear (SimDataLinear.hh:38-45) */
  int n;
  const double* x;
  const double* y;
} dpm_linear_desc;

typedef struct dpm_alias_desc {    /* a family of linear-pdf alias tables: SimLinAliasData */
  int nmat;                        /* materials (1 for Moller/KN) */
  int ne;                          /* primary energies (log grid) */
  int ns;                          /* points per table */
  const double* energy;            /* [ne] */
  const double* x;                 /* [nmat*ne*ns] */
  const double* y;                 /* [nmat*ne*ns] */
  const double* w;                 /* [nmat*ne*ns] alias keep-probability */
  const int*    alias;             /* [nmat*ne*ns] alias index */
} dpm_alias_desc;

typedef struct dpm_gs_desc {       /* SimGSTables (SimGSTables.cc:19-80), reference material */
  int ne, ns;
  const double* energy;            /* [ne] */
  const double* transf;            /* [ne] transformation parameter a(E_i) */
  const double* u;                 /* [ne*ns] */
  const double* a;                 /* [ne*ns] */
  const double* b;                 /* [ne*ns] */
} dpm_gs_desc;

typedef struct dpm_table_desc {
  int nmat;
  double electron_cut, gamma_cut;          /* SimMaterialData::fElectronCut/fGammaCut */
  const double* density;                   /* [nmat] SimMaterialData::fMaterialDensity */
  const double* moller_scale;              /* [nmat] SimMaterialData::fMollerIMFPScaling */
  const dpm_spline_desc* itr1;             /* [nmat] SimITr1MFPElastic (scattering-power corrected) */
  const dpm_spline_desc* dedx;             /* [nmat] SimStoppingPower */
  const dpm_spline_desc* brem;             /* [nmat] SimIMFPBrem */
  dpm_spline_desc moller;                  /* SimIMFPMoller (reference material) */
  dpm_spline_desc k1;                      /* SimMaxScatStrength (reference material) */
  dpm_linear_desc gmax;                    /* SimIMFPMaxPhoton */
  const dpm_linear_desc* ptotal;           /* [nmat] SimIMFPPhoton(0) */
  const dpm_linear_desc* pcompton;         /* [nmat] SimIMFPPhoton(1) */
  const dpm_linear_desc* ppair;            /* [nmat] SimIMFPPhoton(2) */
  dpm_alias_desc sb;                       /* SimSBTables */
  dpm_alias_desc moller_tab;               /* SimMollerTables */
  dpm_alias_desc kn;                       /* SimKNTables */
  dpm_gs_desc gs;                          /* SimGSTables */
} dpm_table_desc;

typedef struct dpm_tables dpm_tables;

/* Replaces SimElectronData::Load + SimPhotonData::Load + SimMaterialData::Load
 * (SimElectronData.cc:44-115, SimPhotonData.cc:28-78, SimMaterialData.cc:6-53)
 * followed by the upload of their content to the device. */
int  dpm_tables_create(const dpm_table_desc* desc, int device, dpm_tables** out);
/* Same, reading the reference's 14 table files in `dir` with the host readers
 * (identical file names, comment-line counts and fields). */
int  dpm_tables_load(const char* dir, int device, dpm_tables** out);
void dpm_tables_destroy(dpm_tables* t);
int  dpm_tables_num_materials(const dpm_tables* t);

/* ---------------------------------------------------------------------------
 * Geometry. Replaces Geom (Simulation/src/Geom.cc, Geom.hh:52-101).
 * ------------------------------------------------------------------------- */
typedef struct dpm_phantom dpm_phantom;

/* The reference's pre-defined slab geometry: cubic boxes of size `lbox`,
 * material by depth index per `geom_index` 0..3 (Geom.cc:133-164), lateral
 * extent +-100 mm, depth extent [-lbox, 100] mm (Geom.cc:25), depth-dose
 * scoring (Geom.cc:120-130). */
int  dpm_phantom_create_slab(const dpm_tables* t, double lbox, int geom_index, dpm_phantom** out);
/* A voxel phantom of nx*ny*nz cubic voxels of size `lbox`, centred on the
 * origin; the source plane is its -z face. `mat` (material indices) and `rho`
 * (g/cm3) are [nz][ny][nx] (x fastest). Leaving the box ends a track. */
int  dpm_phantom_create_voxel(const dpm_tables* t, int nx, int ny, int nz, double lbox,
                              const uint8_t* mat, const float* rho, dpm_phantom** out);
void dpm_phantom_destroy(dpm_phantom* p);
/* Woodcock majorant scale of the photon steps (KeepTrackingPhoton,
 * SimDPMLike.cc:467-510, tracks against the table maximum
 * SimIMFPMaxPhoton::GetIMFP scaled by this factor). A voxel phantom sets it
 * to max(rho_voxel / rho_nominal) >= 1, so that per-voxel densities above
 * the nominal one stay under the majorant; the slab keeps 1. Any larger
 * scale gives the same dose distribution with more null events; a smaller
 * one is rejected. No counterpart in the reference (whose densities are the
 * nominal ones). */
int dpm_phantom_set_majorant(dpm_phantom* p, double scale);
double dpm_phantom_get_majorant(const dpm_phantom* p);
/* number of tally bins: depth bins (slab) or voxels */
int64_t dpm_phantom_num_bins(const dpm_phantom* p);

/* ---------------------------------------------------------------------------
 * Sources and runs
 * ------------------------------------------------------------------------- */
enum { DPM_SRC_PENCIL = 0, DPM_SRC_FIELD = 1, DPM_SRC_GANTRY = 2 };
enum { DPM_SPEC_FLAT = 0, DPM_SPEC_TRIANGULAR = 1, DPM_SPEC_TABLE = 2 };

typedef struct dpm_source {        /* SimDPMLike.cc:64-75 (pencil), extended */
  int kind;                        /* DPM_SRC_* */
  int particle;                    /* DPM_ELECTRON or DPM_PHOTON */
  double energy;                   /* MeV (mono-energetic when emax <= emin) */
  double emin, emax;               /* spectrum over [emin, emax] when emax > emin */
  double x0, y0;                   /* entry point / field centre [mm] */
  double field_x, field_y;         /* field size [mm] (DPM_SRC_FIELD, DPM_SRC_GANTRY; 0 = pencil) */
  /* DPM_SRC_GANTRY (voxel phantoms): beam direction (sin a, 0, cos a) through
     the phantom centre, the field (or pencil) on the plane through the
     centre perpendicular to it, with axes (cos a, 0, -sin a) and (0, 1, 0);
     primaries start where their ray enters the phantom. a = 0 is the +z
     field of DPM_SRC_FIELD. */
  double angle_deg;
  int spectrum;                    /* DPM_SPEC_FLAT, or DPM_SPEC_TRIANGULAR(emin, emode, emax), or
                                      DPM_SPEC_TABLE */
  double emode;
  /* DPM_SPEC_TABLE (a spectrum file): spec_n histogram bins, bin i covering
     [spec_edges[i], spec_edges[i+1]) MeV with relative weight
     spec_weights[i] >= 0 (edges increasing, weights not all zero; emin and
     emax are not used). A primary's energy takes one uniform u, as the flat
     spectrum does: the bin whose cumulative weight interval holds u, then the
     same fraction of that bin's width. The arrays are read during the call. */
  int spec_n;
  const double* spec_edges;        /* [spec_n + 1] */
  const double* spec_weights;      /* [spec_n] */
  /* Collimation of DPM_SRC_GANTRY fields (0 = off: the parallel field above).
     sad > 0: the particles come from a point source on the beam axis at
     distance sad [mm] upstream of the phantom centre and travel through
     their field point, so the field is the beam's cross-section on the plane
     through the centre (the isocentre plane); sad must put the source
     outside the phantom. jaws != 0: a pair of X jaws and a pair of Y jaws,
     given by their projections on that plane, leave the aperture
     jaw_x1 <= fu < jaw_x2, jaw_y1 <= fv < jaw_y2, where (fu, fv) is the field
     point relative to the field centre (its axes as above); a particle aimed
     outside it crosses one jaw of each pair whose aperture it misses (n = 1
     or 2) and goes on unscattered with probability
     exp(-n jaw_tau / cos theta) -- theta its angle to the beam axis, jaw_tau
     a jaw's optical thickness along the axis -- or is absorbed in the jaws
     (its energy counts as escaped). The transmission test takes one uniform,
     after the source's other draws, only for particles aimed outside. Other
     source kinds must leave sad and jaws zero. */
  double sad;
  int jaws;
  double jaw_x1, jaw_x2, jaw_y1, jaw_y2;
  double jaw_tau;
} dpm_source;

typedef struct dpm_run_options {
  uint64_t seed;                   /* Philox key */
  uint64_t hist_begin;             /* id of the first history (stream selection) */
  int precision;                   /* DPM_FP32 (default) or DPM_FP64 */
  int nbatch;                      /* >1: per-batch sum and sum of squares of the voxel tally */
  int64_t chunk;                   /* histories per queue fill; 0 = automatic (one pipe: up to 16M
                                      histories; two pipes: an even number of chunks of up to 8M).
                                      With nbatch > 1 a chunk holds whole batches, so it exceeds
                                      this when one batch does (queues are sized for it) */
  int waves_per_cu;                /* persistent waves per CU; 0 = automatic */
  double* hist_edep;               /* optional [nhist] per-history scored energy (host) */
  int refill_min;                  /* idle lanes of a wave that trigger its refill (1..64); 0 = 1 */
  int tally_mode;                  /* voxel tally: 0 per-lane write-combining, 1 atomic per deposit,
                                      2 write-combining then a per-wave LDS table of hot cells */
  int graph;                       /* 1: replay the transport rounds from captured hipGraphs (no
                                      per-kernel timings in dpm_stats) */
  int rounds_per_sync;             /* transport rounds enqueued (or captured) per host synchronisation:
                                      even, 2..8; 0 = 8 */
  int table_memory;                /* DPM_TABLES_AUTO (0): FP32 runs copy the small tables into each
                                      block's LDS when they fit -- the electron spline tables
                                      (3*nmat+2 tables of 128 nodes: up to 3 materials) and the photon
                                      IMFP tables (2+3*nmat arrays of 1024 values: up to 3 materials);
                                      DPM_TABLES_GLOBAL (1): read them from device memory. Both give
                                      identical results. */
  int pipes;                       /* 1 (0 = 1) or 2: chunks run on that many pipes -- streams with
                                      their own particle queues and batch rows -- side by side, so
                                      that one pipe's kernel tails, drain rounds and folds overlap the
                                      other's transport; the same results either way */
  int64_t tally_budget;            /* dpm_simulate_beamlets: device bytes for the dense FP64 rows of
                                      the beamlets run together (0 = 16 GiB); beamlets beyond
                                      budget / (8 * voxels) run in further groups, with the same
                                      history ids and the same CSR */
} dpm_run_options;

enum { DPM_TABLES_AUTO = 0, DPM_TABLES_GLOBAL = 1 };

enum {
  DPM_CNT_VOXEL_VISITS = 0, DPM_CNT_DEPOSITS, DPM_CNT_SB, DPM_CNT_MOLLER, DPM_CNT_KN, DPM_CNT_GS,
  DPM_CNT_SECONDARIES, DPM_CNT_E_STEPS, DPM_CNT_G_STEPS, DPM_CNT_TALLY_ATOMICS, DPM_NCNT
};

typedef struct dpm_stats {
  double seconds;                  /* device time of the run (events on the run stream) */
  double e_scored, e_escaped;      /* MeV */
  double e_primary;                /* MeV, sum of the primaries' kinetic energies */
  uint64_t n_pair, n_annih;
  uint64_t counters[DPM_NCNT];     /* byte-model counters */
  int rounds;                      /* queue rounds enqueued (including empty ones after the drain) */
  int hist_size;                   /* slab: rows written by the reference rule */
  double kernel_ms[4];             /* device time in source (0: primaries are generated by the transport
                                      kernels), photon, electron, tally-fold kernels; eager runs only */
  int kernel_calls[4];             /* launches of each of those kernels */
  uint64_t counters_photon[DPM_NCNT];    /* counters of the photon kernel */
  uint64_t counters_electron[DPM_NCNT];  /* counters of the electron kernel */
  int graphs;                      /* hipGraphs captured by the run (graph = 1) */
  int lds_tables;                  /* bit 0: the electron kernel read its spline tables from LDS;
                                      bit 1: the photon kernel its linear IMFP tables */
  uint64_t lane_slots[2];          /* photon / electron kernels: 64 per transport-loop trip of a
                                      wave; the steps (counters_*[DPM_CNT_G_STEPS / _E_STEPS]) over
                                      this are the lane utilisation */
  int beamlet_groups;              /* dpm_simulate_beamlets: groups the beamlets ran in */
  uint64_t guard_exits;            /* particles ended by a termination guard instead of the physics:
                                      the per-particle cap of 2e5 transport steps, or a boundary /
                                      locate loop out of tolerance pushes (16); their energy is counted
                                      as escaped. 0 in every run that follows the reference */
} dpm_stats;

/* Replaces Simulate(nprimary, eprimary, iselectron, lbox, mat, el, ph,
 * geomIndex, filename) (SimDPMLike.hh:46). Runs `nhist` histories of the
 * source in the phantom and returns the tally in caller-owned host arrays:
 *   slab:  tally[nbins] = energy per depth bin (MeV, summed over histories),
 *          steps[nbins] = number of scoring calls per bin (may be NULL);
 *   voxel: tally[nbins] = energy per voxel (MeV); with nbatch>1, steps (may be
 *          NULL) receives the per-voxel sum over batches of the squared batch
 *          tally.
 * tally may be NULL: the result then stays on the device (until the next run
 * on this handle) and dpm_tally_fetch copies it out. */
int dpm_simulate(dpm_tables* t, const dpm_phantom* p, const dpm_source* src, uint64_t nhist,
                 const dpm_run_options* opt, double* tally, double* steps, dpm_stats* stats);

/* Several fields of one source (treatment-plan style: gantry angles, uneven
 * history counts) in one run into one tally, as dpm_simulate: field i runs
 * fields[i].nhist histories with its own entry point and angle, history ids
 * consecutive from opt->hist_begin in field order (so the result equals the
 * sum of dpm_simulate calls with those hist_begin values). `src` gives kind,
 * particle, energy or spectrum and field size; its x0, y0, angle_deg are
 * ignored. No counterpart in the reference (one pencil beam per run). */
typedef struct dpm_field {
  double x0, y0;
  double angle_deg;
  uint64_t nhist;
} dpm_field;

int dpm_simulate_fields(dpm_tables* t, const dpm_phantom* p, const dpm_source* src, const dpm_field* fields, int nf,
                        const dpm_run_options* opt, double* tally, double* steps, dpm_stats* stats);

/* No counterpart in the reference: the form of dpm_simulate for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc). It calls no hipMalloc,
 * hipFree or allocating library routine: the particle queues and the source
 * description live in a static device array of 128 MiB per GPU (held from the
 * library's first use on that GPU until the process ends), and the results go
 * to caller-owned DEVICE arrays of dpm_phantom_num_bins doubles, which the
 * call clears first: d_tally (energy per depth bin or voxel, MeV) and, for the
 * slab, d_steps (scoring calls per bin; required there, ignored for voxels).
 * Same kernels and per-history results as dpm_simulate. Limits, each an
 * error: no batch statistics (nbatch <= 1), no per-history tally, one pipe,
 * no graph replay, a spectrum table of at most ~4000 bins (64 KiB of source
 * description), and queue chunks of at most what the scratch holds at the
 * launch grid (waves_per_cu 0 means 4 here; chunk 0 takes the maximum: on
 * 256 CUs about 31000 FP32 / 12700 FP64 histories at 4 waves per CU, 19000
 * FP32 at 8; runs of any length are cut into such chunks). Every call on a GPU shares that GPU's scratch: calls that
 * overlap on one GPU (from several threads or tables handles) race. The call
 * returns when the run is complete. */
int dpm_simulate_noalloc(dpm_tables* t, const dpm_phantom* p, const dpm_source* src, uint64_t nhist,
                         const dpm_run_options* opt, double* d_tally, double* d_steps, dpm_stats* stats);

/* Phase-space source: a list of particles (a phase-space file, e.g. scored
 * below a treatment head), each one history that starts where and as the
 * record says -- no counterpart in the reference, whose source is the fixed
 * pencil of SimDPMLike.cc:64-75 (one primary per history at a given position,
 * direction and energy; a record generalises it). The tally is returned as by
 * dpm_simulate (with nbatch > 1 on voxel phantoms, `steps` receives the sum
 * of squares). History ids: the photon records first, then the e-/e+
 * records, each in the order given, from opt->hist_begin; each history's
 * random stream starts unused (no source draws). Positions are phantom
 * coordinates (voxel phantoms: centred on the origin; slab: the reference's
 * Geom frame). In a voxel phantom a record outside the phantom moves along its
 * direction to where it enters (1e-3 mm inside); a record whose ray misses the
 * phantom, or a slab record outside the slab's extent, leaves at once (its
 * energy counts as escaped). Directions must be unit vectors (within 1e-6),
 * energies above the tracking cut of the particle and below the tables'
 * maximum; at most 2^30 records per call. */
typedef struct dpm_phsp_particle {
  int particle;                    /* DPM_ELECTRON, DPM_PHOTON or DPM_POSITRON */
  double energy;                   /* MeV */
  double x, y, z;                  /* mm */
  double u, v, w;                  /* direction cosines */
} dpm_phsp_particle;

int dpm_simulate_phsp(dpm_tables* t, const dpm_phantom* p, const dpm_phsp_particle* rec, uint64_t n,
                      const dpm_run_options* opt, double* tally, double* steps, dpm_stats* stats);

/* Copies the tally of the last dpm_simulate on this handle (and, if
 * `second` is not NULL, its steps / sum-of-squares array) into host arrays of
 * dpm_phantom_num_bins elements. No counterpart in the reference (which writes
 * hist.sim directly); it separates the run from the result transfer. */
int dpm_tally_fetch(dpm_tables* t, double* tally, double* second);

/* Batched beamlets: `nb` pencil beamlets (entry point, particle, energy,
 * history count) in one run -- or in groups of as many as the dense rows of
 * opt->tally_budget hold, each group one run -- scored into a sparse
 * dose-influence matrix:
 * for each beamlet the voxels above `rel_threshold` of its maximum, as CSR
 * (row_ptr[nb+1]; nnz = row_ptr[nb]). The column indices and values stay on
 * the device in the tables handle until dpm_beamlets_fetch copies them into
 * caller arrays of nnz entries (col: voxel index, val: MeV). */
typedef struct dpm_beamlet {
  double x0, y0;
  int particle;
  double energy;
  uint64_t nhist;
} dpm_beamlet;

int dpm_simulate_beamlets(dpm_tables* t, const dpm_phantom* p, const dpm_beamlet* b, int nb,
                          const dpm_run_options* opt, double rel_threshold,
                          int64_t* row_ptr, dpm_stats* stats);
int dpm_beamlets_fetch(dpm_tables* t, int32_t* col, double* val);

/* ---------------------------------------------------------------------------
 * Autotuner (no counterpart in the reference, whose transport has no launch
 * schedule): times candidate schedules of the persistent kernels -- pipes,
 * waves per CU, refill threshold (compaction interval) and voxel-tally mode,
 * each with the automatic chunk plan of its pipe count; then, at the best of
 * those, the queue chunk size (caps of 2^21..2^24 histories that leave at
 * least two chunks of whole batches) -- on a sample of `nsample` histories of
 * the given source (or beamlets, scaled to about nsample histories in total)
 * and returns the fastest. `chunk` is 0 when the automatic plan won (a fixed
 * chunk must be faster by 2 %). Copy the fields into dpm_run_options for the
 * real run; sample with the real run's batch size so that a chunk holds the
 * same whole batches in both.
 * ------------------------------------------------------------------------- */
typedef struct dpm_schedule {
  int64_t chunk;
  int waves_per_cu;
  int refill_min;
  int tally_mode;
  int candidates;                  /* schedules timed */
  double histories_per_s;          /* device rate of the best on the sample */
  double default_histories_per_s;  /* device rate of the default schedule on the sample */
  int pipes;                       /* 1 or 2 (dpm_run_options.pipes) */
} dpm_schedule;

/* nbatch (> 1, voxel phantoms): time the candidates with the batch
 * statistics of a run with that many batches (their tally rows and folds). */
int dpm_autotune(dpm_tables* t, const dpm_phantom* p, const dpm_source* src, uint64_t nsample, int precision,
                 int nbatch, dpm_schedule* best);
int dpm_autotune_beamlets(dpm_tables* t, const dpm_phantom* p, const dpm_beamlet* b, int nb, uint64_t nsample,
                          int precision, dpm_schedule* best);

/* ---------------------------------------------------------------------------
 * Interaction samplers with explicit uniforms (the reference's own model
 * tests drive PerformBrem/PerformMoller/PerformMSCAngularDeflection this way,
 * testAPI.cc:92-199). Uniforms are 24-bit integers k (uniform = k * 2^-24).
 *   kind 0 SB    : SimSBTables::SampleEnergyTransfer(E, imat, r1, r2, r3)   SimSBTables.cc:93
 *   kind 1 Moller: SimMollerTables::SampleEnergyTransfer(E, r1, r2, r3)    SimMollerTables.cc:69
 *   kind 2 KN    : SimKNTables::SampleEnergyTransfer(E, r1, r2, r3)        SimKNTables.cc:69
 *   kind 3 GS    : SimGSTables::SampleAngularDeflection(E, r1, r2)         SimGSTables.cc:84
 *   kind 4 GS    : the same sample as 1 - cos(theta) (the quantity the MSC
 *                  hinge applies; free of the cancellation of 1 - out[i] of kind 3)
 * out[i] is the sampled value; ebin[i] the chosen energy-grid table and
 * abin[i] the alias (or cumulative) bin, for exactness checks. */
int dpm_sample(dpm_tables* t, int kind, int precision, int n, const double* energy, const int* imat,
               const uint32_t* r1, const uint32_t* r2, const uint32_t* r3,
               double* out, int32_t* ebin, int32_t* abin);

/* Final states of `n` single interactions of a primary of `energy` moving
 * along +z in material `imat` (used by kind 0 only), each with its own
 * Philox stream (key `seed`, history = sample index): the device form of the
 * reference's model tests BremTest / MollerTest / MSCAngularDeflectionTest
 * (testAPI.cc:92-199), which drive PerformBrem (SimDPMLike.cc:661),
 * PerformMoller (:699) and PerformMSCAngularDeflection (:731); kind 3 is the
 * Compton final state of KeepTrackingPhoton (:512-570).
 *   kind 0 brem:    out_value = photon energy,      out_cos = photon cos(theta)
 *   kind 1 Moller:  out_value = secondary energy,   out_cos = secondary cos(theta)
 *   kind 2 MSC:     out_value = 1 - cos(theta),     out_cos = cos(theta)
 *   kind 3 Compton: out_value = photon energy,      out_cos = photon cos(theta) */
int dpm_perform(dpm_tables* t, int kind, int precision, int n, double energy, int imat, uint64_t seed,
                double* out_value, double* out_cos);

/* The interpolations SimDPMLike.cc uses, on the device:
 *   which 0 ITr1(E,m) 1 dE/dx(E,m) 2 brem IMFP(E,m) 3 Moller IMFP(E) 4 K1(E)
 *         5 max total IMFP(E) 6 total(E,m) 7 Compton(E,m) 8 pair(E,m) */
int dpm_eval(dpm_tables* t, int which, int precision, int n, const double* energy, const int* imat, double* out);

/* Page-locked host memory for result arrays (no counterpart in the
 * reference, whose results go to a file). Tally copies into such memory run
 * at the link's DMA rate; into pageable memory they are staged by the
 * runtime. Any host pointer is accepted by the entry points above. */
int dpm_host_alloc(size_t bytes, void** out);
int dpm_host_free(void* p);

/* Measurement helper, no counterpart in the reference: the device time of
 * `nops` FP64 atomic additions (no return value, device scope) to cells spread
 * pseudo-randomly over a fresh buffer of `ncells` doubles (< 2^32), i.e. the
 * access pattern of the voxel tally; nops / *seconds is the atomic throughput
 * the tally runs against. With `check` not NULL, *check receives the sum of
 * the buffer (every addition adds 1.0, including those of an untimed warm-up
 * pass of min(nops, 4096 * CUs) additions). */
int dpm_probe_atomics(dpm_tables* t, int64_t ncells, int64_t nops, double* seconds, double* check);

const char* dpm_last_error(void);
int dpm_abi_version(void);

#ifdef __cplusplus
}
#endif
#endif /* DPM_AMD_H */
