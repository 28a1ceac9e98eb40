/*
 * dpm_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU FP64 restatement of the reference DPM transport (bl406/dpm-g4cpp,
 * Simulation/src/SimDPMLike.cc, Geom.cc and the Sim* table classes), used as
 * the parity oracle for the GPU engine. Only tests/, __graft_entry__.smoke()
 * and bench.py's CPU-baseline leg may load it; the product never does.
 *
 * Two random-number modes:
 *  - ORACLE_RNG_MT19937: std::mt19937 + uniform_real_distribution<double>
 *    exactly as Simulation/src/Random.cc:3-4, with the uniform-consumption
 *    order g++ emits for the reference's multi-argument sampler calls; this
 *    mode reproduces the reference's hist.sim bit for bit.
 *  - ORACLE_RNG_PHILOX: the GPU engine's counter-based Philox4x32-10 layout
 *    (one stream per particle keyed by seed, history and particle serial,
 *    24-bit uniforms), so that GPU and oracle follow the same histories.
 */
#ifndef DPM_ORACLE_H
#define DPM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORACLE_RNG_MT19937 = 0, ORACLE_RNG_PHILOX = 1 };
enum { ORACLE_GEOM_SLAB = 0, ORACLE_GEOM_VOXEL = 1 };
enum { ORACLE_SRC_PENCIL = 0, ORACLE_SRC_FIELD = 1, ORACLE_SRC_GANTRY = 2, ORACLE_SRC_PHSP = 3 };
enum { ORACLE_SPEC_FLAT = 0, ORACLE_SPEC_TRIANGULAR = 1, ORACLE_SPEC_TABLE = 2 };

typedef struct oracle_cfg {
  int      rng_mode;      /* ORACLE_RNG_* */
  int      arg_rtl;       /* mt19937 mode: 1 = multi-argument uniforms drawn right-to-left */
  uint64_t seed;          /* mt19937 seed (reference: 123) or Philox key */
  int      geom_mode;     /* ORACLE_GEOM_* */
  double   lbox;          /* voxel size [mm] */
  int      geom_index;    /* slab: reference pre-defined geometry index 0..3 */
  int      nx, ny, nz;    /* voxel phantom */
  const uint8_t* vox_mat; /* [nx*ny*nz] material index */
  const float*   vox_rho; /* [nx*ny*nz] density [g/cm3] */
  double   majorant_scale;/* Woodcock majorant scale (>=1) */
  int      src_kind;      /* ORACLE_SRC_* */
  int      ptype;         /* -1 e-, 0 gamma, +1 e+ */
  double   e0, emin, emax;/* mono energy, or uniform [emin,emax] when emax>emin */
  double   fx, fy;        /* field size [mm] (field source) */
  double   x0, y0;        /* pencil entry point [mm] */
  uint64_t hist_begin;    /* first global history id (Philox) */
  uint64_t nhist;
  int      nbatch;        /* >0: per-batch sum / sum^2 of the voxel tally */
  /* gantry field (voxel phantom): beam direction (sin a, 0, cos a) through the
     phantom centre, field on the plane through the centre, primaries start
     where their ray enters the phantom */
  double   angle_deg;
  int      spectrum;      /* ORACLE_SPEC_*: with emax > emin, flat or triangular(emin, emode, emax);
                             table: histogram bins [edges[i], edges[i+1]) of relative weight w[i] */
  double   emode;
  int      spec_n;
  const double* spec_edges;   /* [spec_n + 1] */
  const double* spec_weights; /* [spec_n] */
  /* ORACLE_SRC_PHSP: [nhist][8] records (particle -1/0/+1, E, x, y, z, u, v, w);
     history ih is record ih (id hist_begin + ih), no source draws; in a voxel
     phantom a record outside moves along its direction to where it enters */
  const double* phsp;
  /* gantry-field collimation: sad > 0 a point source at that distance [mm]
     upstream of the phantom centre on the beam axis, the particle travelling
     through its field point; jaws != 0 the aperture [jaw_x1, jaw_x2) x
     [jaw_y1, jaw_y2) of the field point relative to the field centre, outside
     which the particle goes on with probability exp(-n jaw_tau / cos theta)
     (n jaw pairs missed, theta its angle to the axis; one uniform drawn after
     the energy) and is absorbed otherwise */
  double sad;
  int jaws;
  double jaw_x1, jaw_x2, jaw_y1, jaw_y2;
  double jaw_tau;
} oracle_cfg;

typedef struct oracle_out {
  double*  edep;          /* slab: depth histogram [hist_cap]; voxel: [nvox] */
  double*  steps;         /* slab: step-count histogram [hist_cap] (may be NULL) */
  double*  batch_sum;     /* voxel, nbatch>0: sum over batches of the batch tally [nvox] */
  double*  batch_sum2;    /* voxel, nbatch>0: sum of squares [nvox] */
  double*  hist_edep;     /* optional [nhist]: energy scored per history */
  int      hist_cap;      /* slab histogram capacity (input) */
  int      hist_size;     /* slab histogram size after the reference growth rule */
  int      max_iz;        /* largest depth index scored */
  double   e_escaped;     /* kinetic energy that left the geometry or entered vacuum */
  double   e_scored;      /* total scored energy (all indices >= 0) */
  uint64_t n_pair, n_annih;
  /* byte-model counters: voxel visits, deposits, SB, Moller, KN, GS samples,
     secondaries created, electron steps, photon steps */
  uint64_t cnt[9];
  double   e_primary;     /* sum of the primaries' kinetic energies */
  double*  prim;          /* optional [nhist][7]: each primary as generated, x y z u v w E */
  uint64_t n_jaw;         /* primaries absorbed in the jaws (collimated gantry field) */
} oracle_out;

int  oracle_load(const char* dir);          /* 0 on success */
const char* oracle_error(void);
int  oracle_num_materials(void);
int  oracle_run(const oracle_cfg* cfg, oracle_out* out);

/* sampler entry points taking explicit uniforms (reference argument order) */
double oracle_sample_sb(double e, int imat, double r1, double r2, double r3, int* ebin, int* abin);
double oracle_sample_moller(double e, double r1, double r2, double r3, int* ebin, int* abin);
double oracle_sample_kn(double e, double r1, double r2, double r3, int* ebin, int* abin);
double oracle_sample_gs(double e, double r1, double r2, int* ebin, int* cbin);
/* the same GS sample as 1 - cos(theta) (no cancellation for small angles) */
double oracle_sample_gs_omu(double e, double r1, double r2, int* ebin, int* cbin);
int oracle_sample_many(int kind, int n, double e, int imat, const double* r1, const double* r2, const double* r3,
                       double* out);
/* table interpolations as used by the transport (with the reference clamps) */
double oracle_eval(int which, double e, int imat);
/* test hook: plant a known mistake, 0 = off: 1 drop the Compton electron,
   2 batch rows shifted by one history, 3 field source half as wide,
   4 Woodcock interaction probabilities from the unscaled majorant,
   5 jaw transmission without the obliquity 1/cos theta */
void   oracle_set_fault(int f);
/* Philox4x32-10 block for tests */
void   oracle_philox(uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* geometry probe: distance to the next voxel face from r along v (r is
 * pushed by the tolerance rule), box of r; -1 when r is outside. Slab: the
 * reference Geom with boxes of lbox; voxel: a centred nx*ny*nz phantom */
double oracle_distance(int geom_mode, double lbox, int nx, int ny, int nz, double* r, const double* v, int* box);
/* table spectrum: energies for uniforms u[m] from n bins (edges[n+1], weights[n]) */
void   oracle_table_energy(int n, const double* edges, const double* weights, int m, const double* u, double* e);
/* RotateToLabFrame: uvw (in the frame of axis u) -> lab frame, in place */
void   oracle_rotate(double* uvw, double u1, double u2, double u3);

#ifdef __cplusplus
}
#endif
#endif
