// dpm_oracle_cli.cc -- TEST INFRASTRUCTURE ONLY.
//
// Command-line front end of the FP64 oracle, used as the CPU baseline timer
// and for reference-format depth-dose output. Flags:
//   --data DIR        table directory
//   --mode mt|philox  random-number mode (mt = reference mt19937(123) stream)
//   --seed S          seed / Philox key
//   --geom slab|voxel slab = reference Geom (needs --lbox, --gidx);
//                     voxel = homogeneous material-0 cube --nx N (cubic)
//   --lbox L --gidx G --nx N
//   --ptype -1|0      e- or gamma
//   --e0 E --emin A --emax B --field F   (field source when F>0)
//   --hist-begin H --n N
//   --hist FILE       write the reference hist.sim (slab only)
//   --dose FILE       write the voxel energy tally (MeV per voxel, float64,
//                     [nz][ny][nx], raw binary; voxel only)
// Prints one JSON line with the wall time and histories/s.
#include "dpm_oracle.h"

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

int main(int argc, char** argv) {
  std::string data = "./data", mode = "philox", geom = "slab", hist, dose;
  double lbox = 1.0, e0 = 15.0, emin = 0, emax = 0, field = 0, rho = 1.0;
  int gidx = 3, nx = 0, ptype = -1;
  unsigned long long seed = 1, hbegin = 0, n = 1000;
  for (int i = 1; i + 1 < argc; i += 2) {
    const std::string k = argv[i];
    const char* v = argv[i + 1];
    if (k == "--data") data = v;
    else if (k == "--mode") mode = v;
    else if (k == "--seed") seed = std::strtoull(v, nullptr, 10);
    else if (k == "--geom") geom = v;
    else if (k == "--lbox") lbox = std::atof(v);
    else if (k == "--gidx") gidx = std::atoi(v);
    else if (k == "--nx") nx = std::atoi(v);
    else if (k == "--rho") rho = std::atof(v);
    else if (k == "--ptype") ptype = std::atoi(v);
    else if (k == "--e0") e0 = std::atof(v);
    else if (k == "--emin") emin = std::atof(v);
    else if (k == "--emax") emax = std::atof(v);
    else if (k == "--field") field = std::atof(v);
    else if (k == "--hist-begin") hbegin = std::strtoull(v, nullptr, 10);
    else if (k == "--n") n = std::strtoull(v, nullptr, 10);
    else if (k == "--hist") hist = v;
    else if (k == "--dose") dose = v;
    else { std::fprintf(stderr, "unknown flag %s\n", k.c_str()); return 2; }
  }
  if (oracle_load(data.c_str()) != 0) { std::fprintf(stderr, "load: %s\n", oracle_error()); return 1; }
  oracle_cfg cfg;
  std::memset(&cfg, 0, sizeof(cfg));
  cfg.rng_mode = mode == "mt" ? ORACLE_RNG_MT19937 : ORACLE_RNG_PHILOX;
  cfg.arg_rtl = 1;
  cfg.seed = seed;
  cfg.lbox = lbox;
  cfg.geom_index = gidx;
  cfg.majorant_scale = 1.0;
  cfg.src_kind = field > 0 ? ORACLE_SRC_FIELD : ORACLE_SRC_PENCIL;
  cfg.ptype = ptype;
  cfg.e0 = e0; cfg.emin = emin; cfg.emax = emax;
  cfg.fx = field; cfg.fy = field;
  cfg.hist_begin = hbegin;
  cfg.nhist = n;
  std::vector<uint8_t> vm;
  std::vector<float> vr;
  oracle_out out;
  std::memset(&out, 0, sizeof(out));
  std::vector<double> edep, steps;
  if (geom == "voxel") {
    cfg.geom_mode = ORACLE_GEOM_VOXEL;
    cfg.nx = cfg.ny = cfg.nz = nx;
    const size_t nv = (size_t)nx * nx * nx;
    vm.assign(nv, 0); vr.assign(nv, (float)rho);
    cfg.vox_mat = vm.data(); cfg.vox_rho = vr.data();
    edep.assign(nv, 0.0);
  } else {
    cfg.geom_mode = ORACLE_GEOM_SLAB;
    out.hist_cap = 4096;
    edep.assign(out.hist_cap, 0.0); steps.assign(out.hist_cap, 0.0);
    out.steps = steps.data();
  }
  out.edep = edep.data();
  const auto t0 = std::chrono::steady_clock::now();
  if (oracle_run(&cfg, &out) != 0) { std::fprintf(stderr, "run: %s\n", oracle_error()); return 1; }
  const double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  if (!hist.empty() && cfg.geom_mode == ORACLE_GEOM_SLAB) {
    // same layout as Geom::Write (Geom.cc:167-192); density per depth bin from
    // the probe of the loaded material table is not exposed, so the writer
    // prints the raw per-primary energy per bin in MeV/cm (density 1)
    FILE* f = std::fopen(hist.c_str(), "w");
    if (f) {
      double sum = 0;
      for (int i = 0; i < out.hist_size; ++i) sum += edep[i];
      std::fprintf(f, "# === Mean energy deposit in the target: %13.4e  [MeV/event]\n", sum / n);
      std::fprintf(f, "# === Energy deposit as a function of the depth: \n");
      std::fprintf(f, "# === Index       Depth [cm]    Edep [MeV cm]  #Simulation-Steps \n");
      const double norm = 1. / (n * lbox * 0.1);
      for (int i = 0; i < out.hist_size; ++i)
        std::fprintf(f, " %10d    %13.4e    %13.4e    %13.4e\n", i, (i + 0.5) * lbox * 0.1, edep[i] * norm, steps[i] / n);
      std::fclose(f);
    }
  }
  if (!dose.empty() && cfg.geom_mode == ORACLE_GEOM_VOXEL) {
    FILE* f = std::fopen(dose.c_str(), "wb");
    if (!f || std::fwrite(edep.data(), sizeof(double), edep.size(), f) != edep.size()) {
      std::fprintf(stderr, "cannot write %s\n", dose.c_str());
      if (f) std::fclose(f);
      return 1;
    }
    std::fclose(f);
  }
  std::printf("{\"histories\": %llu, \"seconds\": %.6f, \"hist_per_s\": %.3f, \"e_scored\": %.9e, \"e_escaped\": %.9e}\n",
              n, dt, n / dt, out.e_scored, out.e_escaped);
  return 0;
}
