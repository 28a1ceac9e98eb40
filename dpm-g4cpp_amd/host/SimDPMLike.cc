// SimDPMLike.cc -- GPU-backed Simulate() with the reference's signature.
#include "SimDPMLike.hh"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <vector>

SimulateOptions& GetSimulateOptions() {
  static SimulateOptions opts;
  return opts;
}

namespace {

// material index of depth row iz in the pre-defined slab (Geom.cc:133-164)
int SlabMaterial(int iz, double lbox, int geomIndex) {
  const double inv = 1.0 / lbox;
  if (iz < 0) return -1;
  switch (geomIndex) {
    case 1: return iz < 10.0 * inv ? 0 : (iz < 30.0 * inv ? 1 : 0);
    case 2: return iz < 20.0 * inv ? 0 : (iz < 40.0 * inv ? 1 : 0);
    case 3: return iz < 10.0 * inv ? 0 : (iz < 15.0 * inv ? 1 : (iz < 25.0 * inv ? 2 : 0));
    default: return 0;
  }
}

[[noreturn]] void Fail(const char* what) {
  std::cerr << " *** ERROR in Simulate(): " << what << ": " << dpm_last_error() << std::endl;
  std::exit(EXIT_FAILURE);
}

}  // namespace

bool WriteHistSim(const std::string& fname, const double* edep, const double* steps, int size, int nprimaries,
                  double lbox, const double* density) {
  FILE* f = std::fopen(fname.c_str(), "w");
  if (!f) {
    std::cerr << " *** ERROR in Geom::Write(): " << " file " << fname << " could not be created! " << std::endl;
    return false;
  }
  const double toCm = 0.1;
  const double norm = 1. / (nprimaries * lbox * toCm);
  double sumEdep = 0.0;
  for (int i = 0; i < size; ++i) sumEdep += edep[i];
  std::fprintf(f, "# === Mean energy deposit in the target: %13.4e  [MeV/event]\n", sumEdep / nprimaries);
  std::fprintf(f, "# === Energy deposit as a function of the depth: \n");
  std::fprintf(f, "# === Index       Depth [cm]    Edep [MeV cm2/g]  #Simulation-Steps \n");
  for (int i = 0; i < size; ++i)
    std::fprintf(f, " %10d    %13.4e    %13.4e    %13.4e\n", i, (i + 0.5) * lbox * toCm, edep[i] * norm / density[i],
                 steps[i] / nprimaries);
  std::fclose(f);
  std::cout << " === Energy deposit histogram is written to the file:  " << fname << "\n" << std::endl;
  return true;
}

void Simulate(int nprimary, double eprimary, bool iselectron, double lbox, SimMaterialData& matData,
              SimElectronData& elData, SimPhotonData& phData, int geomIndex, std::string filename) {
  const SimulateOptions& o = GetSimulateOptions();
  if (nprimary <= 0) {
    std::cerr << " *** ERROR in Simulate(): the number of primaries must be positive" << std::endl;
    std::exit(EXIT_FAILURE);
  }
  SimTableView view;
  FlattenTables(matData, elData, phData, view);
  dpm_tables* tables = nullptr;
  if (dpm_tables_create(&view.desc, o.device, &tables) != DPM_OK) Fail("table upload");
  dpm_phantom* geom = nullptr;
  if (dpm_phantom_create_slab(tables, lbox, geomIndex, &geom) != DPM_OK) Fail("geometry");
  dpm_source src = {};
  src.kind = DPM_SRC_PENCIL;
  src.particle = iselectron ? DPM_ELECTRON : DPM_PHOTON;
  src.energy = eprimary;
  dpm_run_options opt = {};
  opt.seed = o.seed;
  opt.precision = o.precision;
  const int64_t nbins = dpm_phantom_num_bins(geom);
  std::vector<double> edep(nbins, 0.0), steps(nbins, 0.0);
  dpm_stats st = {};
  std::cout << "\n === Start simulation of N = " << nprimary << " events === \n" << std::endl;
  if (dpm_simulate(tables, geom, &src, (uint64_t)nprimary, &opt, edep.data(), steps.data(), &st) != DPM_OK)
    Fail("transport");
  std::cout << "\n === End simulation of N = " << nprimary << " events === \n" << std::endl;
  std::cout << "     device time " << st.seconds << " s, " << nprimary / std::max(st.seconds, 1e-12)
            << " histories/s" << std::endl;
  const int size = (int)std::min<int64_t>(st.hist_size, nbins);
  std::vector<double> dens(size);
  for (int i = 0; i < size; ++i) {
    const int m = SlabMaterial(i, lbox, geomIndex);
    dens[i] = m > -1 ? matData.fMaterialDensity[m] : 1.0E-40;
  }
  WriteHistSim(filename, edep.data(), steps.data(), size, nprimary, lbox, dens.data());
  dpm_phantom_destroy(geom);
  dpm_tables_destroy(tables);
}
