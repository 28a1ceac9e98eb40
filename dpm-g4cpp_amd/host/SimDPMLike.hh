// SimDPMLike.hh -- drop-in for the reference's transport entry point.
//
// Simulate() keeps the exact signature of the reference
// (Simulation/inc/SimDPMLike.hh:46) so that a dpm_Simulate main written for
// the reference compiles unchanged against it; the transport runs on the GPU
// through the C ABI of include/dpm_amd.h.
#ifndef DPM_AMD_SIMDPMLIKE_HH
#define DPM_AMD_SIMDPMLIKE_HH

#include <cstdint>
#include <string>

#include "SimData.hh"

// Engine settings the reference signature has no room for (defaults: FP32
// transport, Philox seed 1, device 0).
struct SimulateOptions {
  int precision = DPM_FP32;
  uint64_t seed = 1;
  int device = 0;
};
SimulateOptions& GetSimulateOptions();

// Simulation/inc/SimDPMLike.hh:46 -- transports `nprimary` histories of a
// pencil beam (e- when `iselectron`, else gamma) of energy `eprimary` [MeV]
// through the pre-defined slab geometry `geomIndex` of `lbox` [mm] voxels and
// writes the depth dose to `filename` in the reference's hist.sim format
// (Geom::Write, Geom.cc:167-192). Errors end the process with EXIT_FAILURE,
// as the reference's loaders do.
void Simulate(int nprimary, double eprimary, bool iselectron, double lbox, SimMaterialData& matData,
              SimElectronData& elData, SimPhotonData& phData, int geomIndex, std::string filename);

// Geom::Write (Geom.cc:167-192): `edep`/`steps` are the depth tallies of
// `size` rows, `density` of each row's material (the slab layer rule).
bool WriteHistSim(const std::string& fname, const double* edep, const double* steps, int size, int nprimaries,
                  double lbox, const double* density);

#endif
