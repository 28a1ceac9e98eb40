// DoseIO.hh -- output formats beyond the reference's hist.sim.
//
// The reference writes only the 1-D depth dose (Geom::Write, Geom.cc:167-192).
// The voxel-phantom and beamlet forms of the engine write:
//
//  * voxel dose: a text header followed by raw little-endian float64 arrays
//      # dpm-amd voxel dose v1
//      nx ny nz lbox_mm nhist nbatch
//      arrays: dose (Gy per history) [nz][ny][nx]; rel_sigma [nz][ny][nx]
//    dose = E_dep / (rho * V) per history; rel_sigma from the batch estimates
//    s_b (batch tallies): var(mean) = (sum s_b^2 / B - (sum s_b / B)^2) / (B-1)
//  * dose-influence matrix (beamlets): text header, then raw arrays
//      # dpm-amd csr v1
//      nbeamlets nvox nnz
//      row_ptr int64[nbeamlets+1], col int32[nnz], val float64[nnz] (MeV)
#ifndef DPM_AMD_DOSEIO_HH
#define DPM_AMD_DOSEIO_HH

#include <cstdint>
#include <string>
#include <vector>

// per-voxel dose [Gy per history] and relative standard deviation of the
// mean from `nbatch` batch tallies (sum2 = sum of squared batch tallies)
void VoxelDoseAndSigma(const double* edep, const double* sum2, int nbatch, const float* rho, int64_t nvox,
                       double lbox_mm, uint64_t nhist, std::vector<double>& dose, std::vector<double>& rel_sigma);

bool WriteVoxelDose(const std::string& fname, int nx, int ny, int nz, double lbox_mm, uint64_t nhist, int nbatch,
                    const std::vector<double>& dose, const std::vector<double>& rel_sigma);

bool WriteCSR(const std::string& fname, int nbeamlets, int64_t nvox, const int64_t* row_ptr, const int32_t* col,
              const double* val);

#endif
