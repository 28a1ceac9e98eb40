// DoseIO.cc -- voxel-dose / uncertainty and CSR writers.
#include "DoseIO.hh"

#include <cmath>
#include <cstdio>

void VoxelDoseAndSigma(const double* edep, const double* sum2, int nbatch, const float* rho, int64_t nvox,
                       double lbox_mm, uint64_t nhist, std::vector<double>& dose, std::vector<double>& rel_sigma) {
  // 1 MeV = 1.602176634e-13 J; voxel mass [kg] = rho [g/cm3] * V [cm3] * 1e-3
  const double vcm3 = std::pow(lbox_mm * 0.1, 3);
  const double mev_j = 1.602176634e-13;
  dose.assign(nvox, 0.0);
  rel_sigma.assign(nvox, 0.0);
  for (int64_t k = 0; k < nvox; ++k) {
    const double mass = (double)rho[k] * vcm3 * 1e-3;
    dose[k] = edep[k] * mev_j / mass / (double)nhist;
    if (nbatch > 1 && sum2 && edep[k] > 0.0) {
      const double b = (double)nbatch;
      const double mean = edep[k] / b;
      const double var_b = sum2[k] / b - mean * mean;
      const double var_mean = var_b > 0.0 ? var_b / (b - 1.0) : 0.0;
      rel_sigma[k] = std::sqrt(var_mean) / mean;
    }
  }
}

bool WriteVoxelDose(const std::string& fname, int nx, int ny, int nz, double lbox_mm, uint64_t nhist, int nbatch,
                    const std::vector<double>& dose, const std::vector<double>& rel_sigma) {
  FILE* f = std::fopen(fname.c_str(), "wb");
  if (!f) return false;
  std::fprintf(f, "# dpm-amd voxel dose v1\n%d %d %d %.17g %llu %d\n", nx, ny, nz, lbox_mm, (unsigned long long)nhist,
               nbatch);
  const bool ok = std::fwrite(dose.data(), sizeof(double), dose.size(), f) == dose.size() &&
                  std::fwrite(rel_sigma.data(), sizeof(double), rel_sigma.size(), f) == rel_sigma.size();
  return std::fclose(f) == 0 && ok;
}

bool WriteCSR(const std::string& fname, int nbeamlets, int64_t nvox, const int64_t* row_ptr, const int32_t* col,
              const double* val) {
  FILE* f = std::fopen(fname.c_str(), "wb");
  if (!f) return false;
  const int64_t nnz = row_ptr[nbeamlets];
  std::fprintf(f, "# dpm-amd csr v1\n%d %lld %lld\n", nbeamlets, (long long)nvox, (long long)nnz);
  bool ok = std::fwrite(row_ptr, sizeof(int64_t), nbeamlets + 1, f) == (size_t)(nbeamlets + 1);
  ok = ok && std::fwrite(col, sizeof(int32_t), nnz, f) == (size_t)nnz;
  ok = ok && std::fwrite(val, sizeof(double), nnz, f) == (size_t)nnz;
  return std::fclose(f) == 0 && ok;
}
