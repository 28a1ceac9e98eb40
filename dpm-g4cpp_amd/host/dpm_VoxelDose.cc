// dpm_VoxelDose.cc -- voxel-phantom runs on the GPU engine, writing the 3-D
// dose with its batch uncertainty, or the dose-influence matrix of a beamlet
// set. Companion of dpm_Simulate (which keeps the reference's slab geometry
// and hist.sim output, Geom.cc:167-192); the file formats are those of
// DoseIO.hh.
//
//   -d data directory           -x -y -z voxels per axis   -l voxel size [mm]
//   -m materials (raw u8, z-major [z][y][x]; default: all material 0)
//   -r densities (raw f32, same order; default: nominal density of each voxel's material)
//   -p particle (e-, gamma, e+) -e energy [MeV]  or  -E emin,emax (uniform spectrum)
//   -F field side [mm] (0: pencil)  -X -Y entry point [mm]
//   -A gantry angle [deg] (field through the phantom centre)  -T mode of a triangular spectrum (with -E)
//   -S spectrum file (lines "E_low E_high weight" [MeV], contiguous bins)
//   -n histories  -k batches  -s seed  -f FP64 transport  -g device  -o output
//   -B beamlet list (lines "x y particle energy histories"; writes CSR)  -t CSR threshold
//   -P phase-space file (lines "particle E x y z u v w": one history per record)
#include <getopt.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "DoseIO.hh"
#include "SimData.hh"
#include "dpm_amd.h"

namespace {

struct Args {
  std::string data = "./data", mat_file, rho_file, beamlets, out = "dose.bin", particle = "gamma", spectrum, phsp;
  int nx = 0, ny = 0, nz = 0, nbatch = 10, device = 0;
  double lbox = 2.0, energy = 10.0, emin = 0.0, emax = 0.0, field = 0.0, x0 = 0.0, y0 = 0.0, threshold = 1e-3;
  double angle = 0.0, emode = 0.0;
  bool gantry = false, triangular = false;
  double sad = 0.0, jaw[4] = {0, 0, 0, 0}, jaw_tau = 0.0;   // collimation of the gantry field
  bool jaws = false;
  uint64_t nhist = 100000, seed = 1;
  bool fp64 = false;
};

void Help() {
  std::cout << "\n  dpm_VoxelDose: voxel-phantom dose (and batch uncertainty) or beamlet dose-influence matrix\n\n"
               "  -d dir     table data directory                     (default ./data)\n"
               "  -x -y -z   voxels per axis                          (required)\n"
               "  -l mm      voxel size                               (default 2)\n"
               "  -m file    materials, raw u8 [z][y][x]              (default: material 0)\n"
               "  -r file    densities, raw f32 [z][y][x] in g/cm3    (default: nominal)\n"
               "  -p name    primary: e-, gamma, e+                   (default gamma)\n"
               "  -e MeV     energy                                   (default 10)\n"
               "  -E lo,hi   uniform energy spectrum instead of -e\n"
               "  -F mm      square field side, 0 = pencil            (default 0)\n"
               "  -X -Y mm   entry point                              (default 0 0)\n"
               "  -A deg     gantry angle: beam (sin a, 0, cos a) through the phantom centre\n"
               "  -D mm      with -A: point source at this distance upstream of the phantom centre\n"
               "             (divergent field; -F is then the field at the centre plane)\n"
               "  -J x1,x2,y1,y2  with -A: jaw aperture on the centre plane, relative to the field\n"
               "             centre (mm); outside it particles cross the jaws unscattered with\n"
               "             probability exp(-n tau / cos theta) or are absorbed\n"
               "  -K tau     jaw optical thickness on the axis for -J     (default 0)\n"
               "  -T MeV     mode of a triangular spectrum over -E lo,hi\n"
               "  -S file    spectrum file: lines \"E_low E_high weight\" (MeV), contiguous bins\n"
               "  -n N       histories                                (default 1e5)\n"
               "  -k K       history batches for the uncertainty      (default 10)\n"
               "  -s seed    Philox seed                              (default 1)\n"
               "  -f         FP64 transport\n"
               "  -g dev     HIP device                               (default 0)\n"
               "  -B file    beamlet list: lines 'x y particle energy histories' -> CSR output\n"
               "  -t rel     CSR threshold relative to each beamlet's maximum (default 1e-3)\n"
               "  -P file    phase-space file: lines 'particle E x y z u v w' (MeV, mm, direction\n"
               "             cosines; '#' comments), one history per record, instead of -p -e -F ...\n"
               "  -o file    output                                   (default dose.bin)\n"
               "  -h         this help\n\n";
}

int ParticleCode(const std::string& p) {
  if (p == "e-") return DPM_ELECTRON;
  if (p == "gamma") return DPM_PHOTON;
  if (p == "e+") return DPM_POSITRON;
  std::cerr << " *** unknown particle " << p << std::endl;
  std::exit(EXIT_FAILURE);
}

template <typename T>
std::vector<T> ReadRaw(const std::string& f, size_t n) {
  std::vector<T> v(n);
  std::ifstream in(f, std::ios::binary);
  if (!in || !in.read(reinterpret_cast<char*>(v.data()), (std::streamsize)(n * sizeof(T)))) {
    std::cerr << " *** cannot read " << n << " values from " << f << std::endl;
    std::exit(EXIT_FAILURE);
  }
  return v;
}

void Check(int rc, const char* what) {
  if (rc != DPM_OK) {
    std::cerr << " *** " << what << ": " << dpm_last_error() << std::endl;
    std::exit(EXIT_FAILURE);
  }
}

}  // namespace

// phase-space file: one record per line, "particle E x y z u v w" (particle
// e-, gamma or e+; MeV; mm in phantom coordinates; direction cosines); '#'
// starts a comment line; any other line that is not exactly these eight
// fields is an error
bool ReadPhaseSpace(const std::string& path, std::vector<dpm_phsp_particle>& rec) {
  std::ifstream f(path);
  if (!f) {
    std::cerr << " *** cannot read the phase-space file " << path << std::endl;
    return false;
  }
  std::string line;
  int lineno = 0;
  while (std::getline(f, line)) {
    ++lineno;
    const size_t first = line.find_first_not_of(" \t\r");
    if (first == std::string::npos || line[first] == '#') continue;
    std::istringstream is(line);
    std::string p, rest;
    dpm_phsp_particle r{};
    if (!(is >> p >> r.energy >> r.x >> r.y >> r.z >> r.u >> r.v >> r.w) || (is >> rest) ||
        (p != "e-" && p != "gamma" && p != "e+")) {
      std::cerr << " *** " << path << ":" << lineno << ": expected 'particle E x y z u v w', got '" << line << "'"
                << std::endl;
      return false;
    }
    r.particle = p == "e-" ? DPM_ELECTRON : (p == "gamma" ? DPM_PHOTON : DPM_POSITRON);
    rec.push_back(r);
  }
  if (rec.empty()) {
    std::cerr << " *** no records in the phase-space file " << path << std::endl;
    return false;
  }
  return true;
}

// spectrum file: one bin per line, "E_low E_high weight" in MeV, each bin
// starting where the previous ends; '#' starts a comment line
static bool ReadSpectrum(const std::string& path, std::vector<double>& edges, std::vector<double>& weights) {
  std::ifstream f(path);
  if (!f) {
    std::cerr << " *** cannot read the spectrum file " << path << std::endl;
    return false;
  }
  std::string line;
  int lineno = 0;
  while (std::getline(f, line)) {
    ++lineno;
    const size_t first = line.find_first_not_of(" \t\r");
    if (first == std::string::npos || line[first] == '#') continue;   // blank or comment
    // every other line is exactly three numbers: a typo or a truncated line
    // is an error, not a silently dropped bin
    std::istringstream is(line);
    double lo, hi, w;
    std::string rest;
    if (!(is >> lo >> hi >> w) || (is >> rest)) {
      std::cerr << " *** " << path << ":" << lineno << ": expected 'E_low E_high weight', got '" << line << "'"
                << std::endl;
      return false;
    }
    if (edges.empty()) edges.push_back(lo);
    else if (lo != edges.back()) {
      std::cerr << " *** spectrum bins must be contiguous (" << path << ": " << lo << ")" << std::endl;
      return false;
    }
    edges.push_back(hi);
    weights.push_back(w);
  }
  if (weights.empty()) {
    std::cerr << " *** no bins in the spectrum file " << path << std::endl;
    return false;
  }
  return true;
}

int main(int argc, char* argv[]) {
  Args a;
  int c;
  while ((c = getopt(argc, argv, "hd:x:y:z:l:m:r:p:e:E:F:X:Y:A:T:S:n:k:s:fg:B:t:o:P:D:J:K:")) != -1) {
    switch (c) {
      case 'd': a.data = optarg; break;
      case 'x': a.nx = std::atoi(optarg); break;
      case 'y': a.ny = std::atoi(optarg); break;
      case 'z': a.nz = std::atoi(optarg); break;
      case 'l': a.lbox = std::atof(optarg); break;
      case 'm': a.mat_file = optarg; break;
      case 'r': a.rho_file = optarg; break;
      case 'p': a.particle = optarg; break;
      case 'e': a.energy = std::atof(optarg); break;
      case 'E': {
        if (std::sscanf(optarg, "%lf,%lf", &a.emin, &a.emax) != 2 || !(a.emax > a.emin)) {
          std::cerr << " *** -E needs lo,hi with hi > lo" << std::endl;
          return EXIT_FAILURE;
        }
        break;
      }
      case 'F': a.field = std::atof(optarg); break;
      case 'X': a.x0 = std::atof(optarg); break;
      case 'Y': a.y0 = std::atof(optarg); break;
      case 'A': a.angle = std::atof(optarg); a.gantry = true; break;
      case 'T': a.emode = std::atof(optarg); a.triangular = true; break;
      case 'D': a.sad = std::atof(optarg); break;
      case 'J': {
        if (std::sscanf(optarg, "%lf,%lf,%lf,%lf", &a.jaw[0], &a.jaw[1], &a.jaw[2], &a.jaw[3]) != 4) {
          std::cerr << " *** -J needs x1,x2,y1,y2" << std::endl;
          return EXIT_FAILURE;
        }
        a.jaws = true;
        break;
      }
      case 'K': a.jaw_tau = std::atof(optarg); break;
      case 'S': a.spectrum = optarg; break;
      case 'P': a.phsp = optarg; break;
      case 'n': a.nhist = (uint64_t)std::atof(optarg); break;
      case 'k': a.nbatch = std::atoi(optarg); break;
      case 's': a.seed = (uint64_t)std::atoll(optarg); break;
      case 'f': a.fp64 = true; break;
      case 'g': a.device = std::atoi(optarg); break;
      case 'B': a.beamlets = optarg; break;
      case 't': a.threshold = std::atof(optarg); break;
      case 'o': a.out = optarg; break;
      case 'h': Help(); return EXIT_SUCCESS;
      default: Help(); return EXIT_FAILURE;
    }
  }
  if (a.nx <= 0 || a.ny <= 0 || a.nz <= 0) {
    std::cerr << " *** -x -y -z are required" << std::endl;
    Help();
    return EXIT_FAILURE;
  }
  const size_t nvox = (size_t)a.nx * a.ny * a.nz;

  dpm_tables* t = nullptr;
  // input files are checked before the device is touched
  std::vector<double> sp_edges, sp_weights;
  if (!a.spectrum.empty() && !ReadSpectrum(a.spectrum, sp_edges, sp_weights)) return EXIT_FAILURE;
  std::vector<dpm_phsp_particle> phsp;
  if (!a.phsp.empty() && !ReadPhaseSpace(a.phsp, phsp)) return EXIT_FAILURE;
  Check(dpm_tables_load(a.data.c_str(), a.device, &t), "loading tables");
  std::vector<uint8_t> mat = a.mat_file.empty() ? std::vector<uint8_t>(nvox, 0) : ReadRaw<uint8_t>(a.mat_file, nvox);
  std::vector<float> rho;
  if (!a.rho_file.empty()) {
    rho = ReadRaw<float>(a.rho_file, nvox);
  } else {
    // nominal density of each voxel's material (SimMaterialData, mat.dat)
    SimMaterialData md;
    std::string err;
    if (!md.Load(a.data, err)) {
      std::cerr << " *** " << err << std::endl;
      return EXIT_FAILURE;
    }
    rho.resize(nvox);
    for (size_t i = 0; i < nvox; ++i) {
      if (mat[i] >= md.fMaterialDensity.size()) {
        std::cerr << " *** material index " << (int)mat[i] << " not in the tables" << std::endl;
        return EXIT_FAILURE;
      }
      rho[i] = (float)md.fMaterialDensity[mat[i]];
    }
  }
  dpm_phantom* ph = nullptr;
  Check(dpm_phantom_create_voxel(t, a.nx, a.ny, a.nz, a.lbox, mat.data(), rho.data(), &ph), "creating the phantom");

  dpm_run_options opt = {};
  opt.seed = a.seed;
  opt.precision = a.fp64 ? DPM_FP64 : DPM_FP32;

  if (!a.beamlets.empty()) {
    std::ifstream in(a.beamlets);
    std::vector<dpm_beamlet> bl;
    std::string line;
    while (std::getline(in, line)) {
      if (line.empty() || line[0] == '#') continue;
      std::istringstream ss(line);
      dpm_beamlet b{};
      std::string p;
      double n = 0;
      if (!(ss >> b.x0 >> b.y0 >> p >> b.energy >> n)) {
        std::cerr << " *** bad beamlet line: " << line << std::endl;
        return EXIT_FAILURE;
      }
      b.particle = ParticleCode(p);
      b.nhist = (uint64_t)n;
      bl.push_back(b);
    }
    std::vector<int64_t> row_ptr(bl.size() + 1);
    dpm_stats st;
    opt.graph = 1;
    Check(dpm_simulate_beamlets(t, ph, bl.data(), (int)bl.size(), &opt, a.threshold, row_ptr.data(), &st),
          "beamlet run");
    const int64_t nnz = row_ptr.back();
    std::vector<int32_t> col(nnz);
    std::vector<double> val(nnz);
    Check(dpm_beamlets_fetch(t, col.data(), val.data()), "fetching the CSR");
    if (!WriteCSR(a.out, (int)bl.size(), (int64_t)nvox, row_ptr.data(), col.data(), val.data())) {
      std::cerr << " *** cannot write " << a.out << std::endl;
      return EXIT_FAILURE;
    }
    std::cout << " beamlets " << bl.size() << ", nonzeros " << nnz << ", device time " << st.seconds << " s -> "
              << a.out << std::endl;
  } else {
    dpm_source src = {};
    src.kind = a.gantry ? DPM_SRC_GANTRY : (a.field > 0 ? DPM_SRC_FIELD : DPM_SRC_PENCIL);
    src.angle_deg = a.angle;
    src.spectrum = a.triangular ? DPM_SPEC_TRIANGULAR : DPM_SPEC_FLAT;
    src.emode = a.emode;
    if (!a.spectrum.empty()) {
      src.spectrum = DPM_SPEC_TABLE;
      src.spec_n = (int)sp_weights.size();
      src.spec_edges = sp_edges.data();
      src.spec_weights = sp_weights.data();
    }
    src.particle = ParticleCode(a.particle);
    src.energy = a.energy;
    src.emin = a.emin;
    src.emax = a.emax;
    src.x0 = a.x0;
    src.y0 = a.y0;
    src.field_x = a.field;
    src.field_y = a.field;
    if ((a.sad > 0.0 || a.jaws) && !a.gantry) {
      std::cerr << " *** -D and -J collimate a gantry field: give -A" << std::endl;
      return EXIT_FAILURE;
    }
    src.sad = a.sad;
    src.jaws = a.jaws ? 1 : 0;
    src.jaw_x1 = a.jaw[0]; src.jaw_x2 = a.jaw[1]; src.jaw_y1 = a.jaw[2]; src.jaw_y2 = a.jaw[3];
    src.jaw_tau = a.jaw_tau;
    opt.nbatch = a.nbatch;
    std::vector<double> edep(nvox), sum2(nvox);
    dpm_stats st;
    if (!phsp.empty()) {
      a.nhist = phsp.size();   // one history per record
      Check(dpm_simulate_phsp(t, ph, phsp.data(), phsp.size(), &opt, edep.data(), sum2.data(), &st), "run");
    } else {
      Check(dpm_simulate(t, ph, &src, a.nhist, &opt, edep.data(), sum2.data(), &st), "run");
    }
    std::vector<double> dose, rel;
    VoxelDoseAndSigma(edep.data(), a.nbatch > 1 ? sum2.data() : nullptr, a.nbatch, rho.data(), (int64_t)nvox, a.lbox,
                      a.nhist, dose, rel);
    if (!WriteVoxelDose(a.out, a.nx, a.ny, a.nz, a.lbox, a.nhist, a.nbatch, dose, rel)) {
      std::cerr << " *** cannot write " << a.out << std::endl;
      return EXIT_FAILURE;
    }
    std::cout << " histories " << a.nhist << ", scored " << st.e_scored << " MeV, device time " << st.seconds
              << " s -> " << a.out << std::endl;
  }
  dpm_phantom_destroy(ph);
  dpm_tables_destroy(t);
  return EXIT_SUCCESS;
}
