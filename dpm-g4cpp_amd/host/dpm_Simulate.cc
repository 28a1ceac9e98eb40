// dpm_Simulate.cc -- the reference command line (dpm_Simulate.cc:27-158) on
// top of the GPU engine. Same flags and defaults:
//   -p primary particle (e-, gamma)  -e energy [MeV]     -n histories
//   -d data directory                -b voxel size [mm]  -c configuration index
//   -o output file                   -h help
// plus engine flags: -f (FP64 transport), -s seed, -g device.
#include <getopt.h>

#include <cstdlib>
#include <iomanip>
#include <iostream>
#include <string>
#include <vector>

#include "SimDPMLike.hh"

static std::string gPrimaryParticle("e-");
static std::string gInputDataDir("./data");
static std::string gOutputFileName("hist.sim");
static double gPrimaryEnergy = 15.0;
static double gVoxelSize = 1.0;
static int gNumPrimaries = 1.0E+5;
static int gConfigIndex = 0;

static struct option options[] = {
    {"primary-particle    (possible particle names: e-, gamma)            - default: e-", required_argument, 0, 'p'},
    {"primary-energy      (in [MeV] units)                                - default: 15", required_argument, 0, 'e'},
    {"number-of-histories (number of primary events to simulate)          - default: 1.0E+5", required_argument, 0, 'n'},
    {"input-data-dir      (where the pre-generated data are located)      - default: ./data", required_argument, 0, 'd'},
    {"voxel-size          (size of the voxel/box in [mm])                 - default: 1.0", required_argument, 0, 'b'},
    {"configuration-index (one of the pre-defined configuration index)    - default: 0", required_argument, 0, 'c'},
    {"output-filename     (the filename of the result)                    - default: hist.sim", required_argument, 0, 'o'},
    {"fp64                (FP64 transport instead of FP32)                 - default: off", no_argument, 0, 'f'},
    {"seed                (Philox seed)                                   - default: 1", required_argument, 0, 's'},
    {"gpu                 (HIP device index)                              - default: 0", required_argument, 0, 'g'},
    {"help", no_argument, 0, 'h'},
    {0, 0, 0, 0}};

// Configuration::PreDefinedConfigurations (Configuration.hh:69-113): only the
// geometry index is used at run time; the material list is informational
static int GeomIndexOf(int config, std::vector<std::string>& mats) {
  switch (config) {
    case 1: mats = {"G4_WATER", "G4_AIR"}; return 1;
    case 2: mats = {"G4_WATER", "G4_BONE_CORTICAL_ICRP"}; return 2;
    case 3: mats = {"G4_WATER", "G4_Ti", "G4_BONE_COMPACT_ICRU"}; return 3;
    case 4: mats = {"G4_Ti"}; return 0;
    case 5: mats = {"G4_W"}; return 0;
    default: mats = {"G4_WATER"}; return 0;
  }
}

static void Help() {
  std::cout << "\n " << std::setw(120) << std::setfill('=') << "" << std::setfill(' ') << std::endl;
  std::cout << "  The dpm++ simulation phase (MI355X engine)." << std::endl;
  std::cout << "\n  Usage: dpm_Simulate [OPTIONS] \n" << std::endl;
  for (int i = 0; options[i].name != NULL; i++) printf("\t-%c  --%s\n", options[i].val, options[i].name);
  std::cout << "\n " << std::setw(120) << std::setfill('=') << "" << std::setfill(' ') << std::endl;
}

static void GetOpt(int argc, char* argv[]) {
  SimulateOptions& so = GetSimulateOptions();
  while (true) {
    int c, optidx = 0;
    c = getopt_long(argc, argv, "hp:e:n:d:b:c:o:fs:g:", options, &optidx);
    if (c == -1) break;
    switch (c) {
      case 0: c = options[optidx].val; /* fall through */
      case 'p':
        gPrimaryParticle = optarg;
        if (!(gPrimaryParticle == "e-" || gPrimaryParticle == "gamma")) {
          std::cout << " *** Unknown primary particle name -p: " << optarg << std::endl;
          Help();
          exit(-1);
        }
        break;
      case 'o': gOutputFileName = optarg; break;
      case 'e': gPrimaryEnergy = std::stod(optarg); break;
      case 'n': gNumPrimaries = std::stoi(optarg); break;
      case 'd': gInputDataDir = optarg; break;
      case 'b': gVoxelSize = std::stod(optarg); break;
      case 'c': gConfigIndex = std::stoi(optarg); break;
      case 'f': so.precision = DPM_FP64; break;
      case 's': so.seed = std::stoull(optarg); break;
      case 'g': so.device = std::stoi(optarg); break;
      case 'h': Help(); exit(-1); break;
      default: printf(" *** Unknown input argument: %c\n", c); Help(); exit(-1);
    }
  }
  std::cout << "\n === The dpm++ simulation confguration: \n"
            << "\n     - input data directory  : " << gInputDataDir << "\n     - output filename  : " << gOutputFileName
            << "\n     - primary particle      : " << gPrimaryParticle << "\n     - primary energy        : "
            << gPrimaryEnergy << " [MeV]"
            << "\n     - number of histories   : " << gNumPrimaries << "\n     - geometry voxel size   : " << gVoxelSize
            << " [mm]"
            << "\n     - confoguration index   : " << gConfigIndex
            << "\n     - transport precision   : " << (so.precision == DPM_FP64 ? "FP64" : "FP32")
            << "\n     - Philox seed           : " << so.seed << std::endl;
}

int main(int argc, char* argv[]) {
  GetOpt(argc, argv);
  std::vector<std::string> mats;
  const int geomIndex = GeomIndexOf(gConfigIndex, mats);
  std::cout << "\n === Pre-defined configuration " << gConfigIndex << ": geometry index " << geomIndex
            << ", materials:";
  for (size_t i = 0; i < mats.size(); ++i) std::cout << "\n             [" << i << "] = " << mats[i];
  std::cout << "\n" << std::endl;
  std::string err;
  SimElectronData theSimElectronData;
  if (!theSimElectronData.Load(gInputDataDir, err)) { std::cerr << " *** ERROR: " << err << std::endl; return EXIT_FAILURE; }
  SimPhotonData theSimPhotonData;
  if (!theSimPhotonData.Load(gInputDataDir, err)) { std::cerr << " *** ERROR: " << err << std::endl; return EXIT_FAILURE; }
  SimMaterialData theSimMaterialData;
  if (!theSimMaterialData.Load(gInputDataDir, err)) { std::cerr << " *** ERROR: " << err << std::endl; return EXIT_FAILURE; }
  const bool isElectron = (gPrimaryParticle == "e-");
  Simulate(gNumPrimaries, gPrimaryEnergy, isElectron, gVoxelSize, theSimMaterialData, theSimElectronData,
           theSimPhotonData, geomIndex, gOutputFileName);
  return 0;
}
