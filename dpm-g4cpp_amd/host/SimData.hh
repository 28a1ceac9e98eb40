// SimData.hh -- host-side run-time data layer of the MI355X DPM engine.
//
// Mirrors the reference's data classes (Simulation/inc/SimMaterialData.hh,
// SimElectronData.hh, SimPhotonData.hh and the table classes they own): the
// same Load(dataDir) entry points reading the same 14 files with the same
// comment-line counts and fields. Interpolation and sampling do not happen
// here: the loaded node data is flattened into a dpm_table_desc and handed to
// the device engine (include/dpm_amd.h), which derives the interpolation
// coefficients and runs every Get*/Sample* on the GPU.
#ifndef DPM_AMD_SIMDATA_HH
#define DPM_AMD_SIMDATA_HH

#include <string>
#include <vector>

#include "dpm_amd.h"

// node data of one interpolated quantity
struct SimNodes {
  std::vector<double> x, y;
};

// one family of linear-pdf alias tables (SimLinAliasData sets)
struct SimAliasSet {
  int nmat = 0, ne = 0, ns = 0;
  std::vector<double> energy, x, y, w;
  std::vector<int> alias;
};

// SimMaterialData (SimMaterialData.hh:13-47, SimMaterialData.cc:6-53)
class SimMaterialData {
 public:
  // returns false (and fills `err`) where the reference calls exit(EXIT_FAILURE)
  bool Load(const std::string& dataDir, std::string& err, int verbose = 0);

  double fElectronCut = 0, fGammaCut = 0;
  double fMSCStepParSLow = 0, fMSCStepParSHigh = 0, fMSCStepParEcross = 0;
  int fNumMaterial = 0;
  std::vector<std::string> fMaterialName;
  std::vector<double> fMaterialDensity;
  std::vector<double> fMollerIMFPScaling;
};

// SimElectronData (SimElectronData.hh:31-103, SimElectronData.cc:44-115)
class SimElectronData {
 public:
  bool Load(const std::string& dataDir, std::string& err, int verbose = 0);

  std::vector<SimNodes> fITr1MFP;     // el_itr1mfp.dat, column 3, per material
  SimNodes fMaxScatStrength;          // el_scatStrength.dat, material 0
  SimNodes fIMFPMoller;               // imfp_moller.dat, material 0
  std::vector<SimNodes> fIMFPBrem;    // imfp_brem.dat, per material
  std::vector<SimNodes> fDEDX;        // eloss_rdedx.dat, per material
  SimAliasSet fMollerTables;          // ioni_MollerDtrData.dat
  SimAliasSet fSBTables;              // brem_SBDtrData.dat
  // GS tables (el_GSDtrData.dat, material 0)
  int fGSNumEnergies = 0, fGSNumCum = 0;
  std::vector<double> fGSEnergy, fGSTransf, fGSU, fGSA, fGSB;
};

// SimPhotonData (SimPhotonData.hh:22-71, SimPhotonData.cc:28-78)
class SimPhotonData {
 public:
  bool Load(const std::string& dataDir, std::string& err, int verbose = 0);

  SimNodes fIMFPTotalMax;             // imfp_globalMax.dat
  std::vector<SimNodes> fIMFPTotal;   // imfp_total.dat
  std::vector<SimNodes> fIMFPCompton; // imfp_compton.dat
  std::vector<SimNodes> fIMFPPairProd;// imfp_pairp.dat
  SimAliasSet fKNTables;              // compton_KNDtrData.dat
};

// Flattens the three loaded objects into the engine's plain-C descriptor.
// The descriptor points into `keep`, which must outlive its use.
struct SimTableView {
  std::vector<dpm_spline_desc> itr1, dedx, brem;
  std::vector<dpm_linear_desc> ptot, pcomp, ppair;
  dpm_table_desc desc;
};
void FlattenTables(const SimMaterialData& mat, const SimElectronData& el, const SimPhotonData& ph,
                   SimTableView& keep);

#endif
