// SimData.hh -- host-side run-time data layer of the MI355X DPM engine.
//
// Mirrors the reference's data classes (Simulation/inc/SimMaterialData.hh,
// SimElectronData.hh, SimPhotonData.hh and the table classes they own,
// SimTables.hh): the same Load(dataDir, verbose) entry points reading the same
// 14 files with the same comment-line counts and fields, the same getters.
// Load(dataDir[, verbose]) ends the process with EXIT_FAILURE on a missing or
// malformed file, as the reference's loaders do; Load(dataDir, err[, verbose])
// returns false with the reason instead (the form the library uses). The
// loaded node data is flattened into a dpm_table_desc for the device engine
// (include/dpm_amd.h), which runs the transport's Get*/Sample* on the GPU.
#ifndef DPM_AMD_SIMDATA_HH
#define DPM_AMD_SIMDATA_HH

#include <string>
#include <vector>

#include "SimTables.hh"
#include "dpm_amd.h"

// SimMaterialData (SimMaterialData.hh:13-47, SimMaterialData.cc:6-53)
class SimMaterialData {
 public:
  void Load(const std::string& dataDir, int verbose = 0);
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
  void Load(const std::string& dataDir, int verbose = 0);
  bool Load(const std::string& dataDir, std::string& err, int verbose = 0);

  // SimElectronData.hh:42-49
  SimITr1MFPElastic* GetITr1MFPElastic() const { return &fITr1MFPElasticTab; }
  SimMaxScatStrength* GetMaxScatStrength() const { return &fMaxScatStrengthTab; }
  SimIMFPMoller* GetIMFPMoller() const { return &fIMFPMollerTab; }
  SimIMFPBrem* GetIMFPBrem() const { return &fIMFPBremTab; }
  SimStoppingPower* GetDEDX() const { return &fDEDXTab; }
  SimMollerTables* GetTheMollerTables() const { return &fMollerTablesTab; }
  SimSBTables* GetTheSBTables() const { return &fSBTablesTab; }
  SimGSTables* GetTheGSTables() const { return &fGSTablesTab; }

  // node data as read (flattened for the device by FlattenTables)
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

 private:
  mutable SimITr1MFPElastic fITr1MFPElasticTab;
  mutable SimMaxScatStrength fMaxScatStrengthTab;
  mutable SimIMFPMoller fIMFPMollerTab;
  mutable SimIMFPBrem fIMFPBremTab;
  mutable SimStoppingPower fDEDXTab;
  mutable SimMollerTables fMollerTablesTab;
  mutable SimSBTables fSBTablesTab;
  mutable SimGSTables fGSTablesTab;
};

// SimPhotonData (SimPhotonData.hh:22-71, SimPhotonData.cc:28-78)
class SimPhotonData {
 public:
  void Load(const std::string& dataDir, int verbose = 0);
  bool Load(const std::string& dataDir, std::string& err, int verbose = 0);

  // SimPhotonData.hh:33-39
  SimIMFPMaxPhoton* GetIMFPTotalMax() const { return &fIMFPTotalMaxTab; }
  SimIMFPPhoton* GetIMFPTotal() const { return &fIMFPTotalTab; }
  SimIMFPPhoton* GetIMFPCompton() const { return &fIMFPComptonTab; }
  SimIMFPPhoton* GetIMFPPairProd() const { return &fIMFPPairProdTab; }
  SimKNTables* GetTheKNTables() const { return &fKNTablesTab; }

  SimNodes fIMFPTotalMax;             // imfp_globalMax.dat
  std::vector<SimNodes> fIMFPTotal;   // imfp_total.dat
  std::vector<SimNodes> fIMFPCompton; // imfp_compton.dat
  std::vector<SimNodes> fIMFPPairProd;// imfp_pairp.dat
  SimAliasSet fKNTables;              // compton_KNDtrData.dat

 private:
  mutable SimIMFPMaxPhoton fIMFPTotalMaxTab;
  mutable SimIMFPPhoton fIMFPTotalTab{0};
  mutable SimIMFPPhoton fIMFPComptonTab{1};
  mutable SimIMFPPhoton fIMFPPairProdTab{2};
  mutable SimKNTables fKNTablesTab;
};

// Readers of the single table files (file names, comment-line counts and
// columns of the reference loader each cites); false and `err` on failure.
bool SimReadITr1MFP(const std::string& dataDir, std::vector<SimNodes>& out, std::string& err);
bool SimReadMaxScatStrength(const std::string& dataDir, SimNodes& out, std::string& err);
bool SimReadIMFPMoller(const std::string& dataDir, SimNodes& out, std::string& err);
bool SimReadIMFPBrem(const std::string& dataDir, std::vector<SimNodes>& out, std::string& err);
bool SimReadDEDX(const std::string& dataDir, std::vector<SimNodes>& out, std::string& err);
bool SimReadMollerTables(const std::string& dataDir, SimAliasSet& out, std::string& err);
bool SimReadSBTables(const std::string& dataDir, SimAliasSet& out, std::string& err);
bool SimReadGSTables(const std::string& dataDir, int& ne, int& ns, std::vector<double>& energy,
                     std::vector<double>& transf, std::vector<double>& u, std::vector<double>& a,
                     std::vector<double>& b, std::string& err);
bool SimReadIMFPMax(const std::string& dataDir, SimNodes& out, std::string& err);
bool SimReadIMFPPhoton(const std::string& dataDir, int type, std::vector<SimNodes>& out, std::string& err);
bool SimReadKNTables(const std::string& dataDir, SimAliasSet& out, std::string& err);

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
