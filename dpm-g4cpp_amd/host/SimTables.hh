// SimTables.hh -- the reference's run-time table classes on the host.
//
// Same class names and public interface as the reference's Simulation/inc
// headers, for code written against its data layer (dpm_Simulate's main, the
// model tests): the interpolators SimITr1MFPElastic, SimMaxScatStrength,
// SimIMFPMoller, SimIMFPBrem, SimStoppingPower, SimIMFPMaxPhoton,
// SimIMFPPhoton (Get*, with the reference's energy clamp and 1e-20 floor) and
// the samplers SimMollerTables, SimSBTables, SimKNTables, SimGSTables
// (SampleEnergyTransfer / SampleAngularDeflection with explicit uniforms).
// They evaluate in FP64 on the host with the reference's arithmetic; the
// transport never calls them -- it runs on the GPU (include/dpm_amd.h), whose
// dpm_eval / dpm_sample are the device forms of the same functions.
#ifndef DPM_AMD_SIMTABLES_HH
#define DPM_AMD_SIMTABLES_HH

#include <string>
#include <vector>

// node data of one interpolated quantity
struct SimNodes {
  std::vector<double> x, y;
};

// one family of linear-pdf alias tables
struct SimAliasSet {
  int nmat = 0, ne = 0, ns = 0;
  std::vector<double> energy, x, y, w;
  std::vector<int> alias;
};

// natural cubic spline coefficients of SimDataSpline::SetUp
// (SimDataSpline.cc:33-67), in its operation order
void SimSplineCoefficients(const double* x, const double* y, int n, std::vector<double>& B, std::vector<double>& C,
                           std::vector<double>& D);

// SimDataSpline (SimDataSpline.hh:20-60): log-spaced nodes, cubic spline
class SimDataSpline {
 public:
  void Set(const SimNodes& nodes);
  double GetValueAt(double xval, int ilow) const;
  double GetValue(double xval, double logxval) const;
  double GetValue(double xval) const;
  int fNumData = 0;
  double fLogMinX = 0, fInvLogDelta = 0;
  std::vector<double> fX, fY, fB, fC, fD;
};

// SimDataLinear (SimDataLinear.hh:20-55): equally spaced nodes, linear
class SimDataLinear {
 public:
  void Set(const SimNodes& nodes);
  double GetValueAt(double xval, int ilow) const;
  double GetValue(double xval) const;
  int fNumData = 0;
  double fMinX = 0, fInvDelta = 0;
  std::vector<double> fX, fY;
};

// Per-material log-grid quantities: clamp to [Emin, Emax - 1e-6] of material
// 0 (the reference keeps the first material's range for all), floor 1e-20
class SimSplineSet {
 public:
  void Set(const std::vector<SimNodes>& perMaterial);
  double Value(double ekin, int imat) const;
  double Value(double ekin, double logekin, int imat) const;
  double ValueAt(double ekin, int ilow, int imat) const;
  std::vector<SimDataSpline> fDataPerMaterial;
  double fEmin = 0, fEmax = 0;
};

// SimITr1MFPElastic (SimITr1MFPElastic.hh:22-60)
class SimITr1MFPElastic : public SimSplineSet {
 public:
  void LoadData(const std::string& dataDir, int verbose);
  double GetITr1MFPPerDensity(double ekin, int imat) { return Value(ekin, imat); }
  double GetITr1MFPPerDensity(double ekin, double logekin, int imat) { return Value(ekin, logekin, imat); }
  double GetITr1MFPPerDensity(double ekin, int ilow, int imat) { return ValueAt(ekin, ilow, imat); }
};

// SimIMFPBrem (SimIMFPBrem.hh:22-66): 1e-20 for imat < 0
class SimIMFPBrem : public SimSplineSet {
 public:
  void LoadData(const std::string& dataDir, int verbose);
  double GetIMFPPerDensity(double ekin, int imat) { return imat < 0 ? 1.0E-20 : Value(ekin, imat); }
  double GetIMFPPerDensity(double ekin, double logekin, int imat) {
    return imat < 0 ? 1.0E-20 : Value(ekin, logekin, imat);
  }
  double GetIMFPPerDensity(double ekin, int ilow, int imat) { return imat < 0 ? 1.0E-20 : ValueAt(ekin, ilow, imat); }
};

// SimStoppingPower (SimStoppingPower.hh:22-66): 1e-20 for imat < 0
class SimStoppingPower : public SimSplineSet {
 public:
  void LoadData(const std::string& dataDir, int verbose);
  double GetDEDXPerDensity(double ekin, int imat) { return imat < 0 ? 1.0E-20 : Value(ekin, imat); }
  double GetDEDXPerDensity(double ekin, double logekin, int imat) {
    return imat < 0 ? 1.0E-20 : Value(ekin, logekin, imat);
  }
  double GetDEDXPerDensity(double ekin, int ilow, int imat) { return imat < 0 ? 1.0E-20 : ValueAt(ekin, ilow, imat); }
};

// SimIMFPMoller (SimIMFPMoller.hh:20-58): reference material only
class SimIMFPMoller : public SimSplineSet {
 public:
  void LoadData(const std::string& dataDir, int verbose);
  double GetIMFPPerDensity(double ekin) { return Value(ekin, 0); }
  double GetIMFPPerDensity(double ekin, double logekin) { return Value(ekin, logekin, 0); }
  double GetIMFPPerDensity(double ekin, int ilow) { return ValueAt(ekin, ilow, 0); }
};

// SimMaxScatStrength (SimMaxScatStrength.hh:22-62): reference material only
class SimMaxScatStrength : public SimSplineSet {
 public:
  void LoadData(const std::string& dataDir, int verbose);
  double GetMaxScatStrength(double ekin) { return Value(ekin, 0); }
  double GetMaxScatStrength(double ekin, double logekin) { return Value(ekin, logekin, 0); }
  double GetMaxScatStrength(double ekin, int ilow) { return ValueAt(ekin, ilow, 0); }
};

// Per-material linear-grid photon quantities, clamped like SimSplineSet
class SimLinearSet {
 public:
  void Set(const std::vector<SimNodes>& perMaterial);
  double Value(double ekin, int imat) const;
  double ValueAt(double ekin, int ilow, int imat) const;
  std::vector<SimDataLinear> fDataPerMaterial;
  double fEmin = 0, fEmax = 0;
};

// SimIMFPMaxPhoton (SimIMFPMaxPhoton.hh:20-52)
class SimIMFPMaxPhoton : public SimLinearSet {
 public:
  void LoadData(const std::string& dataDir, int verbose);
  double GetIMFP(double ekin) { return Value(ekin, 0); }
  double GetIMFP(double ekin, int ilow) { return ValueAt(ekin, ilow, 0); }
};

// SimIMFPPhoton (SimIMFPPhoton.hh:22-66): type 0 total, 1 Compton, 2 pair
class SimIMFPPhoton : public SimLinearSet {
 public:
  explicit SimIMFPPhoton(int type = 0) : fType(type) {}
  void LoadData(const std::string& dataDir, int verbose);
  double GetIMFPPerDensity(double ekin, int imat) { return imat < 0 ? 1.0E-20 : Value(ekin, imat); }
  double GetIMFPPerDensity(double ekin, int ilow, int imat) { return imat < 0 ? 1.0E-20 : ValueAt(ekin, ilow, imat); }
  int fType;
};

// SimLinAliasData::Sample (SimLinAliasData.cc:25-42) of one table of `n`
// points starting at `k` in a SimAliasSet
double SimAliasSample(const SimAliasSet& s, size_t k, double rndm1, double rndm2);

// the log-energy grid of a sampling-table family: table ie or ie + 1 with
// the probability of the position between them (e.g. SimSBTables.cc:96-102)
struct SimEnergyGrid {
  void Set(const std::vector<double>& energy);
  int Pick(double lpenergy, double rndm1, int* ielow = nullptr) const;
  double fMinPrimaryEnergy = 0, fLogMinPrimaryEnergy = 0, fInvLogDeltaPrimaryEnergy = 0;
};

// SimMollerTables (SimMollerTables.hh:22-74)
class SimMollerTables {
 public:
  void LoadData(const std::string& dataDir, int verbose);
  void Set(const SimAliasSet& s);
  double SampleEnergyTransfer(double eprim, double rndm1, double rndm2, double rndm3);
  SimAliasSet fTables;
  SimEnergyGrid fGrid;
};

// SimSBTables (SimSBTables.hh:22-80)
class SimSBTables {
 public:
  void LoadData(const std::string& dataDir, int verbose);
  void Set(const SimAliasSet& s);
  double SampleEnergyTransfer(double eprim, int imat, double rndm1, double rndm2, double rndm3);
  SimAliasSet fTables;
  SimEnergyGrid fGrid;
};

// SimKNTables (SimKNTables.hh:22-74)
class SimKNTables {
 public:
  void LoadData(const std::string& dataDir, int verbose);
  void Set(const SimAliasSet& s);
  double SampleEnergyTransfer(double egamma, double rndm1, double rndm2, double rndm3);
  SimAliasSet fTables;
  SimEnergyGrid fGrid;
};

// SimGSTables (SimGSTables.hh:22-100), reference material
class SimGSTables {
 public:
  void LoadData(const std::string& dataDir, int verbose);
  void Set(int ne, int ns, const std::vector<double>& energy, const std::vector<double>& transf,
           const std::vector<double>& u, const std::vector<double>& a, const std::vector<double>& b);
  double SampleAngularDeflection(double ekin, double rndm1, double rndm2);
  int fNumPrimaryEnergies = 0, fNumCum = 0;
  double fDeltaCum = 0;
  std::vector<double> fEnergy, fTransf, fU, fA, fB;
  SimEnergyGrid fGrid;
};

#endif
