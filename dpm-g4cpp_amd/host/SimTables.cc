// SimTables.cc -- host FP64 evaluation of the reference's table classes
// (interface in SimTables.hh). Each function follows the arithmetic of the
// reference function it names, in the same operation order, so that its
// results equal the reference's bit for bit on the same table files.
#include "SimTables.hh"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <iostream>

#include "SimData.hh"

// SimDataSpline::SetUp (SimDataSpline.cc:33-67)
void SimSplineCoefficients(const double* x, const double* y, int n, std::vector<double>& B, std::vector<double>& C,
                           std::vector<double>& D) {
  B.assign(n, 0.0); C.assign(n, 0.0); D.assign(n, 0.0);
  const int last = n - 1;
  double s = 0.0, r = 0.0;
  for (int m = 0; m < last; ++m) {
    D[m] = x[m + 1] - x[m];
    r = (y[m + 1] - y[m]) / D[m];
    C[m] = r - s;
    s = r;
  }
  r = 0.0; s = 0.0;
  C[0] = 0.0; C[last] = 0.0;
  for (int m = 1; m < last; ++m) {
    C[m] = C[m] + r * C[m - 1];
    B[m] = 2.0 * (x[m - 1] - x[m + 1]) - r * s;
    s = D[m];
    r = s / B[m];
  }
  for (int m = last - 1; m >= 1; --m) C[m] = (D[m] * C[m + 1] - C[m]) / B[m];
  for (int m = 0; m < last; ++m) {
    s = D[m];
    r = C[m + 1] - C[m];
    D[m] = r / s;
    C[m] = 3.0 * C[m];
    B[m] = (y[m + 1] - y[m]) / s - (C[m] + r) * s;
  }
}

// SimDataSpline::FillData (SimDataSpline.cc:9-30): log grid from the first two nodes
void SimDataSpline::Set(const SimNodes& nodes) {
  fNumData = (int)nodes.x.size();
  fX = nodes.x;
  fY = nodes.y;
  fLogMinX = std::log(fX[0]);
  fInvLogDelta = 1. / (std::log(fX[1]) - fLogMinX);
  SimSplineCoefficients(fX.data(), fY.data(), fNumData, fB, fC, fD);
}

double SimDataSpline::GetValueAt(double xval, int ilow) const {
  xval -= fX[ilow];
  return fY[ilow] + xval * (fB[ilow] + xval * (fC[ilow] + xval * fD[ilow]));
}

double SimDataSpline::GetValue(double xval, double logxval) const {
  int ilow = (int)((logxval - fLogMinX) * fInvLogDelta);
  ilow = std::max(0, std::min(fNumData - 2, ilow));
  return GetValueAt(xval, ilow);
}

double SimDataSpline::GetValue(double xval) const { return GetValue(xval, std::log(xval)); }

// SimDataLinear::FillData (SimDataLinear.cc): linear grid from the first two nodes
void SimDataLinear::Set(const SimNodes& nodes) {
  fNumData = (int)nodes.x.size();
  fX = nodes.x;
  fY = nodes.y;
  fMinX = fX[0];
  fInvDelta = 1. / (fX[1] - fMinX);
}

double SimDataLinear::GetValueAt(double xval, int ilow) const {
  return (fY[ilow + 1] - fY[ilow]) * (xval - fX[ilow]) / (fX[ilow + 1] - fX[ilow]) + fY[ilow];
}

double SimDataLinear::GetValue(double xval) const {
  int ilow = (int)((xval - fMinX) * fInvDelta);
  ilow = std::max(0, std::min(fNumData - 2, ilow));
  return GetValueAt(xval, ilow);
}

void SimSplineSet::Set(const std::vector<SimNodes>& perMaterial) {
  fDataPerMaterial.assign(perMaterial.size(), SimDataSpline());
  for (size_t im = 0; im < perMaterial.size(); ++im) fDataPerMaterial[im].Set(perMaterial[im]);
  fEmin = perMaterial[0].x.front();
  fEmax = perMaterial[0].x.back();
}

// e.g. SimStoppingPower::GetDEDXPerDensity (SimStoppingPower.hh:36-54)
double SimSplineSet::Value(double ekin, int imat) const {
  const double e = std::min(fEmax - 1.0E-6, std::max(fEmin, ekin));
  return std::max(1.0E-20, fDataPerMaterial[imat].GetValue(e));
}
double SimSplineSet::Value(double ekin, double logekin, int imat) const {
  const double e = std::min(fEmax - 1.0E-6, std::max(fEmin, ekin));
  return std::max(1.0E-20, fDataPerMaterial[imat].GetValue(e, logekin));
}
double SimSplineSet::ValueAt(double ekin, int ilow, int imat) const {
  return std::max(1.0E-20, fDataPerMaterial[imat].GetValueAt(ekin, ilow));
}

void SimLinearSet::Set(const std::vector<SimNodes>& perMaterial) {
  fDataPerMaterial.assign(perMaterial.size(), SimDataLinear());
  for (size_t im = 0; im < perMaterial.size(); ++im) fDataPerMaterial[im].Set(perMaterial[im]);
  fEmin = perMaterial[0].x.front();
  fEmax = perMaterial[0].x.back();
}

// e.g. SimIMFPPhoton::GetIMFPPerDensity (SimIMFPPhoton.hh:40-46)
double SimLinearSet::Value(double ekin, int imat) const {
  const double e = std::min(fEmax - 1.0E-6, std::max(fEmin, ekin));
  return std::max(1.0E-20, fDataPerMaterial[imat].GetValue(e));
}
double SimLinearSet::ValueAt(double ekin, int ilow, int imat) const {
  return std::max(1.0E-20, fDataPerMaterial[imat].GetValueAt(ekin, ilow));
}

// SimLinAliasData::Sample (SimLinAliasData.cc:25-42)
double SimAliasSample(const SimAliasSet& s, size_t k, double rndm1, double rndm2) {
  const double rest = rndm1 * (s.ns - 1);
  int indxl = (int)(rest);
  if (s.w[k + indxl] < rest - indxl) indxl = s.alias[k + indxl];
  const double xval = s.x[k + indxl];
  const double xdelta = s.x[k + indxl + 1] - xval;
  if (s.y[k + indxl] > 0.0) {
    const double dum = (s.y[k + indxl + 1] - s.y[k + indxl]) / s.y[k + indxl];
    if (std::abs(dum) > 0.1) return xval - xdelta / dum * (1.0 - std::sqrt(1.0 + rndm2 * dum * (dum + 2.0)));
    return xval + rndm2 * xdelta * (1.0 - 0.5 * dum * (rndm2 - 1.0) * (1.0 + dum * rndm2));
  }
  return xval + xdelta * std::sqrt(rndm2);
}

void SimEnergyGrid::Set(const std::vector<double>& energy) {
  fMinPrimaryEnergy = energy[0];
  fLogMinPrimaryEnergy = std::log(energy[0]);
  fInvLogDeltaPrimaryEnergy = 1. / (std::log(energy[1]) - fLogMinPrimaryEnergy);
}

// e.g. SimSBTables.cc:96-102: the lower grid table or the next one
int SimEnergyGrid::Pick(double lpenergy, double rndm1, int* ielow) const {
  double phigher = (lpenergy - fLogMinPrimaryEnergy) * fInvLogDeltaPrimaryEnergy;
  int penergyindx = (int)phigher;
  if (ielow) *ielow = penergyindx;
  phigher -= penergyindx;
  if (rndm1 < phigher) ++penergyindx;
  return penergyindx;
}

void SimMollerTables::Set(const SimAliasSet& s) {
  fTables = s;
  fGrid.Set(s.energy);
}
// SimMollerTables::SampleEnergyTransfer (SimMollerTables.cc:69-89)
double SimMollerTables::SampleEnergyTransfer(double eprim, double rndm1, double rndm2, double rndm3) {
  const double lpenergy = std::log(eprim);
  const int ie = fGrid.Pick(lpenergy, rndm1);
  const double xi = SimAliasSample(fTables, (size_t)ie * fTables.ns, rndm2, rndm3);
  const double dum1 = lpenergy - fGrid.fLogMinPrimaryEnergy;
  return std::exp(xi * dum1) * 0.5 * fGrid.fMinPrimaryEnergy;
}

void SimSBTables::Set(const SimAliasSet& s) {
  fTables = s;
  fGrid.Set(s.energy);
}
// SimSBTables::SampleEnergyTransfer (SimSBTables.cc:93-114)
double SimSBTables::SampleEnergyTransfer(double eprim, int imat, double rndm1, double rndm2, double rndm3) {
  if (imat < 0) return 0.0;
  const double lpenergy = std::log(eprim);
  const int ie = fGrid.Pick(lpenergy, rndm1);
  const double xi = SimAliasSample(fTables, ((size_t)imat * fTables.ne + ie) * fTables.ns, rndm2, rndm3);
  const double kappac = fGrid.fMinPrimaryEnergy / eprim;
  const double kappa = kappac * std::exp(-xi * (fGrid.fLogMinPrimaryEnergy - lpenergy));
  return kappa * eprim;
}

void SimKNTables::Set(const SimAliasSet& s) {
  fTables = s;
  fGrid.Set(s.energy);
}
// SimKNTables::SampleEnergyTransfer (SimKNTables.cc:69-90): eps = E1 / E0
double SimKNTables::SampleEnergyTransfer(double egamma, double rndm1, double rndm2, double rndm3) {
  const double kEMC2 = 0.510991;
  const double lpenergy = std::log(egamma);
  const int ie = fGrid.Pick(lpenergy, rndm1);
  const double xi = SimAliasSample(fTables, (size_t)ie * fTables.ns, rndm2, rndm3);
  const double kappa = egamma / kEMC2;
  return std::exp(std::log(1. + 2. * kappa) * (xi - 1.));
}

void SimGSTables::Set(int ne, int ns, const std::vector<double>& energy, const std::vector<double>& transf,
                      const std::vector<double>& u, const std::vector<double>& a, const std::vector<double>& b) {
  fNumPrimaryEnergies = ne;
  fNumCum = ns;
  fDeltaCum = 1.0 / (ns - 1);
  fEnergy = energy; fTransf = transf; fU = u; fA = a; fB = b;
  fGrid.Set(energy);
}
// SimGSTables::SampleAngularDeflection (SimGSTables.cc:84-119): cos(theta)
double SimGSTables::SampleAngularDeflection(double ekin, double rndm1, double rndm2) {
  int ielow = 0;
  const int ie = fGrid.Pick(std::log(ekin), rndm1, &ielow);
  const int indxl = (int)(rndm2 / fDeltaCum);
  const double resid = rndm2 - indxl * fDeltaCum;
  const size_t k = (size_t)ie * fNumCum + indxl;
  const double parA = fA[k], parB = fB[k], u0 = fU[k], u1 = fU[k + 1];
  const double dum1 = (1.0 + parA + parB) * fDeltaCum * resid;
  const double dum2 = fDeltaCum * fDeltaCum + parA * fDeltaCum * resid + parB * resid * resid;
  const double theU = u0 + dum1 / dum2 * (u1 - u0);
  const double a0 = fTransf[ielow], a1 = fTransf[ielow + 1];
  const double e0 = fEnergy[ielow], e1 = fEnergy[ielow + 1];
  const double parTransf = (a1 - a0) / (e1 - e0) * (ekin - e0) + a0;
  return 1. - 2. * parTransf * theU / (1. - theU + parTransf);
}

// ---------------------------------------------------------------------------
// LoadData of the single classes: the file of each through the data-layer
// readers; a failure ends the process as the reference's loaders do
// ---------------------------------------------------------------------------
namespace {
[[noreturn]] void LoadFailed(const std::string& err) {
  std::cerr << "  *** ERROR: " << err << std::endl;
  std::exit(EXIT_FAILURE);
}
void Check(bool ok, const std::string& err) {
  if (!ok) LoadFailed(err);
}
}  // namespace

void SimITr1MFPElastic::LoadData(const std::string& dataDir, int) {
  std::vector<SimNodes> n;
  std::string err;
  Check(SimReadITr1MFP(dataDir, n, err), err);
  Set(n);
}
void SimIMFPBrem::LoadData(const std::string& dataDir, int) {
  std::vector<SimNodes> n;
  std::string err;
  Check(SimReadIMFPBrem(dataDir, n, err), err);
  Set(n);
}
void SimStoppingPower::LoadData(const std::string& dataDir, int) {
  std::vector<SimNodes> n;
  std::string err;
  Check(SimReadDEDX(dataDir, n, err), err);
  Set(n);
}
void SimIMFPMoller::LoadData(const std::string& dataDir, int) {
  std::vector<SimNodes> n(1);
  std::string err;
  Check(SimReadIMFPMoller(dataDir, n[0], err), err);
  Set(n);
}
void SimMaxScatStrength::LoadData(const std::string& dataDir, int) {
  std::vector<SimNodes> n(1);
  std::string err;
  Check(SimReadMaxScatStrength(dataDir, n[0], err), err);
  Set(n);
}
void SimIMFPMaxPhoton::LoadData(const std::string& dataDir, int) {
  std::vector<SimNodes> n(1);
  std::string err;
  Check(SimReadIMFPMax(dataDir, n[0], err), err);
  Set(n);
}
void SimIMFPPhoton::LoadData(const std::string& dataDir, int) {
  std::vector<SimNodes> n;
  std::string err;
  Check(SimReadIMFPPhoton(dataDir, fType, n, err), err);
  Set(n);
}
void SimMollerTables::LoadData(const std::string& dataDir, int) {
  SimAliasSet s;
  std::string err;
  Check(SimReadMollerTables(dataDir, s, err), err);
  Set(s);
}
void SimSBTables::LoadData(const std::string& dataDir, int) {
  SimAliasSet s;
  std::string err;
  Check(SimReadSBTables(dataDir, s, err), err);
  Set(s);
}
void SimKNTables::LoadData(const std::string& dataDir, int) {
  SimAliasSet s;
  std::string err;
  Check(SimReadKNTables(dataDir, s, err), err);
  Set(s);
}
void SimGSTables::LoadData(const std::string& dataDir, int) {
  int ne = 0, ns = 0;
  std::vector<double> e, tr, u, a, b;
  std::string err;
  Check(SimReadGSTables(dataDir, ne, ns, e, tr, u, a, b, err), err);
  Set(ne, ns, e, tr, u, a, b);
}
