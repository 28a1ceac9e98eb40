// SimData.cc -- readers of the reference's run-time table files.
//
// Each reader follows the corresponding reference reader's layout exactly
// (file name, number of comment lines skipped, columns used); see the
// citations per function. Load(dir, err) reports an error as text;
// Load(dir[, verbose]) prints it and ends the process with EXIT_FAILURE, as
// the reference's loaders do.
#include "SimData.hh"

#include <cstdio>
#include <cstdlib>
#include <iostream>

namespace {

// Thin cursor over a text table file with the two primitives the reference
// readers use: skipping whole lines (fgets) and scanning numbers (fscanf,
// where a trailing "\n" in the format swallows following white space).
class TableFile {
 public:
  explicit TableFile(const std::string& path) : path_(path), f_(std::fopen(path.c_str(), "r")) {}
  ~TableFile() { if (f_) std::fclose(f_); }
  bool ok() const { return f_ != nullptr; }
  const std::string& path() const { return path_; }
  void skipLines(int k) {
    char buf[1024];
    for (int i = 0; i < k; ++i)
      if (!std::fgets(buf, sizeof(buf), f_)) return;
  }
  bool num(double& v) { return std::fscanf(f_, "%lg", &v) == 1; }
  bool num(int& v) { return std::fscanf(f_, "%d", &v) == 1; }
  bool word(std::string& s) {
    char buf[512];
    if (std::fscanf(f_, "%511s", buf) != 1) return false;
    s = buf;
    return true;
  }
  void eatSpace() { int r = std::fscanf(f_, " "); (void)r; }

 private:
  std::string path_;
  FILE* f_;
};

bool fail(std::string& err, const std::string& what) {
  err = what;
  return false;
}

// per-material node files: n, #mat on one line, one skipped line, then per
// material 3 comment lines and n rows; `ycol` selects the value column
bool readPerMaterial(const std::string& path, int headLines, int ncols, int ycol, std::vector<SimNodes>& out,
                     std::string& err) {
  TableFile tf(path);
  if (!tf.ok()) return fail(err, "file = " + path + " not found!");
  tf.skipLines(headLines);
  int n = 0, nmat = 0;
  if (!tf.num(n) || !tf.num(nmat) || n < 2 || nmat < 1) return fail(err, "malformed " + path);
  tf.eatSpace();
  tf.skipLines(1);
  out.assign(nmat, SimNodes());
  for (int im = 0; im < nmat; ++im) {
    tf.skipLines(3);
    out[im].x.resize(n);
    out[im].y.resize(n);
    for (int i = 0; i < n; ++i) {
      double c[3];
      for (int k = 0; k < ncols; ++k)
        if (!tf.num(c[k])) return fail(err, "malformed " + path);
      out[im].x[i] = c[0];
      out[im].y[i] = c[ycol];
    }
    tf.eatSpace();
  }
  return true;
}

// reference-material node files: head lines, n, 4 skipped lines, n rows
bool readReference(const std::string& path, int headLines, SimNodes& out, std::string& err) {
  TableFile tf(path);
  if (!tf.ok()) return fail(err, "file = " + path + " not found!");
  tf.skipLines(headLines);
  int n = 0;
  if (!tf.num(n) || n < 2) return fail(err, "malformed " + path);
  tf.eatSpace();
  tf.skipLines(4);
  out.x.resize(n);
  out.y.resize(n);
  for (int i = 0; i < n; ++i)
    if (!tf.num(out.x[i]) || !tf.num(out.y[i])) return fail(err, "malformed " + path);
  return true;
}

bool readAliasRows(TableFile& tf, SimAliasSet& s, size_t base) {
  for (int is = 0; is < s.ns; ++is) {
    const size_t k = base + is;
    if (!tf.num(s.x[k]) || !tf.num(s.y[k]) || !tf.num(s.w[k]) || !tf.num(s.alias[k])) return false;
  }
  return true;
}

void sizeAlias(SimAliasSet& s) {
  const size_t n = (size_t)s.nmat * s.ne * s.ns;
  s.energy.resize(s.ne);
  s.x.resize(n); s.y.resize(n); s.w.resize(n); s.alias.resize(n);
}

// SimMollerTables.cc:20-67 / SimKNTables.cc:20-67: 5 comment lines, ne ns,
// then per energy: E and ns rows
bool readEnergyAlias(const std::string& path, SimAliasSet& s, std::string& err) {
  TableFile tf(path);
  if (!tf.ok()) return fail(err, "file = " + path + " not found!");
  tf.skipLines(5);
  s.nmat = 1;
  if (!tf.num(s.ne) || !tf.num(s.ns) || s.ne < 2 || s.ns < 2) return fail(err, "malformed " + path);
  sizeAlias(s);
  for (int ie = 0; ie < s.ne; ++ie) {
    if (!tf.num(s.energy[ie])) return fail(err, "malformed " + path);
    if (!readAliasRows(tf, s, (size_t)ie * s.ns)) return fail(err, "malformed " + path);
  }
  return true;
}

}  // namespace

// ---------------------------------------------------------------------------
// readers of the single files
// ---------------------------------------------------------------------------
// SimITr1MFPElastic.cc:13-56 -- column 3 (scattering-power corrected)
bool SimReadITr1MFP(const std::string& dataDir, std::vector<SimNodes>& out, std::string& err) {
  return readPerMaterial(dataDir + "/el_itr1mfp.dat", 4, 3, 2, out, err);
}
// SimMaxScatStrength.cc:6-43
bool SimReadMaxScatStrength(const std::string& dataDir, SimNodes& out, std::string& err) {
  return readReference(dataDir + "/el_scatStrength.dat", 4, out, err);
}
// SimIMFPMoller.cc:6-43
bool SimReadIMFPMoller(const std::string& dataDir, SimNodes& out, std::string& err) {
  return readReference(dataDir + "/imfp_moller.dat", 3, out, err);
}
// SimIMFPBrem.cc:18-61
bool SimReadIMFPBrem(const std::string& dataDir, std::vector<SimNodes>& out, std::string& err) {
  return readPerMaterial(dataDir + "/imfp_brem.dat", 3, 2, 1, out, err);
}
// SimStoppingPower.cc:18-61 (two fused comment lines in the file)
bool SimReadDEDX(const std::string& dataDir, std::vector<SimNodes>& out, std::string& err) {
  return readPerMaterial(dataDir + "/eloss_rdedx.dat", 2, 2, 1, out, err);
}
// SimMollerTables.cc:20-67
bool SimReadMollerTables(const std::string& dataDir, SimAliasSet& out, std::string& err) {
  return readEnergyAlias(dataDir + "/ioni_MollerDtrData.dat", out, err);
}
// SimSBTables.cc:21-90
bool SimReadSBTables(const std::string& dataDir, SimAliasSet& s, std::string& err) {
  const std::string path = dataDir + "/brem_SBDtrData.dat";
  TableFile tf(path);
  if (!tf.ok()) return fail(err, "file = " + path + " not found!");
  tf.skipLines(5);
  if (!tf.num(s.ne) || !tf.num(s.ns) || !tf.num(s.nmat) || s.ne < 2 || s.ns < 2 || s.nmat < 1)
    return fail(err, "malformed " + path);
  sizeAlias(s);
  for (int ie = 0; ie < s.ne; ++ie)
    if (!tf.num(s.energy[ie])) return fail(err, "malformed " + path);
  tf.eatSpace();
  for (int im = 0; im < s.nmat; ++im) {
    tf.skipLines(3);
    for (int ie = 0; ie < s.ne; ++ie)
      if (!readAliasRows(tf, s, ((size_t)im * s.ne + ie) * s.ns)) return fail(err, "malformed " + path);
    tf.eatSpace();
  }
  return true;
}
// SimGSTables.cc:19-80 (reference material only)
bool SimReadGSTables(const std::string& dataDir, int& ne, int& ns, std::vector<double>& energy,
                     std::vector<double>& transf, std::vector<double>& u, std::vector<double>& a,
                     std::vector<double>& b, std::string& err) {
  const std::string path = dataDir + "/el_GSDtrData.dat";
  TableFile tf(path);
  if (!tf.ok()) return fail(err, "file = " + path + " not found!");
  tf.skipLines(4);
  double emin = 0, emax = 0;
  if (!tf.num(ne) || !tf.num(emin) || !tf.num(emax) || !tf.num(ns) || ne < 2 || ns < 2)
    return fail(err, "malformed " + path);
  tf.eatSpace();
  tf.skipLines(6);
  const size_t n = (size_t)ne * ns;
  energy.resize(ne);
  transf.resize(ne);
  u.resize(n); a.resize(n); b.resize(n);
  for (int ie = 0; ie < ne; ++ie) {
    int idx = 0;
    if (!tf.num(idx) || !tf.num(energy[ie]) || !tf.num(transf[ie])) return fail(err, "malformed " + path);
    for (int is = 0; is < ns; ++is) {
      const size_t k = (size_t)ie * ns + is;
      if (!tf.num(u[k]) || !tf.num(a[k]) || !tf.num(b[k])) return fail(err, "malformed " + path);
    }
  }
  return true;
}
// SimIMFPMaxPhoton.cc:6-38: 3 comment lines, n, 1 comment line, n rows
bool SimReadIMFPMax(const std::string& dataDir, SimNodes& out, std::string& err) {
  const std::string path = dataDir + "/imfp_globalMax.dat";
  TableFile tf(path);
  if (!tf.ok()) return fail(err, "file = " + path + " not found!");
  tf.skipLines(3);
  int n = 0;
  if (!tf.num(n) || n < 2) return fail(err, "malformed " + path);
  tf.eatSpace();
  tf.skipLines(1);
  out.x.resize(n);
  out.y.resize(n);
  for (int i = 0; i < n; ++i)
    if (!tf.num(out.x[i]) || !tf.num(out.y[i])) return fail(err, "malformed " + path);
  return true;
}
// SimIMFPPhoton.cc:19-63 (type 0 total, 1 Compton, 2 pair production)
bool SimReadIMFPPhoton(const std::string& dataDir, int type, std::vector<SimNodes>& out, std::string& err) {
  static const char* kFile[3] = {"/imfp_total.dat", "/imfp_compton.dat", "/imfp_pairp.dat"};
  if (type < 0 || type > 2) return fail(err, "unknown photon IMFP type");
  return readPerMaterial(dataDir + kFile[type], 3, 2, 1, out, err);
}
// SimKNTables.cc:20-67
bool SimReadKNTables(const std::string& dataDir, SimAliasSet& out, std::string& err) {
  return readEnergyAlias(dataDir + "/compton_KNDtrData.dat", out, err);
}

namespace {
// the reference's loaders print the reason and end the process
void LoadOrExit(bool ok, const std::string& err) {
  if (ok) return;
  std::cerr << "  *** ERROR: " << err << std::endl;
  std::exit(EXIT_FAILURE);
}
}  // namespace

// SimMaterialData::Load (SimMaterialData.cc:6-53)
void SimMaterialData::Load(const std::string& dataDir, int verbose) {
  std::string err;
  LoadOrExit(Load(dataDir, err, verbose), err);
}

bool SimMaterialData::Load(const std::string& dataDir, std::string& err, int verbose) {
  const std::string path = dataDir + "/mat.dat";
  TableFile tf(path);
  if (!tf.ok()) return fail(err, "file = " + path + " not found!");
  tf.skipLines(1);
  if (!tf.num(fElectronCut) || !tf.num(fGammaCut)) return fail(err, "malformed " + path);
  tf.eatSpace();
  tf.skipLines(1);
  if (!tf.num(fMSCStepParSLow) || !tf.num(fMSCStepParSHigh) || !tf.num(fMSCStepParEcross))
    return fail(err, "malformed " + path);
  tf.eatSpace();
  tf.skipLines(1);
  if (!tf.num(fNumMaterial) || fNumMaterial < 1 || fNumMaterial > 255) return fail(err, "malformed " + path);
  tf.eatSpace();
  fMaterialName.resize(fNumMaterial);
  fMaterialDensity.resize(fNumMaterial);
  fMollerIMFPScaling.resize(fNumMaterial);
  for (int im = 0; im < fNumMaterial; ++im) {
    tf.skipLines(2);
    int idx = 0;
    double dummy = 0;
    if (!tf.num(idx) || !tf.word(fMaterialName[im]) || !tf.num(fMaterialDensity[im]))
      return fail(err, "malformed " + path);
    tf.eatSpace();
    tf.skipLines(3);
    if (!tf.num(fMollerIMFPScaling[im]) || !tf.num(dummy)) return fail(err, "malformed " + path);
    tf.eatSpace();
    if (verbose > 0) std::cout << "    --- Material data for: " << fMaterialName[im] << std::endl;
  }
  return true;
}

// SimElectronData::Load (SimElectronData.cc:44-115) and the loaders it calls
void SimElectronData::Load(const std::string& dataDir, int verbose) {
  std::string err;
  LoadOrExit(Load(dataDir, err, verbose), err);
}

bool SimElectronData::Load(const std::string& dataDir, std::string& err, int verbose) {
  if (!SimReadITr1MFP(dataDir, fITr1MFP, err) || !SimReadMaxScatStrength(dataDir, fMaxScatStrength, err) ||
      !SimReadIMFPMoller(dataDir, fIMFPMoller, err) || !SimReadIMFPBrem(dataDir, fIMFPBrem, err) ||
      !SimReadDEDX(dataDir, fDEDX, err) || !SimReadMollerTables(dataDir, fMollerTables, err) ||
      !SimReadSBTables(dataDir, fSBTables, err) ||
      !SimReadGSTables(dataDir, fGSNumEnergies, fGSNumCum, fGSEnergy, fGSTransf, fGSU, fGSA, fGSB, err))
    return false;
  fITr1MFPElasticTab.Set(fITr1MFP);
  fMaxScatStrengthTab.Set({fMaxScatStrength});
  fIMFPMollerTab.Set({fIMFPMoller});
  fIMFPBremTab.Set(fIMFPBrem);
  fDEDXTab.Set(fDEDX);
  fMollerTablesTab.Set(fMollerTables);
  fSBTablesTab.Set(fSBTables);
  fGSTablesTab.Set(fGSNumEnergies, fGSNumCum, fGSEnergy, fGSTransf, fGSU, fGSA, fGSB);
  if (verbose > 0) std::cout << " == Electron data loaded from " << dataDir << std::endl;
  return true;
}

// SimPhotonData::Load (SimPhotonData.cc:28-78)
void SimPhotonData::Load(const std::string& dataDir, int verbose) {
  std::string err;
  LoadOrExit(Load(dataDir, err, verbose), err);
}

bool SimPhotonData::Load(const std::string& dataDir, std::string& err, int verbose) {
  if (!SimReadIMFPMax(dataDir, fIMFPTotalMax, err) || !SimReadIMFPPhoton(dataDir, 0, fIMFPTotal, err) ||
      !SimReadIMFPPhoton(dataDir, 1, fIMFPCompton, err) || !SimReadIMFPPhoton(dataDir, 2, fIMFPPairProd, err) ||
      !SimReadKNTables(dataDir, fKNTables, err))
    return false;
  fIMFPTotalMaxTab.Set({fIMFPTotalMax});
  fIMFPTotalTab.Set(fIMFPTotal);
  fIMFPComptonTab.Set(fIMFPCompton);
  fIMFPPairProdTab.Set(fIMFPPairProd);
  fKNTablesTab.Set(fKNTables);
  if (verbose > 0) std::cout << " == Photon data loaded from " << dataDir << std::endl;
  return true;
}

static dpm_spline_desc splineOf(const SimNodes& n) {
  dpm_spline_desc d;
  d.n = (int)n.x.size();
  d.x = n.x.data();
  d.y = n.y.data();
  return d;
}
static dpm_linear_desc linearOf(const SimNodes& n) {
  dpm_linear_desc d;
  d.n = (int)n.x.size();
  d.x = n.x.data();
  d.y = n.y.data();
  return d;
}
static dpm_alias_desc aliasOf(const SimAliasSet& s) {
  dpm_alias_desc d;
  d.nmat = s.nmat; d.ne = s.ne; d.ns = s.ns;
  d.energy = s.energy.data();
  d.x = s.x.data(); d.y = s.y.data(); d.w = s.w.data(); d.alias = s.alias.data();
  return d;
}

void FlattenTables(const SimMaterialData& mat, const SimElectronData& el, const SimPhotonData& ph,
                   SimTableView& keep) {
  const int nm = mat.fNumMaterial;
  keep.itr1.clear(); keep.dedx.clear(); keep.brem.clear();
  keep.ptot.clear(); keep.pcomp.clear(); keep.ppair.clear();
  for (int im = 0; im < nm; ++im) {
    keep.itr1.push_back(splineOf(el.fITr1MFP[im]));
    keep.dedx.push_back(splineOf(el.fDEDX[im]));
    keep.brem.push_back(splineOf(el.fIMFPBrem[im]));
    keep.ptot.push_back(linearOf(ph.fIMFPTotal[im]));
    keep.pcomp.push_back(linearOf(ph.fIMFPCompton[im]));
    keep.ppair.push_back(linearOf(ph.fIMFPPairProd[im]));
  }
  dpm_table_desc& d = keep.desc;
  d.nmat = nm;
  d.electron_cut = mat.fElectronCut;
  d.gamma_cut = mat.fGammaCut;
  d.density = mat.fMaterialDensity.data();
  d.moller_scale = mat.fMollerIMFPScaling.data();
  d.itr1 = keep.itr1.data();
  d.dedx = keep.dedx.data();
  d.brem = keep.brem.data();
  d.moller = splineOf(el.fIMFPMoller);
  d.k1 = splineOf(el.fMaxScatStrength);
  d.gmax = linearOf(ph.fIMFPTotalMax);
  d.ptotal = keep.ptot.data();
  d.pcompton = keep.pcomp.data();
  d.ppair = keep.ppair.data();
  d.sb = aliasOf(el.fSBTables);
  d.moller_tab = aliasOf(el.fMollerTables);
  d.kn = aliasOf(ph.fKNTables);
  d.gs.ne = el.fGSNumEnergies;
  d.gs.ns = el.fGSNumCum;
  d.gs.energy = el.fGSEnergy.data();
  d.gs.transf = el.fGSTransf.data();
  d.gs.u = el.fGSU.data();
  d.gs.a = el.fGSA.data();
  d.gs.b = el.fGSB.data();
}
