// host_api_driver.cc -- TEST DRIVER for the host data layer (host/SimData.hh,
// host/SimTables.hh): loads a table directory through the reference-style
// Load(dataDir) of the three data classes and answers queries read from
// stdin through their getters, one result per line (%.17g):
//   sb E imat r1 r2 r3 | mo E r1 r2 r3 | kn E r1 r2 r3 | gs E r1 r2
//   ev which E imat     (which: 0 ITr1 1 dE/dx 2 brem 3 Moller 4 K1 5 max
//                        total IMFP 6 total 7 Compton 8 pair)
#include <cstdio>
#include <cstring>
#include <string>

#include "SimElectronData.hh"
#include "SimMaterialData.hh"
#include "SimPhotonData.hh"

int main(int argc, char** argv) {
  if (argc < 2) return 2;
  SimMaterialData mat;
  SimElectronData el;
  SimPhotonData ph;
  mat.Load(argv[1]);
  el.Load(argv[1]);
  ph.Load(argv[1]);
  char kind[16];
  while (std::scanf("%15s", kind) == 1) {
    double e, r1, r2, r3, v = 0;
    int m = 0, w = 0;
    if (!std::strcmp(kind, "sb")) {
      if (std::scanf("%lf %d %lf %lf %lf", &e, &m, &r1, &r2, &r3) != 5) return 3;
      v = el.GetTheSBTables()->SampleEnergyTransfer(e, m, r1, r2, r3);
    } else if (!std::strcmp(kind, "mo")) {
      if (std::scanf("%lf %lf %lf %lf", &e, &r1, &r2, &r3) != 4) return 3;
      v = el.GetTheMollerTables()->SampleEnergyTransfer(e, r1, r2, r3);
    } else if (!std::strcmp(kind, "kn")) {
      if (std::scanf("%lf %lf %lf %lf", &e, &r1, &r2, &r3) != 4) return 3;
      v = ph.GetTheKNTables()->SampleEnergyTransfer(e, r1, r2, r3);
    } else if (!std::strcmp(kind, "gs")) {
      if (std::scanf("%lf %lf %lf", &e, &r1, &r2) != 3) return 3;
      v = el.GetTheGSTables()->SampleAngularDeflection(e, r1, r2);
    } else if (!std::strcmp(kind, "ev")) {
      if (std::scanf("%d %lf %d", &w, &e, &m) != 3) return 3;
      switch (w) {
        case 0: v = el.GetITr1MFPElastic()->GetITr1MFPPerDensity(e, m); break;
        case 1: v = el.GetDEDX()->GetDEDXPerDensity(e, m); break;
        case 2: v = el.GetIMFPBrem()->GetIMFPPerDensity(e, m); break;
        case 3: v = el.GetIMFPMoller()->GetIMFPPerDensity(e); break;
        case 4: v = el.GetMaxScatStrength()->GetMaxScatStrength(e); break;
        case 5: v = ph.GetIMFPTotalMax()->GetIMFP(e); break;
        case 6: v = ph.GetIMFPTotal()->GetIMFPPerDensity(e, m); break;
        case 7: v = ph.GetIMFPCompton()->GetIMFPPerDensity(e, m); break;
        case 8: v = ph.GetIMFPPairProd()->GetIMFPPerDensity(e, m); break;
        default: return 3;
      }
    } else {
      return 3;
    }
    std::printf("%.17g\n", v);
  }
  return 0;
}
