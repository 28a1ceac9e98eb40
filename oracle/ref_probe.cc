// ref_probe.cc -- TEST INFRASTRUCTURE ONLY.
//
// Small driver linked against the reference's own run-time data classes
// (Simulation/src/Sim*.cc, compiled where they lie under the reference tree by
// oracle/Makefile, target `ref`). It loads a table directory with the
// reference loaders and answers queries read from stdin, one per line:
//
//   sb  E imat r1 r2 r3     -> SimSBTables::SampleEnergyTransfer      (SimSBTables.cc:93)
//   mo  E r1 r2 r3          -> SimMollerTables::SampleEnergyTransfer  (SimMollerTables.cc:69)
//   kn  E r1 r2 r3          -> SimKNTables::SampleEnergyTransfer      (SimKNTables.cc:69)
//   gs  E r1 r2             -> SimGSTables::SampleAngularDeflection   (SimGSTables.cc:84)
//   ev  which E imat        -> the interpolation getters used by SimDPMLike.cc
//
// and prints each result as a C99 hex-float so that no digits are lost.
#include "SimElectronData.hh"
#include "SimPhotonData.hh"
#include "SimMaterialData.hh"
#include "SimITr1MFPElastic.hh"
#include "SimMaxScatStrength.hh"
#include "SimIMFPMoller.hh"
#include "SimIMFPBrem.hh"
#include "SimStoppingPower.hh"
#include "SimMollerTables.hh"
#include "SimSBTables.hh"
#include "SimGSTables.hh"
#include "SimIMFPMaxPhoton.hh"
#include "SimIMFPPhoton.hh"
#include "SimKNTables.hh"

#include <cstdio>
#include <cstring>
#include <string>

int main(int argc, char** argv) {
  if (argc < 2) { std::fprintf(stderr, "usage: ref_probe DATADIR < queries\n"); return 2; }
  SimElectronData el; el.Load(argv[1], 0);
  SimPhotonData ph;   ph.Load(argv[1], 0);
  SimMaterialData mat; mat.Load(argv[1], 0);
  char kind[16];
  while (std::scanf("%15s", kind) == 1) {
    double e, r1, r2, r3; int im, which;
    double v = 0.0;
    if (!std::strcmp(kind, "sb")) {
      if (std::scanf("%lg %d %lg %lg %lg", &e, &im, &r1, &r2, &r3) != 5) return 1;
      v = el.GetTheSBTables()->SampleEnergyTransfer(e, im, r1, r2, r3);
    } else if (!std::strcmp(kind, "mo")) {
      if (std::scanf("%lg %lg %lg %lg", &e, &r1, &r2, &r3) != 4) return 1;
      v = el.GetTheMollerTables()->SampleEnergyTransfer(e, r1, r2, r3);
    } else if (!std::strcmp(kind, "kn")) {
      if (std::scanf("%lg %lg %lg %lg", &e, &r1, &r2, &r3) != 4) return 1;
      v = ph.GetTheKNTables()->SampleEnergyTransfer(e, r1, r2, r3);
    } else if (!std::strcmp(kind, "gs")) {
      if (std::scanf("%lg %lg %lg", &e, &r1, &r2) != 3) return 1;
      v = el.GetTheGSTables()->SampleAngularDeflection(e, r1, r2);
    } else if (!std::strcmp(kind, "ev")) {
      if (std::scanf("%d %lg %d", &which, &e, &im) != 3) return 1;
      switch (which) {
        case 0: v = el.GetITr1MFPElastic()->GetITr1MFPPerDensity(e, im); break;
        case 1: v = el.GetDEDX()->GetDEDXPerDensity(e, im); break;
        case 2: v = el.GetIMFPBrem()->GetIMFPPerDensity(e, im); break;
        case 3: v = el.GetIMFPMoller()->GetIMFPPerDensity(e); break;
        case 4: v = el.GetMaxScatStrength()->GetMaxScatStrength(e); break;
        case 5: v = ph.GetIMFPTotalMax()->GetIMFP(e); break;
        case 6: v = ph.GetIMFPTotal()->GetIMFPPerDensity(e, im); break;
        case 7: v = ph.GetIMFPCompton()->GetIMFPPerDensity(e, im); break;
        case 8: v = ph.GetIMFPPairProd()->GetIMFPPerDensity(e, im); break;
        case 9: v = mat.fMaterialDensity[im]; break;
        case 10: v = mat.fMollerIMFPScaling[im]; break;
        default: return 1;
      }
    } else {
      return 1;
    }
    std::printf("%a\n", v);
  }
  return 0;
}
