// SimElectronData.hh -- the reference's header name (Simulation/inc/SimElectronData.hh) for code
// written against its data layer; the class is declared in SimData.hh.
#ifndef DPM_AMD_SimElectronData_HH
#define DPM_AMD_SimElectronData_HH
#include "SimData.hh"
#endif
