// SimPhotonData.hh -- the reference's header name (Simulation/inc/SimPhotonData.hh) for code
// written against its data layer; the class is declared in SimData.hh.
#ifndef DPM_AMD_SimPhotonData_HH
#define DPM_AMD_SimPhotonData_HH
#include "SimData.hh"
#endif
