// SimStoppingPower.hh -- the reference's header name (Simulation/inc/SimStoppingPower.hh); the class is
// declared in SimTables.hh.
#ifndef DPM_AMD_SimStoppingPower_HH
#define DPM_AMD_SimStoppingPower_HH
#include "SimTables.hh"
#endif
