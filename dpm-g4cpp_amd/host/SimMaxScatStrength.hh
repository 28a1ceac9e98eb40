// SimMaxScatStrength.hh -- the reference's header name (Simulation/inc/SimMaxScatStrength.hh); the class is
// declared in SimTables.hh.
#ifndef DPM_AMD_SimMaxScatStrength_HH
#define DPM_AMD_SimMaxScatStrength_HH
#include "SimTables.hh"
#endif
