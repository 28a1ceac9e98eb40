// SimKNTables.hh -- the reference's header name (Simulation/inc/SimKNTables.hh); the class is
// declared in SimTables.hh.
#ifndef DPM_AMD_SimKNTables_HH
#define DPM_AMD_SimKNTables_HH
#include "SimTables.hh"
#endif
