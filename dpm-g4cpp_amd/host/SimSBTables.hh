// SimSBTables.hh -- the reference's header name (Simulation/inc/SimSBTables.hh); the class is
// declared in SimTables.hh.
#ifndef DPM_AMD_SimSBTables_HH
#define DPM_AMD_SimSBTables_HH
#include "SimTables.hh"
#endif
