// SimGSTables.hh -- the reference's header name (Simulation/inc/SimGSTables.hh); the class is
// declared in SimTables.hh.
#ifndef DPM_AMD_SimGSTables_HH
#define DPM_AMD_SimGSTables_HH
#include "SimTables.hh"
#endif
