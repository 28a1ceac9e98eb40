// SimMollerTables.hh -- the reference's header name (Simulation/inc/SimMollerTables.hh); the class is
// declared in SimTables.hh.
#ifndef DPM_AMD_SimMollerTables_HH
#define DPM_AMD_SimMollerTables_HH
#include "SimTables.hh"
#endif
