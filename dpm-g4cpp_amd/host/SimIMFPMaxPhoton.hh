// SimIMFPMaxPhoton.hh -- the reference's header name (Simulation/inc/SimIMFPMaxPhoton.hh); the class is
// declared in SimTables.hh.
#ifndef DPM_AMD_SimIMFPMaxPhoton_HH
#define DPM_AMD_SimIMFPMaxPhoton_HH
#include "SimTables.hh"
#endif
