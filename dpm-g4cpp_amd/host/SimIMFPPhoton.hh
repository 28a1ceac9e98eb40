// SimIMFPPhoton.hh -- the reference's header name (Simulation/inc/SimIMFPPhoton.hh); the class is
// declared in SimTables.hh.
#ifndef DPM_AMD_SimIMFPPhoton_HH
#define DPM_AMD_SimIMFPPhoton_HH
#include "SimTables.hh"
#endif
