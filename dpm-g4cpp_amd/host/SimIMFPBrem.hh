// SimIMFPBrem.hh -- the reference's header name (Simulation/inc/SimIMFPBrem.hh); the class is
// declared in SimTables.hh.
#ifndef DPM_AMD_SimIMFPBrem_HH
#define DPM_AMD_SimIMFPBrem_HH
#include "SimTables.hh"
#endif
