// SimIMFPMoller.hh -- the reference's header name (Simulation/inc/SimIMFPMoller.hh); the class is
// declared in SimTables.hh.
#ifndef DPM_AMD_SimIMFPMoller_HH
#define DPM_AMD_SimIMFPMoller_HH
#include "SimTables.hh"
#endif
