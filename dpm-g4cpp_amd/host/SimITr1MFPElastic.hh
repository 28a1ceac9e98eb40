// SimITr1MFPElastic.hh -- the reference's header name (Simulation/inc/SimITr1MFPElastic.hh); the class is
// declared in SimTables.hh.
#ifndef DPM_AMD_SimITr1MFPElastic_HH
#define DPM_AMD_SimITr1MFPElastic_HH
#include "SimTables.hh"
#endif
