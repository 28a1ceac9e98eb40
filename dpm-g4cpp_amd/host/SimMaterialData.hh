// SimMaterialData.hh -- the reference's header name (Simulation/inc/SimMaterialData.hh) for code
// written against its data layer; the class is declared in SimData.hh.
#ifndef DPM_AMD_SimMaterialData_HH
#define DPM_AMD_SimMaterialData_HH
#include "SimData.hh"
#endif
