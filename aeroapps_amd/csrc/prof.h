// Kernel classes timed by the optional HIP-event profiler (prof.hip).
#pragma once
#include <hip/hip_runtime.h>

namespace aa {
enum ProfKind {
  PROF_MIE = 0,     // Chem_Mie optics + tracer sum
  PROF_CHECK,       // input checks
  PROF_OPTICS,      // layer optics, delta-M, Greek
  PROF_GEOM,        // Chapman + first-order sun paths
  PROF_ITEM,        // per-column recursions
  PROF_LAYER,       // per-layer eigensystem + particular solutions
  PROF_BVP,         // band LU + back-substitution + source + convergence
  PROF_BRDF,        // BRDF supplement (MODIS land surface)
  PROF_SABDAB,      // FP32 SAB/DAB contraction of the mixed path (matrix cores)
  PROF_KINDS
};
bool prof_on();
void prof_begin(int kind, hipStream_t st);
void prof_end(int kind, hipStream_t st);
}  // namespace aa
