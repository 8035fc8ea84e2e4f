// Shared host/device definitions of the MI355X aerosol-optics + radiance library.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <cstddef>

#define AA_CHECK(call)                                                                      \
  do {                                                                                      \
    hipError_t aa_err_ = (call);                                                            \
    if (aa_err_ != hipSuccess) {                                                            \
      std::fprintf(stderr, "aeroapps_amd: HIP error %s at %s:%d\n", hipGetErrorString(aa_err_), \
                   __FILE__, __LINE__);                                                     \
      return -2;                                                                            \
    }                                                                                       \
  } while (0)

namespace aa {

// Compiled maxima of the radiance kernels.
constexpr int MAX_LAYERS = 72;       // V283/vlidort_def/vlidort_pars.f90:140
constexpr int MAX_STREAMS = 20;        // streams per hemisphere, scalar; the workloads use 2..8
constexpr int MAX_STREAMS_VECTOR = 6;  // streams per hemisphere with three Stokes components (n = 18)
constexpr int MAX_MOMENTS_IN = 1000; // V283/vlidort_def/vlidort_pars.f90:149
constexpr int MAX_OBSGEOMS = 2;      // MAX_SZANGLES / MAX_USER_OBSGEOMS, V283/vlidort_def/vlidort_pars.f90:162,177
constexpr int NRH_MAP = 991;         // CB/Chem_MieTableMod.F90:81-82

// VLIDORT constants (V283/vlidort_def/vlidort_pars.f90:393-444) and the AeroApps
// fixed options (V283/VLIDORT_Mod.F90:200-309).
constexpr double DEG_TO_RAD = 1.7453292519943e-02;
constexpr double PIE = 180.0 * DEG_TO_RAD;
constexpr double PI4 = 4.0 * PIE;
constexpr double TAYLOR_SMALL = 1.0e-04;
constexpr double MAX_TAU_PATH = 32.0;
constexpr double OMEGA_SMALLNUM = 1.0e-15;
constexpr int TAYLOR_ORDER = 3;
constexpr int NFINE = 3;
constexpr double EARTH_RADIUS = 6371.0;
constexpr double ASYMTX_TOL = 1.0e-12;

// Device-resident Chem_Mie table set.  Tracer-major arrays (all tables'
// bins concatenated):
//   bext/bsca/g/bbck/refr/refi: [tracer][ch][rh]          (float)
//   pmom:                      [tracer][ch][rh][pol][mom] (float)
//   rhi/rha:                   [tracer][991]               (int, float)
struct MieDev {
  int ntracer = 0, nch = 0, nrh = 0, nmom = 0, npol = 0;
  const float* bext = nullptr;
  const float* bsca = nullptr;
  const float* g = nullptr;
  const float* bbck = nullptr;
  const float* refr = nullptr;
  const float* refi = nullptr;
  const float* pmom = nullptr;
  const int* rhi = nullptr;
  const float* rha = nullptr;
  const int* nrh_t = nullptr;     // per-tracer RH node count
  const int* nmom_t = nullptr;    // per-tracer moments in the source table
  const int* npol_t = nullptr;    // per-tracer polarizations in the source table
};

}  // namespace aa
