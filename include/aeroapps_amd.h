/* C-ABI of the MI355X aerosol-optics + discrete-ordinate radiance library.
 *
 * Every entry point takes plain pointers and sizes.  Host-pointer entry points
 * ("drop-in" surface) accept arrays in the reference's Fortran column-major
 * layout and are synchronous, like the Fortran routines they replace.  The
 * "_dev" entry points take device pointers in the same layouts and enqueue on a
 * caller stream (a hipStream_t passed as void*); they return once the work is
 * enqueued.
 *
 * Return value of every function: 0 on success, otherwise an AA_E* code below
 * (an argument/setup error detected before any kernel ran).  Per-element
 * failures follow the reference convention instead: the element's outputs are
 * set to MISSING and the call still returns 0.
 *
 * Reference paths are relative to the GEOS-ESM/AeroApps source tree:
 *   CB   = src/Shared/GMAO_Shared/Chem_Base
 *   GEO  = src/Components/rtms/geo_vlidort
 *   V283 = src/Components/rtms/RTSI/VLIDORT2p8p3
 */
#ifndef AEROAPPS_AMD_H
#define AEROAPPS_AMD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  AA_OK = 0,
  AA_EARG = -1,        /* bad size / null pointer / unsupported option */
  AA_EHIP = -2,        /* HIP runtime error (allocation, launch) */
  AA_ENOTABLE = -3,    /* table handle has no species table */
  AA_EDIMS = -4,       /* dimension beyond the compiled maxima (streams: 20 per hemisphere scalar,
                          6 with three Stokes components; layers; moments) */
  AA_EUNSUP = -5,      /* a reference option this library does not implement (DO_BOA) */
  AA_ENOTRES = -6,     /* optics tables not yet resident on the device (see the _noalloc entry) */
  AA_ENOFILE = -7      /* a file cannot be read */
};

/* Reference return codes reproduced in the rc out-arguments. */
#define AA_RC_CHANNEL 99    /* GEO/VLIDORT_MieMod.F90:225  channel not in the Mie tables       */
#define AA_RC_NMOM 666      /* GEO/VLIDORT_MieMod.F90:260  table has fewer moments than asked   */
#define AA_RC_NPOL 777      /* GEO/VLIDORT_MieMod.F90:276  inconsistent number of polarizations */

/* ------------------------------------------------------------------ Chem_Mie
 * Device-resident optics tables.  Replaces the Chem_Mie object of
 * CB/Chem_MieMod.F90:65-99 (type Chem_Mie) built by Chem_MieCreate
 * (CB/Chem_MieMod.F90:124-411) from Chem_MieTableRead tables
 * (CB/Chem_MieTableMod.F90:238-719).  The tables are uploaded once and stay
 * resident until aa_mie_destroy. */
typedef struct aa_mie aa_mie;

/* Empty table set for nch channels given in metres (Chem_Mie%channels).
 * Replaces Chem_MieCreate (CB/Chem_MieMod.F90:124). */
int aa_mie_create(aa_mie** out, int nch, const float* channels_m);

/* Adds one species table already interpolated to the nch channels (the result
 * of Chem_MieTableRead's polint step, CB/Chem_MieTableMod.F90:560-589).  Arrays
 * are Fortran-ordered float32: rh(nrh), bext/bsca/bbck/g/refr/refi(nrh,nch,nbin),
 * pmom(nrh,nch,nbin,nmom,npol).  bin_names is nbin names of 16 characters
 * (blank padded), the tracer names served by this table (Chem_Mie%vname).
 * The 991-slot RH hash (CB/Chem_MieTableMod.F90:612-627) is built here. */
int aa_mie_add_table(aa_mie* t, int nrh, int nbin, int nmom, int npol, const float* rh,
                     const float* bext, const float* bsca, const float* bbck, const float* g,
                     const float* refr, const float* refi, const float* pmom, const char* bin_names);

/* Adds one species table from the arrays of its optics-table file, doing
 * what Chem_MieTableRead (CB/Chem_MieTableMod.F90:238-719) does after the
 * netCDF reads: fill values of absent variables (pback 1, growth_factor /
 * rhop / refreal / refimag -999; refimag made non-negative; rhod taken from
 * the first RH row of 'rhop'), wet volume gf^3/rhod and area
 * vol/(4/3 reff), the polint linear interpolation (float32 slope and offset)
 * from the file wavelengths lambda_table [m] to the table set's channels, and
 * the RH map.  File layouts (Fortran, float64): rh(nrh), reff/gf/rhop(nrh,nbin),
 * bext/bsca/bbck/g/refr/refi(nch_table,nrh,nbin),
 * pback(nch_table,nrh,nbin,npol_table), pmom(nch_table,nrh,nbin,nmom_table,
 * npol_table); optional ones may be NULL (pmom only when nmom = 0).  nmom is
 * the number of moments kept; rc = 99 when it exceeds nmom_table (as the
 * reference, nothing is added). */
int aa_mie_add_table_native(aa_mie* t, int nch_table, const double* lambda_table, int nrh, int nbin,
                            int nmom_table, int npol_table, int nmom, const double* rh, const double* reff,
                            const double* bext, const double* bsca, const double* bbck, const double* g,
                            const double* pback, const double* pmom, const double* gf, const double* rhop,
                            const double* refr, const double* refi, const char* bin_names, int* rc);

/* The Chem_MieQuery outputs besides the optics (CB/Chem_MieMod.F90:885-943)
 * for n relative humidities: rEff [um], p11, p22, growth factor, dry
 * density, wet volume and area (any may be NULL), for tables added with
 * aa_mie_add_table_native.  Host pointers, synchronous, no device work. */
int aa_mie_query_extra(const aa_mie* t, int idx, float channel, int n, const float* rh, float* reff, float* p11,
                       float* p22, float* gf, float* rhod, float* vol, float* area);

/* Tracer index (0-based) of qname, or -1.  Replaces Chem_MieQueryIdx
 * (CB/Chem_MieMod.F90:720-777): case-insensitive, "XX::" qualifier stripped. */
int aa_mie_query_idx(const aa_mie* t, const char* qname);

/* Chem_MieQuery over n (q_mass, rh) pairs of one tracer and channel
 * (CB/Chem_MieMod.F90:789-965).  channel is the 1-based table channel number
 * (Chem_MieQuery's real CHANNEL argument, rounded with nint).  Outputs may be
 * NULL when not wanted; pmom is (n, nmom, npol) Fortran order and only filled
 * for nmom <= the table's moments.  Host pointers, synchronous. */
int aa_mie_query(const aa_mie* t, int idx, float channel, int n, const float* q_mass, const float* rh,
                 float* tau, float* ssa, float* gasym, float* bext, float* bsca, float* bbck,
                 float* refr, float* refi, int nmom, int npol, float* pmom);

/* Number of channels / moments / polarizations of the table set. */
int aa_mie_dims(const aa_mie* t, int* nch, int* nmom_min, int* npol_max, int* ntracer);

/* The nch channels [m] of the table set (Chem_Mie%channels). */
int aa_mie_channels(const aa_mie* t, float* channels_m);

/* Table of tracer idx (0-based, tables numbered in the order they were
 * added), its 0-based bin in that table and the table's bins, RH nodes,
 * moments and polarizations: what Chem_MieQueryIdx's vindex and vtableUse
 * (CB/Chem_MieMod.F90:720-777) hold for a tracer.  Any output may be NULL. */
int aa_mie_tracer_table(const aa_mie* t, int idx, int* table, int* bin, int* nbin, int* nrh, int* nmom,
                        int* npol);

/* ---- Construction from resource files (Chem_MieCreateFromRC,
 * CB/Chem_MieMod.F90:124-411).  Species codes of the seven optics tables. */
enum { AA_SP_DU = 0, AA_SP_SS = 1, AA_SP_SU = 2, AA_SP_OC = 3, AA_SP_BC = 4, AA_SP_NI = 5, AA_SP_BRC = 6 };

/* The arrays of one optics-table file, in the layouts of
 * aa_mie_add_table_native (float64, Fortran order); a reader fills them and
 * keeps them valid until it is called again or aa_mie_create_rc returns. */
typedef struct aa_mie_table_file {
  int nch, nrh, nbin, nmom, npol;     /* file dimensions (nmom = npol = 0 without moments) */
  const double *lambda, *rh, *reff, *bext, *bsca, *bbck, *g, *pback, *pmom, *gf, *rhop, *refr, *refi;
} aa_mie_table_file;
/* Reads the optics table `path` (the file named in the resource file); 0 on
 * success.  The reference reads these netCDF files with Chem_MieTableRead
 * (CB/Chem_MieTableMod.F90:238-719); here the host application supplies the
 * reader. */
typedef int (*aa_mie_table_reader)(void* user, const char* path, aa_mie_table_file* table);

/* The chemistry registry file (Chem_RegistryCreate,
 * CB/Chem_RegistryMod.F90:200-545, i90 format): the tracers of every species
 * with doing_XX: yes, in the registry's order, named from variable_table_XX::
 * or by prefix and bin number (du001 ...).  nq receives their number; names
 * (16 characters each, blank padded) and species (AA_SP_* code, -1 for a
 * species without an optics table) are filled when nq <= nq_max.  rc: 0; 1
 * the file cannot be read; 2 a species is on without nbins or units; 3 its name table
 * has too few rows (the reference stops in each case). */
int aa_mie_registry(const char* path, int nq_max, int* nq, char* names, int* species, int* rc);

/* Chem_MieCreate(rcfile, rc, chemReg): reads n_channels, n_moments,
 * r_channels [m] and filename_optical_properties_{DU,SS,BC,OC,SU,NI,BRC} from
 * rcfile, reads the seven tables with `reader` in the reference's order (DU,
 * SS, SU, OC, BC, NI, BRC) and adds each with aa_mie_add_table_native (nmom =
 * n_moments), its bins served to the registry tracers (names/species as from
 * aa_mie_registry) of that species in order (the sulfate table to
 * so4/SO4/sul/SUL names only, the nitrate table to no3/NO3 only,
 * CB/Chem_MieMod.F90:338-400).  rc: 0; 1 unreadable rc file or a missing
 * label; 2 the reader failed; otherwise aa_mie_add_table_native's rc.  *out
 * is NULL unless rc = 0 (the reference stops with die() instead). */
int aa_mie_create_rc(aa_mie** out, const char* rcfile, int nq, const char* names, const int* species,
                     aa_mie_table_reader reader, void* user, int* rc);

int aa_mie_destroy(aa_mie* t);

/* Replaces VLIDORT_getAOPvector (GEO/VLIDORT_MieMod.F90:124-336).  Fortran
 * arrays: channels(nch) [nm], vname(nq,16) (character array as passed by f2py,
 * i.e. vname[c*nq + iq]), qm(km,nq,nobs), rh(km,nobs) -> tau,ssa,g(km,nch,nobs),
 * pmom(km,nch,nobs,nmom,npol); rc as in the reference (0, 1 when the last
 * tracer name was not found, 99, 666, 777).  Host pointers, synchronous. */
int aa_get_aop_vector(aa_mie* t, int km, int nobs, int nch, int nq, const float* channels,
                      const char* vname, int verbose, const float* qm, const float* rh, int nmom, int npol,
                      float* tau, float* ssa, float* g, float* pmom, int* rc);

/* Replaces VLIDORT_getAOPscalar (GEO/VLIDORT_MieMod.F90:10-120). */
int aa_get_aop_scalar(aa_mie* t, int km, int nobs, int nch, int nq, const float* channels,
                      const char* vname, int verbose, const float* qm, const float* rh,
                      float* tau, float* ssa, float* g, int* rc);

/* Device-pointer form of aa_get_aop_vector: same layouts, tracer indices
 * already resolved (tracer_idx[nq], -1 = skipped) and channel indices
 * resolved (chan_idx[nch], 0-based), both DEVICE arrays; enqueued on stream. */
int aa_get_aop_vector_dev(aa_mie* t, int km, int nobs, int nch, int nq, const int* tracer_idx,
                          const int* chan_idx, const float* d_qm, const float* d_rh, int nmom, int npol,
                          float* d_tau, float* d_ssa, float* d_g, float* d_pmom, void* stream);

/* ------------------------------------------------------------------ VLIDORT drivers
 * Replace VLIDORT_Scalar_Lambert (GEO/VLIDORT_LAMBERT.F90:34-168) and
 * VLIDORT_Vector_Lambert single-geometry (GEO/VLIDORT_LAMBERT.F90:172-342),
 * which run VLIDORT_Run_Scalar / VLIDORT_Run_Vector (V283/VLIDORT_ScatMod.F90:
 * 975-1247 / 284-617) over columns x channels with the fixed options of
 * VLIDORT_Init (V283/VLIDORT_Mod.F90:65-383).  Fortran arrays:
 *   channels(nch), ROT(km,nobs,nch), depol(nch), [alpha(km,nobs,nch)],
 *   tau,ssa,[g](km,nch,nobs), pmom(km,nch,nobs,nMom,nPol), pe,he,te(km+1,nobs),
 *   albedo(nobs,nch), solar_zenith/relat_azymuth/sensor_zenith(nobs),
 *   [flux_factor(nch,nobs)] -> radiance,reflectance,[Q,U](nobs,nch).
 * nfourier (may be NULL) receives the number of Fourier terms used per
 * (obs, channel), 0 where MISSING.  precision: 0 = FP64 throughout (the
 * drop-in), 1 = mixed (FP32 Chem_Mie optics, FP32 phase-matrix (SAB/DAB)
 * contraction on the matrix cores and, from nstreams = 8 on, FP32 source
 * assembly: the user-angle homogeneous solutions and the phase-matrix sums of
 * the beam source; FP64 eigensystem, normalisation, particular-solution
 * coefficients and boundary-value solve).  Host pointers, synchronous. */
int aa_vlidort_scalar_lambert(int km, int nch, int nobs, const double* channels, int nstreams,
                              int plane_parallel, int nmom, int npol, const double* rot, const double* depol,
                              const double* tau, const double* ssa, const double* g, const double* pmom,
                              const double* pe, const double* he, const double* te, const double* albedo,
                              const double* solar_zenith, const double* relat_azymuth,
                              const double* sensor_zenith, double missing, int verbose, int precision,
                              double* radiance, double* reflectance, int* nfourier, int* rc);

int aa_vlidort_vector_lambert(int km, int nch, int nobs, const double* channels, int nstreams,
                              int plane_parallel, int nmom, int npol, const double* rot, const double* depol,
                              const double* alpha, const double* tau, const double* ssa, const double* pmom,
                              const double* pe, const double* he, const double* te, const double* albedo,
                              const double* solar_zenith, const double* relat_azymuth,
                              const double* sensor_zenith, const double* flux_factor, double missing,
                              int verbose, int precision, double* radiance, double* reflectance, double* q,
                              double* u, int* nfourier, int* rc);

/* VLIDORT_Vector_Lambert_SingleGeom with its optional arguments
 * (GEO/VLIDORT_LAMBERT.F90:172-342): do_2os_correction != 0 runs the vector
 * driver with NSTOKES = 1 (what DO_2OS_CORRECTION = .true. selects with the
 * VLIDORT 2.8.3 library, GEO/VLIDORT_LAMBERT.F90:256-265): the Stokes-1
 * radiance of the vector driver's inputs (gas absorption, flux factor), Q and
 * U = 0.  do_boa != 0 (downwelling output at user level 1) is not implemented
 * and returns AA_EUNSUP.  aa_vlidort_vector_lambert is this call with both 0. */
int aa_vlidort_vector_lambert_opt(int km, int nch, int nobs, const double* channels, int nstreams,
                                  int plane_parallel, int nmom, int npol, const double* rot, const double* depol,
                                  const double* alpha, const double* tau, const double* ssa, const double* pmom,
                                  const double* pe, const double* he, const double* te, const double* albedo,
                                  const double* solar_zenith, const double* relat_azymuth,
                                  const double* sensor_zenith, const double* flux_factor, double missing,
                                  int verbose, int precision, int do_2os_correction, int do_boa, double* radiance,
                                  double* reflectance, double* q, double* u, int* nfourier, int* rc);

/* Replaces VLIDORT_Vector_Lambert_MultiGeom (GEO/VLIDORT_LAMBERT.F90:346-520,
 * called by missions/ACCP/accp_polar_vlidort.F90:988): ngeom geometries per
 * observation, solar_zenith/relat_azymuth/sensor_zenith(nobs,ngeom),
 * albedo(nobs,nch,ngeom) -> radiance,reflectance,Q,U,nfourier(nobs,nch,ngeom).
 * As the reference: the missing-geometry and missing-albedo checks use
 * geometry 1, the surface albedo is albedo(:,:,1) for every geometry, an
 * overhead sun in any geometry switches first-order nadir mode on for that
 * and every later observation, and a failed element fails all its
 * geometries.  ngeom above the reference's MAX_SZANGLES (2) gives rc = 6. */
int aa_vlidort_vector_lambert_multigeom(int km, int nch, int nobs, int ngeom, const double* channels, int nstreams,
                                        int plane_parallel, int nmom, int npol, const double* rot,
                                        const double* depol, const double* alpha, const double* tau,
                                        const double* ssa, const double* pmom, const double* pe, const double* he,
                                        const double* te, const double* albedo, const double* solar_zenith,
                                        const double* relat_azymuth, const double* sensor_zenith,
                                        const double* flux_factor, double missing, int verbose, int precision,
                                        int do_2os_correction, int do_boa, double* radiance, double* reflectance,
                                        double* q, double* u, int* nfourier, int* rc);

/* ------------------------------------------------------------------ MODIS land surface (BRDF)
 * Replaces VLIDORT_Scalar_LandMODIS (GEO/VLIDORT_BRDF_MODIS.F90:35-202): per
 * (obs, channel) the BRDF supplement of VLIDORT_LANDMODIS_SingleGeom
 * (V283/VLIDORT_SurfaceMod.F90:1376-1596; kernels Lambertian, Ross-thick,
 * Li-sparse with weights kernel_wt(3,nch,nobs) = (fiso, fgeo, fvol) and
 * Li-sparse parameters param(2,nch,nobs) = (h/b, b/r), 50 azimuth streams,
 * white-sky-albedo check) followed by VLIDORT_Run_Scalar with that BRDF.
 * Same arrays as aa_vlidort_scalar_lambert without albedo; BR(nobs,nch)
 * receives the exact direct-bounce BRDF.  As the reference: BR < 0 gives
 * radiance = reflectance = BR = -500 (VLIDORT not run), a missing geometry
 * gives MISSING (BR = MISSING where the reference leaves it undefined), a
 * MISSING kernel weight is not a skip (the reference's check only leaves its
 * inner loop), nMom above 1000 gives rc = 5 with outputs untouched.  Scalar
 * only (the reference's vector MODIS driver stops with rc = 9). */
int aa_vlidort_scalar_landmodis(int km, int nch, int nobs, const double* channels, int nstreams,
                                int plane_parallel, int nmom, int npol, const double* rot, const double* depol,
                                const double* tau, const double* ssa, const double* g, const double* pmom,
                                const double* pe, const double* he, const double* te, const double* kernel_wt,
                                const double* param, const double* solar_zenith, const double* relat_azymuth,
                                const double* sensor_zenith, double missing, int verbose, int precision,
                                double* radiance, double* reflectance, double* br, int* nfourier, int* rc);

/* The BRDF supplement alone (VBRDF_MAINMASTER under the
 * VLIDORT_LANDMODIS_SingleGeom set-up, V283/vsup/vbrdf/vbrdf_sup_masters.f90:
 * 1559-3293) for ngeom observation geometries (degrees) with their own
 * kernel_wt(3,ngeom) and param(2,ngeom): the Fourier components the solver
 * reads, in Fortran order per geometry g, brdf_f(0:2N-1,N,N,g) (incident
 * stream j -> reflected i at (M,i,j)), brdf_f0(0:2N-1,N,g),
 * user_brdf_f(0:2N-1,N,g), user_brdf_f0(0:2N-1,g), the exact direct bounce
 * dbounce(g) and status(g) (4 = white-sky albedo outside (0,1]: the Li-sparse
 * Fourier terms are left out, as the reference's supplement stops there).
 * Host pointers, synchronous. */
int aa_vbrdf_modis(int nstreams, int ngeom, const double* sza, const double* vza, const double* raa,
                   const double* kernel_wt, const double* param, double* brdf_f, double* brdf_f0,
                   double* user_brdf_f, double* user_brdf_f0, double* dbounce, int* status);

/* ------------------------------------------------------------------ Cox-Munk ocean surface (BRDF)
 * Replaces VLIDORT_Scalar_CX (GEO/VLIDORT_BRDF_CX.F90:420-586): per (obs,
 * channel) the supplement of VLIDORT_CoxMunk_SingleGeom
 * (V283/VLIDORT_SurfaceMod.F90:815-1042: one Cox-Munk glint kernel with
 * shadowing, sigma^2 = 0.003 + 0.00512 |(u10m, v10m)| capped at 50 m/s,
 * refractive index mr, 100 azimuth streams, white-sky check) followed by
 * VLIDORT_Run_Scalar with that BRDF.  Arrays as aa_vlidort_scalar_lambert
 * without albedo, plus u10m(nobs), v10m(nobs) [m/s], mr(nch); brdf(1,nobs,nch)
 * receives the exact direct-bounce BRDF (MISSING where the reference leaves it
 * undefined: missing geometry, wind or refractive index, which give MISSING
 * radiances). */
int aa_vlidort_scalar_cx(int km, int nch, int nobs, const double* channels, int nstreams, int plane_parallel,
                         int nmom, int npol, const double* rot, const double* depol, const double* tau,
                         const double* ssa, const double* g, const double* pmom, const double* pe,
                         const double* he, const double* te, const double* u10m, const double* v10m,
                         const double* mr, const double* solar_zenith, const double* relat_azymuth,
                         const double* sensor_zenith, double missing, int verbose, int precision,
                         double* radiance, double* reflectance, double* brdf, int* nfourier, int* rc);

/* Replaces VLIDORT_Vector_CX_SingleGeom (GEO/VLIDORT_BRDF_CX.F90:587-762; f2py
 * VLIDORT_POLAR_.vector_cx, src/Components/rtms/leo_vlidort/
 * VLIDORT_POLAR_py.F90:830-900): the same surface with NSTOKES = 3 and
 * VLIDORT_Run_Vector, the flux factor VLIDORT_Init's 1.  Arrays as
 * aa_vlidort_scalar_cx without g, pmom(km,nch,nobs,nMom,nPol) with nPol >= 6;
 * q, u (nobs,nch) receive the Stokes components, brdf the exact direct-bounce
 * BRDF, brdf_q and brdf_u 0 (the Cox-Munk kernel reflects I only).  Outputs of
 * elements the reference skips (missing geometry, wind or refractive index)
 * are MISSING. */
int aa_vlidort_vector_cx(int km, int nch, int nobs, const double* channels, int nstreams, int plane_parallel,
                         int nmom, int npol, const double* rot, const double* depol, const double* tau,
                         const double* ssa, const double* pmom, const double* pe, const double* he,
                         const double* te, const double* u10m, const double* v10m, const double* mr,
                         const double* solar_zenith, const double* relat_azymuth, const double* sensor_zenith,
                         double missing, int verbose, int precision, double* radiance, double* reflectance,
                         double* q, double* u, double* brdf, double* brdf_q, double* brdf_u, int* nfourier,
                         int* rc);

/* Replace the polarised land drivers VLIDORT_Vector_Lambert_BPDF
 * (GEO/VLIDORT_LAMBERT_BPDF.F90:23-222: Lambertian kernel of factor
 * albedo(nobs,nch), then BPDF-NDVI) and VLIDORT_Vector_LandMODIS_BPDF
 * (GEO/VLIDORT_BRDF_MODIS_BPDF.F90:24-222: Lambertian fiso, Ross-thick fvol,
 * Li-sparse fgeo with kernel_wt(3,nch,nobs) = (fiso, fgeo, fvol) and
 * rtls_param(2,nch,nobs) = (h/b, b/r), then BPDF-NDVI), surface set-ups
 * VLIDORT_LAMBERTIAN_BPDF / VLIDORT_LANDMODIS_BPDF (V283/VLIDORT_SurfaceMod.F90:
 * 1822-2200): NSTOKES = 3, 50 full-range BRDF azimuths, the white-sky check
 * dropping the BPDF kernel's Fourier terms; VLIDORT_Run_Vector with the flux
 * factor of VLIDORT_Init (1).  bpdf_param(3,nch,nobs) = (refractive index,
 * NDVI, scaling factor C).  Outputs (nobs,nch): radiance, reflectance, Q, U,
 * br, br_q, br_u = the direct-bounce BRDF elements (1,1), (1,2), (1,3);
 * all -500 where the (1,1) direct bounce is negative (the element is not
 * run).  Missing geometry (and, for the Lambertian form, a missing albedo)
 * gives MISSING; missing kernel weights or BPDF parameters are used as given,
 * as in the reference (its checks leave only their own inner loops).
 * nstreams <= 6.  Host pointers, synchronous. */
int aa_vlidort_vector_lambert_bpdf(int km, int nch, int nobs, const double* channels, int nstreams,
                                   int plane_parallel, int nmom, int npol, const double* rot, const double* depol,
                                   const double* tau, const double* ssa, const double* pmom, const double* pe,
                                   const double* he, const double* te, const double* albedo, const double* bpdf_param,
                                   const double* solar_zenith, const double* relat_azymuth,
                                   const double* sensor_zenith, double missing, int verbose, int precision,
                                   double* radiance, double* reflectance, double* br, double* q, double* u,
                                   double* br_q, double* br_u, int* nfourier, int* rc);

int aa_vlidort_vector_landmodis_bpdf(int km, int nch, int nobs, const double* channels, int nstreams,
                                     int plane_parallel, int nmom, int npol, const double* rot, const double* depol,
                                     const double* tau, const double* ssa, const double* pmom, const double* pe,
                                     const double* he, const double* te, const double* kernel_wt,
                                     const double* rtls_param, const double* bpdf_param, const double* solar_zenith,
                                     const double* relat_azymuth, const double* sensor_zenith, double missing,
                                     int verbose, int precision, double* radiance, double* reflectance, double* br,
                                     double* q, double* u, double* br_q, double* br_u, int* nfourier, int* rc);

/* The polarised supplement alone for ngeom geometries (degrees): the BRDF of
 * VLIDORT_LAMBERTIAN_BPDF (kernel_wt NULL: Lambertian factor lamb(g), BPDF)
 * or VLIDORT_LANDMODIS_BPDF (kernel_wt(3,g), rtls_param(2,g): MODIS kernels,
 * BPDF), bpdf_param(3,g), as VBRDF_MAINMASTER returns it
 * (VB/vbrdf_sup_masters.f90:1559-3293) with NSTOKES = 3, in the arrays
 * BS_BRDF_F(0:2N-1, 16, N, N), BS_BRDF_F_0(0:2N-1, 16, N) (first Mueller
 * column), BS_USER_BRDF_F(0:2N-1, 16, N) and BS_DBOUNCE_BRDFUNC(16) per
 * geometry, Mueller index q = 4 (o1 - 1) + o2 (entries of the fourth Stokes
 * component 0).  status(g): 0, 4 (white-sky albedo outside (0, 1]: the BPDF
 * kernel's Fourier terms left out), -1 (negative (1,1) direct bounce:
 * Fourier components not computed, as the drivers skip the element).
 * nstreams <= 6. */
int aa_vbrdf_bpdf(int nstreams, int ngeom, const double* sza, const double* vza, const double* raa,
                  const double* lamb, const double* kernel_wt, const double* rtls_param, const double* bpdf_param,
                  double* brdf_f, double* brdf_f0, double* user_brdf_f, double* dbounce, int* status);

/* The Cox-Munk supplement alone, as aa_vbrdf_modis with u10m, v10m, mr per
 * geometry (status 4: white-sky albedo outside (0,1], the kernel's Fourier
 * terms are then zero). */
int aa_vbrdf_coxmunk(int nstreams, int ngeom, const double* sza, const double* vza, const double* raa,
                     const double* u10m, const double* v10m, const double* mr, double* brdf_f, double* brdf_f0,
                     double* user_brdf_f, double* user_brdf_f0, double* dbounce, int* status);

/* ------------------------------------------------------------------ input preparation
 * The steps either side of the radiance path in a geo_vlidort run
 * (GEO/geo_vlidort_AI.F90:399-430): met fields -> edge variables and tracer
 * column masses -> (optics) -> Rayleigh optical thickness -> radiance. */

/* VLIDORT_ROT_CALC (GEO/VLIDORT_ROT.F90:19-85; f2py VLIDORT_POLAR_.rot_calc,
 * src/Components/rtms/leo_vlidort/VLIDORT_POLAR_py.F90:9-41): Rayleigh optical
 * thickness rot(km,nobs,nch) of each layer and the depolarisation ratio
 * depol(nch) from edge pressure pe [Pa], height he [m] and temperature te [K]
 * (km+1,nobs), channels [nm] (VLIDORT_Rayleigh, coef_depol, F_air,
 * V283/VLIDORT_ScatMod.F90:143-281).  rc = 3 (and AA_EDIMS) when km exceeds
 * the layer dimension, as the drivers' VLIDORT_Init. */
int aa_rot_calc(int km, int nch, int nobs, const double* channels, const double* pe, const double* he,
                const double* te, double missing, int verbose, double* rot, double* depol, int* rc);
/* Device form: d_pe/d_he/d_te/d_rot device pointers, channels and depol host
 * arrays (nch), enqueued on stream. */
int aa_rot_calc_dev(int km, int nch, int nobs, const double* channels, const double* d_pe, const double* d_he,
                    const double* d_te, double* d_rot, double* depol, void* stream);

/* getEdgeVars (CB/MieObs_py.F90:27-83; f2py MieObs_.getedgevars behind
 * mieobs.getEdgeVars, src/Components/misc/scat/mieobs.py:60-76): edge
 * pressure pe [Pa], height above the surface ze [m] and temperature te [K]
 * (km+1,nobs) from air density airdens and pressure thickness delp (km,nobs),
 * top pressure ptop; float32 as the reference. */
int aa_edge_vars(int km, int nobs, const float* airdens, const float* delp, float ptop, float* pe, float* ze,
                 float* te);
int aa_edge_vars_dev(int km, int nobs, const float* d_airdens, const float* d_delp, float ptop, float* d_pe,
                     float* d_ze, float* d_te, void* stream);

/* calc_qm (GEO/geo_vlidort.F90:2124-2144): tracer mass per layer
 * qm(km,nq,nobs) = q(km,nq,nobs) * delp(km,nobs) / 9.81, float32. */
int aa_calc_qm(int km, int nq, int nobs, const float* q, const float* delp, float* qm);
int aa_calc_qm_dev(int km, int nq, int nobs, const float* d_q, const float* d_delp, float* d_qm, void* stream);

/* ------------------------------------------------------------------ fused device path
 * The whole hot path on resident data: Mie optics (tables in t) -> layer
 * optics -> VLIDORT radiance, for nobs columns x nch channels, device pointers
 * in the Fortran layouts above (qm(km,nq,nobs), rh(km,nobs) float32; rot,
 * pe/he/te, albedo, angles, flux_factor float64).  nlayers (may be NULL,
 * int[nobs]) gives a per-column layer count <= km (terrain truncation: the
 * column is layers 1..nlayers counted from the top, its surface is edge
 * nlayers+1); nstreams_col (may be NULL, int[nobs]) a per-column stream count
 * (overrides nstreams).  vector
 * selects NSTOKES=3.  Outputs radiance/reflectance/[q,u](nobs,nch) float64,
 * nfourier(nobs,nch) int32.  Work buffers come from the library's device pool;
 * enqueued on stream. */
int aa_column_radiance_dev(aa_mie* t, int km, int nobs, int nch, int nq, const int* tracer_idx,
                           const int* chan_idx, const float* d_qm, const float* d_rh, int nmom, int npol,
                           int nstreams, int vector, int precision, const int* d_nlayers,
                           const int* d_nstreams_col, const double* d_rot, const double* d_depol,
                           const double* d_pe, const double* d_he, const double* d_te,
                           const double* d_albedo, const double* d_sza, const double* d_raa,
                           const double* d_vza, const double* d_flux_factor, double missing,
                           double* d_radiance, double* d_reflectance, double* d_q, double* d_u,
                           int* d_nfourier, void* stream);

/* The fused device path over a MODIS land surface (the BRDF of
 * aa_vlidort_scalar_landmodis computed on the device per item): d_kernel_wt
 * (3,nch,nobs), d_param (2,nch,nobs) in place of albedo, d_br(nobs,nch) out.
 * Scalar. */
int aa_column_radiance_modis_dev(aa_mie* t, int km, int nobs, int nch, int nq, const int* tracer_idx,
                                 const int* chan_idx, const float* d_qm, const float* d_rh, int nmom,
                                 int nstreams, int precision, const int* d_nlayers, const int* d_nstreams_col,
                                 const double* d_rot, const double* d_depol, const double* d_pe,
                                 const double* d_he, const double* d_te, const double* d_kernel_wt,
                                 const double* d_param, const double* d_sza, const double* d_raa,
                                 const double* d_vza, double missing, double* d_radiance, double* d_reflectance,
                                 double* d_br, int* d_nfourier, void* stream);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): aa_column_radiance_dev with one stream count for every
 * column (no d_nstreams_col), whose work memory is a fixed 512 MiB device
 * array inside the library (one per GPU, held until the process ends) instead
 * of the allocated arenas.  It calls no hipMalloc/hipFree: the optics tables
 * must already be resident on the device (used by an earlier call after their
 * last change), else AA_ENOTRES; a shape whose single column does not fit the
 * array returns AA_EDIMS.  Any number of columns is accepted: they go through
 * the array in blocks, one after another on the stream.  Results are
 * bit-identical to aa_column_radiance_dev's.  All calls on one GPU share its
 * array, so calls that overlap on one GPU (other streams or host threads)
 * race: serialise them.  Profiling (aa_profile_start) creates events and is
 * not for these callers. */
int aa_column_radiance_dev_noalloc(aa_mie* t, int km, int nobs, int nch, int nq, const int* tracer_idx,
                                   const int* chan_idx, const float* d_qm, const float* d_rh, int nmom, int npol,
                                   int nstreams, int vector, int precision, const int* d_nlayers,
                                   const double* d_rot, const double* d_depol, const double* d_pe,
                                   const double* d_he, const double* d_te, const double* d_albedo,
                                   const double* d_sza, const double* d_raa, const double* d_vza,
                                   const double* d_flux_factor, double missing, double* d_radiance,
                                   double* d_reflectance, double* d_q, double* d_u, int* d_nfourier, void* stream);

/* ------------------------------------------------------------------ work memory
 * No counterpart in the reference (its work arrays are Fortran locals).
 * aa_set_arena_bytes() sets the byte budgets of the per-(GPU, stream) work
 * arena and Chem_Mie optics block of later calls, process-wide; 0 restores the
 * default (AA_WORK_BYTES / AA_OPTICS_BYTES from the environment, else 24 GiB
 * and 4 GiB, sized so that several streams can each hold an arena).  A
 * one-stream bulk caller on a 288 GB GPU gains from 64 GiB / 4 GiB: fewer,
 * larger chunks keep the Fourier terms' shrinking active lists filling the
 * GPU.  The work budget is further capped at 60% of the free device memory.
 * The synchronous host-pointer entry points copy their arguments through one
 * further grow-only block per GPU and run one at a time in a process (a lock
 * held for the whole call), so several host threads may call them.
 * aa_release_arenas() synchronises the current GPU and frees every arena the
 * library holds on it, that block included (they are grown again on the next
 * call).  It must not
 * be called while another host thread is inside a library call on that GPU:
 * that call keeps the arena pointers it carved and would launch into freed
 * memory. */
int aa_set_arena_bytes(size_t work_bytes, size_t optics_bytes);
int aa_release_arenas(void);

/* ------------------------------------------------------------------ launch shapes
 * No counterpart in the reference.  Per (nstreams, nstokes, precision) class
 * of the radiance kernels: bvp_lanes, the lanes per item of the band solver
 * (8 for n = nstreams*nstokes <= 8, 16/32/64 above; the alternative 4 for
 * n <= 4 or 16 for 5 <= n <= 8); sab_pairs, the (item, layer) pairs per
 * wavefront of the SAB/DAB contraction (64 when its tiles fit in registers,
 * else 16; the alternative 16); chunk_items, a cap on the items of one work
 * chunk (0: the memory budget alone).  0 in a field keeps the default.  Every
 * setting gives bit-identical results.  The library reads the table from
 * tuning_gfx950.txt beside the shared library on first use (lines
 * "nstreams nstokes precision bvp_lanes sab_pairs chunk_items", '#' comments;
 * written by aeroapps_amd/tune.py from measurements); aa_load_tuning(path)
 * replaces it from another file (NULL: that default file; AA_ENOFILE when
 * unreadable, AA_EARG when malformed), aa_set_tuning() sets one class (t NULL:
 * back to the defaults), aa_get_tuning() reports the shapes a class runs with.
 * The table is process-wide; a call uses the entry current when it starts. */
typedef struct aa_tuning {
  int bvp_lanes, sab_pairs, chunk_items;
} aa_tuning;
int aa_set_tuning(int nstreams, int nstokes, int precision, const aa_tuning* t);
int aa_get_tuning(int nstreams, int nstokes, int precision, aa_tuning* t);
int aa_load_tuning(const char* path);

/* Library version string. */
const char* aa_version(void);

/* ------------------------------------------------------------------ measurement hook
 * No counterpart in the reference.  aa_profile_start() makes every later
 * kernel launch of the library record a pair of HIP events on its launch
 * stream; aa_profile_read() synchronises the device, returns per-class total
 * milliseconds and launch counts for nkinds classes (0 Chem_Mie optics,
 * 1 input checks, 2 layer optics/delta-M, 3 Chapman + first-order geometry,
 * 4 per-column setup, 5 per-layer eigensystems, 6 boundary-value problem +
 * source + convergence, 7 BRDF supplement, 8 the mixed path's SAB/DAB
 * contraction on the matrix cores) and switches recording off. */
int aa_profile_start(void);
int aa_profile_read(double* ms, long long* launches, int nkinds);

#ifdef __cplusplus
}
#endif
#endif /* AEROAPPS_AMD_H */
