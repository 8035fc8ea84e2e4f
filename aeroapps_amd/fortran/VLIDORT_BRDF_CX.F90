! Drop-in replacement of the Cox-Munk ocean drivers of module VLIDORT_BRDF_CX
! (src/Components/rtms/geo_vlidort/VLIDORT_BRDF_CX.F90: VLIDORT_Scalar_CX
! :420-586, VLIDORT_Vector_CX_SingleGeom :587-762 behind the generic
! VLIDORT_Vector_CX :27-30): same module, procedure names and dummy arguments;
! the BRDF supplement and the radiance run on the GPU through the C-ABI.  The GPU
! status of the call (argument / HIP errors) is reported in rc when the
! reference rc would be 0: rc = 1000 - (C-ABI error code).
module VLIDORT_BRDF_CX
  use iso_c_binding
  use aeroapps_amd_capi
  implicit none
  private
  public VLIDORT_Scalar_CX
  public VLIDORT_Vector_CX
  public VLIDORT_Vector_CX_SingleGeom

  interface VLIDORT_Vector_CX
    module procedure VLIDORT_Vector_CX_SingleGeom
  end interface VLIDORT_Vector_CX

contains

  subroutine VLIDORT_Scalar_CX(km, nch, nobs, channels, nstreams, plane_parallel, nMom, &
                               nPol, ROT, depol, tau, ssa, g, pmom, pe, he, te, U10m, V10m, &
                               mr, solar_zenith, relat_azymuth, sensor_zenith, &
                               MISSING, verbose, radiance_VL_SURF, reflectance_VL_SURF, BRDF, rc)
    integer, intent(in) :: km, nch, nobs
    logical, intent(in) :: plane_parallel
    integer, target, intent(in) :: nMom, nPol
    integer, intent(in) :: nstreams
    real*8, target, intent(in) :: channels(nch)
    real*8, target, intent(in) :: ROT(km,nobs,nch), depol(nch)
    real*8, target, intent(in) :: tau(km,nch,nobs), ssa(km,nch,nobs), g(km,nch,nobs)
    real*8, target, intent(in) :: pmom(km,nch,nobs,nMom,nPol)
    real*8, target, intent(in) :: pe(km+1,nobs), he(km+1,nobs), te(km+1,nobs)
    real*8, target, intent(in) :: U10m(nobs), V10m(nobs), mr(nch)
    real*8, target, intent(in) :: MISSING
    real*8, target, intent(in) :: solar_zenith(nobs), relat_azymuth(nobs), sensor_zenith(nobs)
    integer, intent(in) :: verbose
    real*8, intent(out) :: radiance_VL_SURF(nobs,nch), reflectance_VL_SURF(nobs,nch)
    integer, intent(out) :: rc
    real*8, intent(out) :: BRDF(1,nobs,nch)
    integer(c_int) :: e, crc, pp
    pp = 0
    if (plane_parallel) pp = 1
    e = aa_vlidort_scalar_cx(km, nch, nobs, channels, nstreams, pp, nMom, nPol, ROT, depol, tau, ssa, g, pmom, &
                             pe, he, te, U10m, V10m, mr, solar_zenith, relat_azymuth, sensor_zenith, MISSING, &
                             verbose, 0, radiance_VL_SURF, reflectance_VL_SURF, BRDF, c_null_ptr, crc)
    rc = crc
    if (rc == 0 .and. e /= 0) rc = 1000 - e
  end subroutine VLIDORT_Scalar_CX

  subroutine VLIDORT_Vector_CX_SingleGeom(km, nch, nobs, channels, nstreams, plane_parallel, nMom, &
                                          nPol, ROT, depol, tau, ssa, pmom, pe, he, te, U10m, V10m, &
                                          mr, solar_zenith, relat_azymuth, sensor_zenith, &
                                          MISSING, verbose, radiance_VL_SURF, reflectance_VL_SURF, &
                                          Q, U, BRDF, BRDF_Q, BRDF_U, rc)
    integer, intent(in) :: km, nch, nobs
    logical, intent(in) :: plane_parallel
    integer, target, intent(in) :: nMom, nPol
    integer, intent(in) :: nstreams
    real*8, target, intent(in) :: channels(nch)
    real*8, target, intent(in) :: ROT(km,nobs,nch), depol(nch)
    real*8, target, intent(in) :: tau(km,nch,nobs), ssa(km,nch,nobs)
    real*8, target, intent(in) :: pmom(km,nch,nobs,nMom,nPol)
    real*8, target, intent(in) :: MISSING
    real*8, target, intent(in) :: pe(km+1,nobs), he(km+1,nobs), te(km+1,nobs)
    real*8, target, intent(in) :: U10m(nobs), V10m(nobs), mr(nch)
    real*8, target, intent(in) :: solar_zenith(nobs), relat_azymuth(nobs), sensor_zenith(nobs)
    integer, intent(in) :: verbose
    real*8, intent(out) :: radiance_VL_SURF(nobs,nch), reflectance_VL_SURF(nobs,nch)
    integer, intent(out) :: rc
    real*8, intent(out) :: BRDF(nobs,nch), BRDF_Q(nobs,nch), BRDF_U(nobs,nch)
    real*8, intent(out) :: Q(nobs,nch), U(nobs,nch)
    integer(c_int) :: e, crc, pp
    pp = 0
    if (plane_parallel) pp = 1
    e = aa_vlidort_vector_cx(km, nch, nobs, channels, nstreams, pp, nMom, nPol, ROT, depol, tau, ssa, pmom, &
                             pe, he, te, U10m, V10m, mr, solar_zenith, relat_azymuth, sensor_zenith, MISSING, &
                             verbose, 0, radiance_VL_SURF, reflectance_VL_SURF, Q, U, BRDF, BRDF_Q, BRDF_U, &
                             c_null_ptr, crc)
    rc = crc
    if (rc == 0 .and. e /= 0) rc = 1000 - e
  end subroutine VLIDORT_Vector_CX_SingleGeom

end module VLIDORT_BRDF_CX
