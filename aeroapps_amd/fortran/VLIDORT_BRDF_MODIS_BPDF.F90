! Drop-in replacement of module VLIDORT_BRDF_MODIS_BPDF
! (src/Components/rtms/geo_vlidort/VLIDORT_BRDF_MODIS_BPDF.F90:
! VLIDORT_Vector_LandMODIS_BPDF :24-222): same module, procedure name and
! dummy arguments; the polarised BRDF supplement (MODIS kernels + BPDF-NDVI)
! and the vector radiance run on the GPU through the C-ABI.  DO_BOA = .true.
! is not served (rc = 1001).  The GPU status of the call is reported in rc
! when the reference rc would be 0: rc = 1000 - (C-ABI error code).
module VLIDORT_BRDF_MODIS_BPDF
  use iso_c_binding
  use aeroapps_amd_capi
  implicit none
  private
  public VLIDORT_Vector_LandMODIS_BPDF

contains

  subroutine VLIDORT_Vector_LandMODIS_BPDF(km, nch, nobs, channels, nstreams, plane_parallel, nMom, &
                     nPol, ROT, depol, tau, ssa, pmom, pe, he, te, kernel_wt, RTLSparam, BPDFparam, &
                     solar_zenith, relat_azymuth, sensor_zenith, &
                     MISSING, verbose, radiance_VL_SURF, reflectance_VL_SURF, BR, Q, U, BR_Q, BR_U, rc, DO_BOA)
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
    real*8, target, intent(in) :: kernel_wt(3,nch,nobs), RTLSparam(2,nch,nobs), BPDFparam(3,nch,nobs)
    real*8, target, intent(in) :: solar_zenith(nobs), relat_azymuth(nobs), sensor_zenith(nobs)
    integer, intent(in) :: verbose
    real*8, intent(out) :: radiance_VL_SURF(nobs,nch), reflectance_VL_SURF(nobs,nch)
    integer, intent(out) :: rc
    real*8, intent(out) :: BR(nobs,nch), BR_Q(nobs,nch), BR_U(nobs,nch)
    real*8, intent(out) :: Q(nobs,nch), U(nobs,nch)
    logical, optional, intent(in) :: DO_BOA
    integer(c_int) :: e, crc, pp
    rc = 0
    if (present(DO_BOA)) then
      if (DO_BOA) then
        rc = 1001
        return
      end if
    end if
    pp = 0
    if (plane_parallel) pp = 1
    e = aa_vlidort_vector_landmodis_bpdf(km, nch, nobs, channels, nstreams, pp, nMom, nPol, ROT, depol, tau, &
                                         ssa, pmom, pe, he, te, kernel_wt, RTLSparam, BPDFparam, solar_zenith, &
                                         relat_azymuth, sensor_zenith, MISSING, verbose, 0, radiance_VL_SURF, &
                                         reflectance_VL_SURF, BR, Q, U, BR_Q, BR_U, c_null_ptr, crc)
    rc = crc
    if (rc == 0 .and. e /= 0) rc = 1000 - e
  end subroutine VLIDORT_Vector_LandMODIS_BPDF

end module VLIDORT_BRDF_MODIS_BPDF
