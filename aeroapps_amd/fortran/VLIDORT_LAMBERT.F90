! Drop-in replacement of module VLIDORT_LAMBERT
! (src/Components/rtms/geo_vlidort/VLIDORT_LAMBERT.F90:1-520) for the
! Lambertian clear-sky drivers: same module, procedure names and dummy
! arguments (VLIDORT_Vector_Lambert is the generic of the single- and
! multi-geometry forms, with the optional DO_2OS_CORRECTION and DO_BOA); the
! work runs on the GPU through the C-ABI.  DO_2OS_CORRECTION = .true. runs
! the vector driver with NSTOKES = 1, as the reference does with the VLIDORT
! 2.8.3 library; DO_BOA = .true. is not implemented (rc = 1005).  The GPU
! status of the call (argument / HIP errors) is reported in rc when the
! reference rc would be 0: rc = 1000 - (C-ABI error code).
module VLIDORT_LAMBERT
  use iso_c_binding
  use aeroapps_amd_capi
  implicit none
  private
  public VLIDORT_Scalar_Lambert
  public VLIDORT_Vector_Lambert

  interface VLIDORT_Vector_Lambert
    module procedure VLIDORT_Vector_Lambert_SingleGeom
    module procedure VLIDORT_Vector_Lambert_MultiGeom
  end interface

  public VLIDORT_Vector_Lambert_SingleGeom, VLIDORT_Vector_Lambert_MultiGeom

contains

  subroutine VLIDORT_Scalar_Lambert(km, nch, nobs, channels, nstreams, plane_parallel, nMom, &
                                    nPol, ROT, depol, tau, ssa, g, pmom, pe, he, te, albedo, &
                                    solar_zenith, relat_azymuth, sensor_zenith, &
                                    MISSING, verbose, radiance_VL, reflectance_VL, rc)
    integer, intent(in) :: km, nch, nobs
    logical, intent(in) :: plane_parallel
    integer, target, intent(in) :: nMom, nPol
    integer, intent(in) :: nstreams
    real*8, target, intent(in) :: channels(nch)
    real*8, target, intent(in) :: ROT(km,nobs,nch), depol(nch)
    real*8, target, intent(in) :: tau(km,nch,nobs), ssa(km,nch,nobs), g(km,nch,nobs)
    real*8, target, intent(in) :: pmom(km,nch,nobs,nMom,nPol)
    real*8, target, intent(in) :: pe(km+1,nobs), he(km+1,nobs), te(km+1,nobs)
    real*8, target, intent(in) :: MISSING
    real*8, target, intent(in) :: solar_zenith(nobs), relat_azymuth(nobs), sensor_zenith(nobs)
    real*8, target, intent(in) :: albedo(nobs,nch)
    integer, intent(in) :: verbose
    real*8, intent(out) :: radiance_VL(nobs,nch)
    integer, intent(out) :: rc
    real*8, intent(out) :: reflectance_VL(nobs,nch)
    integer(c_int) :: e, crc
    integer(c_int) :: pp
    pp = 0
    if (plane_parallel) pp = 1
    e = aa_vlidort_scalar_lambert(km, nch, nobs, channels, nstreams, pp, nMom, nPol, ROT, depol, tau, ssa, g, &
                                  pmom, pe, he, te, albedo, solar_zenith, relat_azymuth, sensor_zenith, MISSING, &
                                  verbose, 0, radiance_VL, reflectance_VL, c_null_ptr, crc)
    rc = crc
    if (rc == 0 .and. e /= 0) rc = 1000 - e
  end subroutine VLIDORT_Scalar_Lambert

  subroutine VLIDORT_Vector_Lambert_SingleGeom(km, nch, nobs, channels, nstreams, plane_parallel, nMom, &
                                               nPol, ROT, depol, alpha, tau, ssa, pmom, pe, he, te, albedo, &
                                               solar_zenith, relat_azymuth, sensor_zenith, flux_factor, &
                                               MISSING, verbose, radiance_VL, reflectance_VL, Q, U, rc, &
                                               DO_2OS_CORRECTION, DO_BOA)
    integer, intent(in) :: km, nch, nobs
    logical, intent(in) :: plane_parallel
    integer, target, intent(in) :: nMom, nPol
    integer, intent(in) :: nstreams
    real*8, target, intent(in) :: channels(nch)
    real*8, target, intent(in) :: ROT(km,nobs,nch), depol(nch), alpha(km,nobs,nch)
    real*8, target, intent(in) :: tau(km,nch,nobs), ssa(km,nch,nobs)
    real*8, target, intent(in) :: pmom(km,nch,nobs,nMom,nPol)
    real*8, target, intent(in) :: MISSING
    real*8, target, intent(in) :: pe(km+1,nobs), he(km+1,nobs), te(km+1,nobs)
    real*8, target, intent(in) :: solar_zenith(nobs), relat_azymuth(nobs), sensor_zenith(nobs)
    real*8, intent(in) :: flux_factor(nch,nobs)
    real*8, target, intent(in) :: albedo(nobs,nch)
    integer, intent(in) :: verbose
    real*8, intent(out) :: radiance_VL(nobs,nch), reflectance_VL(nobs,nch), Q(nobs,nch), U(nobs,nch)
    integer, intent(out) :: rc
    logical, optional, intent(in) :: DO_2OS_CORRECTION, DO_BOA
    integer(c_int) :: e, crc, pp, do2, doboa
    rc = 0
    pp = 0
    if (plane_parallel) pp = 1
    call opt_flags_(DO_2OS_CORRECTION, DO_BOA, do2, doboa)
    e = aa_vlidort_vector_lambert_opt(km, nch, nobs, channels, nstreams, pp, nMom, nPol, ROT, depol, alpha, tau, &
                                      ssa, pmom, pe, he, te, albedo, solar_zenith, relat_azymuth, sensor_zenith, &
                                      flux_factor, MISSING, verbose, 0, do2, doboa, radiance_VL, reflectance_VL, &
                                      Q, U, c_null_ptr, crc)
    rc = crc
    if (rc == 0 .and. e /= 0) rc = 1000 - e
  end subroutine VLIDORT_Vector_Lambert_SingleGeom

  subroutine VLIDORT_Vector_Lambert_MultiGeom(km, nch, nobs, ngeom, channels, nstreams, plane_parallel, nMom, &
                                              nPol, ROT, depol, alpha, tau, ssa, pmom, pe, he, te, albedo, &
                                              solar_zenith, relat_azymuth, sensor_zenith, flux_factor, &
                                              MISSING, verbose, radiance_VL, reflectance_VL, Q, U, rc, &
                                              DO_2OS_CORRECTION, DO_BOA)
    integer, intent(in) :: km, nch, nobs, ngeom
    logical, intent(in) :: plane_parallel
    integer, target, intent(in) :: nMom, nPol
    integer, intent(in) :: nstreams
    real*8, target, intent(in) :: channels(nch)
    real*8, target, intent(in) :: ROT(km,nobs,nch), depol(nch), alpha(km,nobs,nch)
    real*8, target, intent(in) :: tau(km,nch,nobs), ssa(km,nch,nobs)
    real*8, target, intent(in) :: pmom(km,nch,nobs,nMom,nPol)
    real*8, target, intent(in) :: MISSING
    real*8, target, intent(in) :: pe(km+1,nobs), he(km+1,nobs), te(km+1,nobs)
    real*8, target, intent(in) :: solar_zenith(nobs,ngeom), relat_azymuth(nobs,ngeom), sensor_zenith(nobs,ngeom)
    real*8, intent(in) :: flux_factor(nch,nobs)
    real*8, target, intent(in) :: albedo(nobs,nch,ngeom)
    integer, intent(in) :: verbose
    real*8, intent(out) :: radiance_VL(nobs,nch,ngeom), reflectance_VL(nobs,nch,ngeom)
    real*8, intent(out) :: Q(nobs,nch,ngeom), U(nobs,nch,ngeom)
    integer, intent(out) :: rc
    logical, optional, intent(in) :: DO_2OS_CORRECTION, DO_BOA
    integer(c_int) :: e, crc, pp, do2, doboa
    rc = 0
    pp = 0
    if (plane_parallel) pp = 1
    call opt_flags_(DO_2OS_CORRECTION, DO_BOA, do2, doboa)
    e = aa_vlidort_vector_lambert_multigeom(km, nch, nobs, ngeom, channels, nstreams, pp, nMom, nPol, ROT, depol, &
                                            alpha, tau, ssa, pmom, pe, he, te, albedo, solar_zenith, &
                                            relat_azymuth, sensor_zenith, flux_factor, MISSING, verbose, 0, do2, &
                                            doboa, radiance_VL, reflectance_VL, Q, U, c_null_ptr, crc)
    rc = crc
    if (rc == 0 .and. e /= 0) rc = 1000 - e
  end subroutine VLIDORT_Vector_Lambert_MultiGeom

  subroutine opt_flags_(DO_2OS_CORRECTION, DO_BOA, do2, doboa)
    logical, optional, intent(in) :: DO_2OS_CORRECTION, DO_BOA
    integer(c_int), intent(out) :: do2, doboa
    do2 = 0
    doboa = 0
    if (present(DO_2OS_CORRECTION)) then
      if (DO_2OS_CORRECTION) do2 = 1
    end if
    if (present(DO_BOA)) then
      if (DO_BOA) doboa = 1
    end if
  end subroutine opt_flags_

end module VLIDORT_LAMBERT
