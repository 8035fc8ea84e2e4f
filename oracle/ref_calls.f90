! ORACLE TEST INFRASTRUCTURE -- not part of the product.
!
! C-callable entry points that call the reference's OWN AeroApps driver layer,
! compiled by oracle/Makefile.ref from the reference sources:
!
!   drv_scalar_lambert    -> VLIDORT_Scalar_Lambert            (GEO/VLIDORT_LAMBERT.F90:34-168)
!   drv_vector_lambert    -> VLIDORT_Vector_Lambert_SingleGeom (GEO/VLIDORT_LAMBERT.F90:172-342)
!   drv_vector_lambert_mg -> VLIDORT_Vector_Lambert_MultiGeom  (GEO/VLIDORT_LAMBERT.F90:346-520)
!   drv_rot_calc          -> VLIDORT_ROT_CALC                  (GEO/VLIDORT_ROT.F90:19-85)
!   drv_coef_depol        -> coef_depol                        (V283/VLIDORT_ScatMod.F90:245-260)
!   drv_scalar_landmodis  -> VLIDORT_Scalar_LandMODIS          (GEO/VLIDORT_BRDF_MODIS.F90:35-202)
!   drv_scalar_cx         -> VLIDORT_Scalar_CX                 (GEO/VLIDORT_BRDF_CX.F90:420-586)
!   drv_vector_cx         -> VLIDORT_Vector_CX_SingleGeom      (GEO/VLIDORT_BRDF_CX.F90:587-762)
!
! These drivers run VLIDORT_Run_Scalar / VLIDORT_Run_Vector
! (V283/VLIDORT_ScatMod.F90) and VLIDORT_Init (V283/VLIDORT_Mod.F90) over the
! VLIDORT 2.8.3 core, all reference code.  The wrappers only convert C
! arguments (int flags -> LOGICAL, verbose = 0) and pre-fill the outputs the
! reference leaves undefined with MISSING:
!   * Q/U of an observation with missing geometry in the vector SingleGeom
!     driver (it writes Q(j,i), U(j,i) with the channel index left over from
!     the previous observation, i.e. i = nch+1 after a completed channel loop);
!     the wrapper passes Q/U buffers padded by one channel so that write stays
!     inside the buffer, and refuses (rc = -99) an observation list whose
!     first observation has missing geometry (there i is undefined);
!   * BR of an observation with missing geometry (MODIS and Cox-Munk drivers),
!     and Q, U, BR, BR_Q, BR_U of the elements the vector Cox-Munk driver
!     skips (missing geometry, winds or refractive index).
! The reference drivers do not return the Fourier-term count; the oracle's
! statement-by-statement harness (ref_driver.f90) does, and the tests require
! the two to agree bit for bit, which pins that count.
! (GEO = src/Components/rtms/geo_vlidort, V283 = src/Components/rtms/RTSI/VLIDORT2p8p3)
module ref_calls_m
  use iso_c_binding
  use VLIDORT_LAMBERT
  use VLIDORT_ROT
  use VLIDORT_BRDF_MODIS
  use VLIDORT_BRDF_CX
  use VLIDORT_LAMBERT_BPDF
  use VLIDORT_BRDF_MODIS_BPDF
  use VLIDORT_ScatMod, only: coef_depol
  implicit none

contains

  logical function geom_missing(x, missing)
    real*8, intent(in) :: x, missing
    geom_missing = abs(x/missing-1) < 0.001
  end function geom_missing

  subroutine drv_scalar_lambert(km, nch, nobs, channels, nstreams, plane_parallel, nMom, &
       nPol, ROT, depol, tau, ssa, g, pmom, pe, he, te, albedo, &
       solar_zenith, relat_azymuth, sensor_zenith, MISSING, &
       radiance_VL, reflectance_VL, rc) bind(C, name="drv_scalar_lambert")
    integer(c_int), value, intent(in) :: km, nch, nobs, nstreams, plane_parallel, nMom, nPol
    real(c_double), intent(in) :: channels(nch), ROT(km,nobs,nch), depol(nch)
    real(c_double), intent(in) :: tau(km,nch,nobs), ssa(km,nch,nobs), g(km,nch,nobs)
    real(c_double), intent(in) :: pmom(km,nch,nobs,nMom,nPol)
    real(c_double), intent(in) :: pe(km+1,nobs), he(km+1,nobs), te(km+1,nobs)
    real(c_double), intent(in) :: albedo(nobs,nch)
    real(c_double), intent(in) :: solar_zenith(nobs), relat_azymuth(nobs), sensor_zenith(nobs)
    real(c_double), value, intent(in) :: MISSING
    real(c_double), intent(out) :: radiance_VL(nobs,nch), reflectance_VL(nobs,nch)
    integer(c_int), intent(out) :: rc
    integer :: irc
    real*8 :: miss
    miss = MISSING
    radiance_VL = MISSING
    reflectance_VL = MISSING
    call VLIDORT_Scalar_Lambert(km, nch, nobs, channels, nstreams, plane_parallel /= 0, nMom, &
         nPol, ROT, depol, tau, ssa, g, pmom, pe, he, te, albedo, &
         solar_zenith, relat_azymuth, sensor_zenith, miss, 0, radiance_VL, reflectance_VL, irc)
    rc = irc
  end subroutine drv_scalar_lambert

  subroutine drv_vector_lambert(km, nch, nobs, channels, nstreams, plane_parallel, nMom, &
       nPol, ROT, depol, alpha, tau, ssa, pmom, pe, he, te, albedo, &
       solar_zenith, relat_azymuth, sensor_zenith, flux_factor, MISSING, do_2os, do_boa, &
       radiance_VL, reflectance_VL, Q, U, rc) bind(C, name="drv_vector_lambert")
    integer(c_int), value, intent(in) :: km, nch, nobs, nstreams, plane_parallel, nMom, nPol
    integer(c_int), value, intent(in) :: do_2os, do_boa
    real(c_double), intent(in) :: channels(nch), ROT(km,nobs,nch), depol(nch), alpha(km,nobs,nch)
    real(c_double), intent(in) :: tau(km,nch,nobs), ssa(km,nch,nobs)
    real(c_double), intent(in) :: pmom(km,nch,nobs,nMom,nPol)
    real(c_double), intent(in) :: pe(km+1,nobs), he(km+1,nobs), te(km+1,nobs)
    real(c_double), intent(in) :: albedo(nobs,nch), flux_factor(nch,nobs)
    real(c_double), intent(in) :: solar_zenith(nobs), relat_azymuth(nobs), sensor_zenith(nobs)
    real(c_double), value, intent(in) :: MISSING
    real(c_double), intent(out) :: radiance_VL(nobs,nch), reflectance_VL(nobs,nch)
    real(c_double), intent(out) :: Q(nobs,nch), U(nobs,nch)
    integer(c_int), intent(out) :: rc
    real*8, allocatable :: Qp(:), Up(:)
    integer :: irc
    real*8 :: miss
    miss = MISSING
    radiance_VL = MISSING
    reflectance_VL = MISSING
    Q = MISSING
    U = MISSING
    if (nobs > 0) then
      if (geom_missing(solar_zenith(1),MISSING) .or. geom_missing(sensor_zenith(1),MISSING) .or. &
          geom_missing(relat_azymuth(1),MISSING)) then
        rc = -99
        return
      end if
    end if
    allocate(Qp(nobs*(nch+1)), Up(nobs*(nch+1)))
    Qp = MISSING
    Up = MISSING
    call VLIDORT_Vector_Lambert_SingleGeom(km, nch, nobs, channels, nstreams, plane_parallel /= 0, nMom, &
         nPol, ROT, depol, alpha, tau, ssa, pmom, pe, he, te, albedo, &
         solar_zenith, relat_azymuth, sensor_zenith, flux_factor, miss, 0, &
         radiance_VL, reflectance_VL, Qp, Up, irc, DO_2OS_CORRECTION=(do_2os /= 0), DO_BOA=(do_boa /= 0))
    Q = reshape(Qp(1:nobs*nch), (/nobs, nch/))
    U = reshape(Up(1:nobs*nch), (/nobs, nch/))
    deallocate(Qp, Up)
    rc = irc
  end subroutine drv_vector_lambert

  subroutine drv_vector_lambert_mg(km, nch, nobs, ngeom, channels, nstreams, plane_parallel, nMom, &
       nPol, ROT, depol, alpha, tau, ssa, pmom, pe, he, te, albedo, &
       solar_zenith, relat_azymuth, sensor_zenith, flux_factor, MISSING, do_2os, do_boa, &
       radiance_VL, reflectance_VL, Q, U, rc) bind(C, name="drv_vector_lambert_mg")
    integer(c_int), value, intent(in) :: km, nch, nobs, ngeom, nstreams, plane_parallel, nMom, nPol
    integer(c_int), value, intent(in) :: do_2os, do_boa
    real(c_double), intent(in) :: channels(nch), ROT(km,nobs,nch), depol(nch), alpha(km,nobs,nch)
    real(c_double), intent(in) :: tau(km,nch,nobs), ssa(km,nch,nobs)
    real(c_double), intent(in) :: pmom(km,nch,nobs,nMom,nPol)
    real(c_double), intent(in) :: pe(km+1,nobs), he(km+1,nobs), te(km+1,nobs)
    real(c_double), intent(in) :: albedo(nobs,nch,ngeom), flux_factor(nch,nobs)
    real(c_double), intent(in) :: solar_zenith(nobs,ngeom), relat_azymuth(nobs,ngeom)
    real(c_double), intent(in) :: sensor_zenith(nobs,ngeom)
    real(c_double), value, intent(in) :: MISSING
    real(c_double), intent(out) :: radiance_VL(nobs,nch,ngeom), reflectance_VL(nobs,nch,ngeom)
    real(c_double), intent(out) :: Q(nobs,nch,ngeom), U(nobs,nch,ngeom)
    integer(c_int), intent(out) :: rc
    integer :: irc
    real*8 :: miss
    miss = MISSING
    radiance_VL = MISSING
    reflectance_VL = MISSING
    Q = MISSING
    U = MISSING
    call VLIDORT_Vector_Lambert_MultiGeom(km, nch, nobs, ngeom, channels, nstreams, plane_parallel /= 0, &
         nMom, nPol, ROT, depol, alpha, tau, ssa, pmom, pe, he, te, albedo, &
         solar_zenith, relat_azymuth, sensor_zenith, flux_factor, miss, 0, &
         radiance_VL, reflectance_VL, Q, U, irc, DO_2OS_CORRECTION=(do_2os /= 0), DO_BOA=(do_boa /= 0))
    rc = irc
  end subroutine drv_vector_lambert_mg

  subroutine drv_rot_calc(km, nch, nobs, channels, pe, he, te, MISSING, ROT, depol, rc) &
       bind(C, name="drv_rot_calc")
    integer(c_int), value, intent(in) :: km, nch, nobs
    real(c_double), intent(in) :: channels(nch)
    real(c_double), intent(in) :: pe(km+1,nobs), he(km+1,nobs), te(km+1,nobs)
    real(c_double), value, intent(in) :: MISSING
    real(c_double), intent(out) :: ROT(km,nobs,nch), depol(nch)
    integer(c_int), intent(out) :: rc
    integer :: irc
    real*8 :: miss
    miss = MISSING
    call VLIDORT_ROT_CALC(km, nch, nobs, channels, pe, he, te, miss, 0, ROT, depol, irc)
    rc = irc
  end subroutine drv_rot_calc

  ! The reference's King-factor depolarisation ratio coef_depol
  ! (V283/VLIDORT_ScatMod.F90:245-260), x = wavelength in micron.
  ! VLIDORT_ROT_CALC's depol output is not usable as a fixture: VLIDORT_Rayleigh
  ! returns it through a pointer to its own local variable
  ! (V283/VLIDORT_ScatMod.F90:185-186), which is gone when the caller reads it.
  real(c_double) function drv_coef_depol(x) bind(C, name="drv_coef_depol")
    real(c_double), value, intent(in) :: x
    real*8 :: xx
    xx = x
    drv_coef_depol = coef_depol(xx)
  end function drv_coef_depol

  subroutine drv_scalar_landmodis(km, nch, nobs, channels, nstreams, plane_parallel, nMom, &
       nPol, ROT, depol, tau, ssa, g, pmom, pe, he, te, kernel_wt, param, &
       solar_zenith, relat_azymuth, sensor_zenith, MISSING, &
       radiance_VL, reflectance_VL, BR, rc) bind(C, name="drv_scalar_landmodis")
    integer(c_int), value, intent(in) :: km, nch, nobs, nstreams, plane_parallel, nMom, nPol
    real(c_double), intent(in) :: channels(nch), ROT(km,nobs,nch), depol(nch)
    real(c_double), intent(in) :: tau(km,nch,nobs), ssa(km,nch,nobs), g(km,nch,nobs)
    real(c_double), intent(in) :: pmom(km,nch,nobs,nMom,nPol)
    real(c_double), intent(in) :: pe(km+1,nobs), he(km+1,nobs), te(km+1,nobs)
    real(c_double), intent(in) :: kernel_wt(3,nch,nobs), param(2,nch,nobs)
    real(c_double), intent(in) :: solar_zenith(nobs), relat_azymuth(nobs), sensor_zenith(nobs)
    real(c_double), value, intent(in) :: MISSING
    real(c_double), intent(out) :: radiance_VL(nobs,nch), reflectance_VL(nobs,nch), BR(nobs,nch)
    integer(c_int), intent(out) :: rc
    integer :: irc
    real*8 :: miss
    miss = MISSING
    radiance_VL = MISSING
    reflectance_VL = MISSING
    BR = MISSING
    call VLIDORT_Scalar_LandMODIS(km, nch, nobs, channels, nstreams, plane_parallel /= 0, nMom, &
         nPol, ROT, depol, tau, ssa, g, pmom, pe, he, te, kernel_wt, param, &
         solar_zenith, relat_azymuth, sensor_zenith, miss, 0, radiance_VL, reflectance_VL, BR, irc)
    rc = irc
  end subroutine drv_scalar_landmodis

  subroutine drv_scalar_cx(km, nch, nobs, channels, nstreams, plane_parallel, nMom, &
       nPol, ROT, depol, tau, ssa, g, pmom, pe, he, te, u10m, v10m, mr, &
       solar_zenith, relat_azymuth, sensor_zenith, MISSING, &
       radiance_VL, reflectance_VL, BR, rc) bind(C, name="drv_scalar_cx")
    integer(c_int), value, intent(in) :: km, nch, nobs, nstreams, plane_parallel, nMom, nPol
    real(c_double), intent(in) :: channels(nch), ROT(km,nobs,nch), depol(nch)
    real(c_double), intent(in) :: tau(km,nch,nobs), ssa(km,nch,nobs), g(km,nch,nobs)
    real(c_double), intent(in) :: pmom(km,nch,nobs,nMom,nPol)
    real(c_double), intent(in) :: pe(km+1,nobs), he(km+1,nobs), te(km+1,nobs)
    real(c_double), intent(in) :: u10m(nobs), v10m(nobs), mr(nch)
    real(c_double), intent(in) :: solar_zenith(nobs), relat_azymuth(nobs), sensor_zenith(nobs)
    real(c_double), value, intent(in) :: MISSING
    real(c_double), intent(out) :: radiance_VL(nobs,nch), reflectance_VL(nobs,nch), BR(nobs,nch)
    integer(c_int), intent(out) :: rc
    integer :: irc
    real*8 :: miss
    miss = MISSING
    radiance_VL = MISSING
    reflectance_VL = MISSING
    BR = MISSING
    call VLIDORT_Scalar_CX(km, nch, nobs, channels, nstreams, plane_parallel /= 0, nMom, &
         nPol, ROT, depol, tau, ssa, g, pmom, pe, he, te, u10m, v10m, mr, &
         solar_zenith, relat_azymuth, sensor_zenith, miss, 0, radiance_VL, reflectance_VL, BR, irc)
    rc = irc
  end subroutine drv_scalar_cx

  subroutine drv_vector_cx(km, nch, nobs, channels, nstreams, plane_parallel, nMom, &
       nPol, ROT, depol, tau, ssa, pmom, pe, he, te, u10m, v10m, mr, &
       solar_zenith, relat_azymuth, sensor_zenith, MISSING, &
       radiance_VL, reflectance_VL, Q, U, BR, BR_Q, BR_U, rc) bind(C, name="drv_vector_cx")
    integer(c_int), value, intent(in) :: km, nch, nobs, nstreams, plane_parallel, nMom, nPol
    real(c_double), intent(in) :: channels(nch), ROT(km,nobs,nch), depol(nch)
    real(c_double), intent(in) :: tau(km,nch,nobs), ssa(km,nch,nobs)
    real(c_double), intent(in) :: pmom(km,nch,nobs,nMom,nPol)
    real(c_double), intent(in) :: pe(km+1,nobs), he(km+1,nobs), te(km+1,nobs)
    real(c_double), intent(in) :: u10m(nobs), v10m(nobs), mr(nch)
    real(c_double), intent(in) :: solar_zenith(nobs), relat_azymuth(nobs), sensor_zenith(nobs)
    real(c_double), value, intent(in) :: MISSING
    real(c_double), intent(out) :: radiance_VL(nobs,nch), reflectance_VL(nobs,nch)
    real(c_double), intent(out) :: Q(nobs,nch), U(nobs,nch), BR(nobs,nch), BR_Q(nobs,nch), BR_U(nobs,nch)
    integer(c_int), intent(out) :: rc
    integer :: irc
    real*8 :: miss
    miss = MISSING
    radiance_VL = MISSING
    reflectance_VL = MISSING
    Q = MISSING
    U = MISSING
    BR = MISSING
    BR_Q = MISSING
    BR_U = MISSING
    call VLIDORT_Vector_CX_SingleGeom(km, nch, nobs, channels, nstreams, plane_parallel /= 0, nMom, &
         nPol, ROT, depol, tau, ssa, pmom, pe, he, te, u10m, v10m, mr, &
         solar_zenith, relat_azymuth, sensor_zenith, miss, 0, radiance_VL, reflectance_VL, &
         Q, U, BR, BR_Q, BR_U, irc)
    rc = irc
  end subroutine drv_vector_cx

  subroutine drv_vector_lambert_bpdf(km, nch, nobs, channels, nstreams, plane_parallel, nMom, &
       nPol, ROT, depol, tau, ssa, pmom, pe, he, te, albedo, bpdf, &
       solar_zenith, relat_azymuth, sensor_zenith, MISSING, &
       radiance_VL, reflectance_VL, BR, Q, U, BR_Q, BR_U, rc) bind(C, name="drv_vector_lambert_bpdf")
    integer(c_int), value, intent(in) :: km, nch, nobs, nstreams, plane_parallel, nMom, nPol
    real(c_double), intent(in) :: channels(nch), ROT(km,nobs,nch), depol(nch)
    real(c_double), intent(in) :: tau(km,nch,nobs), ssa(km,nch,nobs)
    real(c_double), intent(in) :: pmom(km,nch,nobs,nMom,nPol)
    real(c_double), intent(in) :: pe(km+1,nobs), he(km+1,nobs), te(km+1,nobs)
    real(c_double), intent(in) :: albedo(nobs,nch), bpdf(3,nch,nobs)
    real(c_double), intent(in) :: solar_zenith(nobs), relat_azymuth(nobs), sensor_zenith(nobs)
    real(c_double), value, intent(in) :: MISSING
    real(c_double), intent(out) :: radiance_VL(nobs,nch), reflectance_VL(nobs,nch)
    real(c_double), intent(out) :: Q(nobs,nch), U(nobs,nch), BR(nobs,nch), BR_Q(nobs,nch), BR_U(nobs,nch)
    integer(c_int), intent(out) :: rc
    integer :: irc
    real*8 :: miss
    miss = MISSING
    radiance_VL = MISSING
    reflectance_VL = MISSING
    Q = MISSING
    U = MISSING
    BR = MISSING
    BR_Q = MISSING
    BR_U = MISSING
    call VLIDORT_Vector_Lambert_BPDF(km, nch, nobs, channels, nstreams, plane_parallel /= 0, nMom, &
         nPol, ROT, depol, tau, ssa, pmom, pe, he, te, albedo, bpdf, &
         solar_zenith, relat_azymuth, sensor_zenith, miss, 0, radiance_VL, reflectance_VL, &
         BR, Q, U, BR_Q, BR_U, irc)
    rc = irc
  end subroutine drv_vector_lambert_bpdf

  subroutine drv_vector_landmodis_bpdf(km, nch, nobs, channels, nstreams, plane_parallel, nMom, &
       nPol, ROT, depol, tau, ssa, pmom, pe, he, te, kernel_wt, param, bpdf, &
       solar_zenith, relat_azymuth, sensor_zenith, MISSING, &
       radiance_VL, reflectance_VL, BR, Q, U, BR_Q, BR_U, rc) bind(C, name="drv_vector_landmodis_bpdf")
    integer(c_int), value, intent(in) :: km, nch, nobs, nstreams, plane_parallel, nMom, nPol
    real(c_double), intent(in) :: channels(nch), ROT(km,nobs,nch), depol(nch)
    real(c_double), intent(in) :: tau(km,nch,nobs), ssa(km,nch,nobs)
    real(c_double), intent(in) :: pmom(km,nch,nobs,nMom,nPol)
    real(c_double), intent(in) :: pe(km+1,nobs), he(km+1,nobs), te(km+1,nobs)
    real(c_double), intent(in) :: kernel_wt(3,nch,nobs), param(2,nch,nobs), bpdf(3,nch,nobs)
    real(c_double), intent(in) :: solar_zenith(nobs), relat_azymuth(nobs), sensor_zenith(nobs)
    real(c_double), value, intent(in) :: MISSING
    real(c_double), intent(out) :: radiance_VL(nobs,nch), reflectance_VL(nobs,nch)
    real(c_double), intent(out) :: Q(nobs,nch), U(nobs,nch), BR(nobs,nch), BR_Q(nobs,nch), BR_U(nobs,nch)
    integer(c_int), intent(out) :: rc
    integer :: irc
    real*8 :: miss
    miss = MISSING
    radiance_VL = MISSING
    reflectance_VL = MISSING
    Q = MISSING
    U = MISSING
    BR = MISSING
    BR_Q = MISSING
    BR_U = MISSING
    call VLIDORT_Vector_LandMODIS_BPDF(km, nch, nobs, channels, nstreams, plane_parallel /= 0, nMom, &
         nPol, ROT, depol, tau, ssa, pmom, pe, he, te, kernel_wt, param, bpdf, &
         solar_zenith, relat_azymuth, sensor_zenith, miss, 0, radiance_VL, reflectance_VL, &
         BR, Q, U, BR_Q, BR_U, irc)
    rc = irc
  end subroutine drv_vector_landmodis_bpdf

end module ref_calls_m
