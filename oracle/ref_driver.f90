! ORACLE TEST INFRASTRUCTURE -- not part of the product.
!
! C-callable harness around the reference VLIDORT 2.8.3 core, built from the
! unmodified reference sources by oracle/Makefile.ref into oracle/_ref/.
! It restates the per-column setup that the AeroApps drivers perform before
! calling VLIDORT_MASTER, because VLIDORT_ScatMod.F90 (which holds that setup)
! does not compile with flang (LOGICAL == comparison in an unused legacy
! routine at VLIDORT_ScatMod.F90:1352).  Everything numerical below the
! setup is the reference's own code; VLIDORT_Init (VLIDORT_Mod.F90:65-383)
! is the reference's own routine as well.
!
! Restated (reference lines followed statement for statement, including the
! single-precision literals the reference writes in double-precision
! context):
!   ref_scalar_lambert  <- GEO/VLIDORT_LAMBERT.F90:34-168   (VLIDORT_Scalar_Lambert)
!                          V283/VLIDORT_ScatMod.F90:975-1247 (VLIDORT_Run_Scalar)
!   ref_vector_lambert  <- GEO/VLIDORT_LAMBERT.F90:172-342  (VLIDORT_Vector_Lambert_SingleGeom)
!                          V283/VLIDORT_ScatMod.F90:284-617  (VLIDORT_Run_Vector_SingleGeom)
!   ref_scalar_landmodis <- GEO/VLIDORT_BRDF_MODIS.F90:35-202 (VLIDORT_Scalar_LandMODIS)
!                          V283/VLIDORT_SurfaceMod.F90:1376-1596 (VLIDORT_LANDMODIS_SingleGeom)
!                          V283/VLIDORT_ScatMod.F90:975-1247 (VLIDORT_Run_Scalar, BRDF branch)
!   ref_vbrdf_modis      <- the BRDF supplement alone for one geometry
!                          (VLIDORT_LANDMODIS_SingleGeom + the reference's VBRDF_MAINMASTER)
!   ref_scalar_cx        <- GEO/VLIDORT_BRDF_CX.F90:420-586 (VLIDORT_Scalar_CX)
!                          V283/VLIDORT_SurfaceMod.F90:815-1042 (VLIDORT_CoxMunk_SingleGeom)
!   ref_vbrdf_coxmunk    <- the Cox-Munk supplement alone for one geometry
!   ref_vector_cx        <- GEO/VLIDORT_BRDF_CX.F90:587-762 (VLIDORT_Vector_CX_SingleGeom)
!                          V283/VLIDORT_SurfaceMod.F90:815-1042 (VLIDORT_CoxMunk_SingleGeom, NSTOKES = 3)
!                          V283/VLIDORT_ScatMod.F90:284-617 (VLIDORT_Run_Vector_SingleGeom, BRDF branch)
!   ref_vector_bpdf      <- GEO/VLIDORT_LAMBERT_BPDF.F90:23-222 (VLIDORT_Vector_Lambert_BPDF) and
!                          GEO/VLIDORT_BRDF_MODIS_BPDF.F90:24-222 (VLIDORT_Vector_LandMODIS_BPDF)
!                          V283/VLIDORT_SurfaceMod.F90:1822-2060, 2062-2200 (VLIDORT_LANDMODIS_BPDF,
!                          VLIDORT_LAMBERTIAN_BPDF; NSTOKES = 3)
!                          V283/VLIDORT_ScatMod.F90:284-617 (VLIDORT_Run_Vector_SingleGeom, BRDF branch)
!   ref_vbrdf_bpdf       <- the polarised supplement alone for one geometry
! Where the reference leaves BR(j,:) undefined (missing geometry) the harness
! writes MISSING.
! (GEO = src/Components/rtms/geo_vlidort, V283 = src/Components/rtms/RTSI/VLIDORT2p8p3)
!
! Extra output: nfourier(nobs,nch) = TS_FOURIER_SAVED(1) + 1, the number of
! Fourier terms the reference executed (0 where MISSING).
module ref_driver_m
  use iso_c_binding
  use VLIDORT_PARS_m
  use VLIDORT_IO_DEFS_m
  use VLIDORT_MASTERS_m
  use VLIDORT_Mod
  use VBRDF_SUP_MOD_m
  implicit none

contains

  logical function is_missing(x, missing)
    real*8, intent(in) :: x, missing
    is_missing = abs(x/missing-1) < 0.001
  end function is_missing

  ! Per-column input setup shared by the scalar and vector paths
  ! (VLIDORT_ScatMod.F90:1045-1137 / 362-441).
  subroutine set_column(V, nstokes, nmom, albedo, sza, vza, raa, pe, ze, te, km)
    type(VLIDORT), intent(inout) :: V
    integer, intent(in) :: nstokes, nmom, km
    real*8, intent(in) :: albedo, sza, vza, raa
    real*8, intent(in) :: pe(km+1), ze(km+1), te(km+1)
    V%VIO%VLIDORT_FixIn%Cont%TS_NSTOKES = nstokes
    V%VIO%VLIDORT_ModIn%MCont%TS_NGREEK_MOMENTS_INPUT = nmom
    V%VIO%VLIDORT_FixIn%Bool%TS_DO_LAMBERTIAN_SURFACE = .true.
    V%VIO%VLIDORT_FixIn%Optical%TS_LAMBERTIAN_ALBEDO = albedo
    V%VIO%VLIDORT_ModIn%MSunRays%TS_SZANGLES(1) = sza
    V%VIO%VLIDORT_ModIn%MUserVal%TS_USER_RELAZMS(1) = raa
    V%VIO%VLIDORT_ModIn%MUserVal%TS_USER_VZANGLES_INPUT(1) = vza
    V%VIO%VLIDORT_ModIn%MUserVal%TS_N_USER_OBSGEOMS = 1
    V%VIO%VLIDORT_ModIn%MUserVal%TS_USER_OBSGEOMS_INPUT(1,1) = sza
    V%VIO%VLIDORT_ModIn%MUserVal%TS_USER_OBSGEOMS_INPUT(1,2) = vza
    V%VIO%VLIDORT_ModIn%MUserVal%TS_USER_OBSGEOMS_INPUT(1,3) = raa
    V%VIO%VLIDORT_ModIn%MUserVal%TS_USER_LEVELS(1) = 0.0
    ! SZA == 0 latches nadir FO for the rest of the driver call
    ! (VLIDORT_ScatMod.F90:1123-1126)
    if (sza == 0.0) then
      V%VIO%VLIDORT_ModIn%MBool%TS_DO_FOCORR_NADIR = .true.
      V%VIO%VLIDORT_ModIn%MBool%TS_DO_FOCORR_OUTGOING = .false.
    end if
    V%VIO%VLIDORT_FixIn%Chapman%TS_height_grid(0:km) = ze * 1.E-3
    V%VIO%VLIDORT_FixIn%Chapman%TS_pressure_grid(0:km) = pe * 1.E-2
    V%VIO%VLIDORT_FixIn%Chapman%TS_temperature_grid(0:km) = te
  end subroutine set_column

  subroutine ref_scalar_lambert(km, nch, nobs, channels, nstreams, plane_parallel, nMom, &
       nPol, ROT, depol, tau, ssa, g, pmom, pe, he, te, albedo, &
       solar_zenith, relat_azymuth, sensor_zenith, MISSING, &
       radiance_VL, reflectance_VL, nfourier, rc) bind(C, name="ref_scalar_lambert")
    integer(c_int), value, intent(in) :: km, nch, nobs, nstreams, plane_parallel, nMom, nPol
    real(c_double), intent(in) :: channels(nch), ROT(km,nobs,nch), depol(nch)
    real(c_double), intent(in) :: tau(km,nch,nobs), ssa(km,nch,nobs), g(km,nch,nobs)
    real(c_double), intent(in) :: pmom(km,nch,nobs,nMom,nPol)
    real(c_double), intent(in) :: pe(km+1,nobs), he(km+1,nobs), te(km+1,nobs)
    real(c_double), intent(in) :: albedo(nobs,nch)
    real(c_double), intent(in) :: solar_zenith(nobs), relat_azymuth(nobs), sensor_zenith(nobs)
    real(c_double), value, intent(in) :: MISSING
    real(c_double), intent(out) :: radiance_VL(nobs,nch), reflectance_VL(nobs,nch)
    integer(c_int), intent(out) :: nfourier(nobs,nch), rc

    type(VLIDORT), allocatable :: V
    integer :: i, j, l, ier, irc
    real*8 :: ray_l, alpha_l, tau_l, ssa_l, tau_ext, tau_scat, ssa_tot
    real*8 :: raysmom2, aerswt, rayswt, depol_ratio
    real*8, allocatable :: greekmat_total_input(:,:,:)
    real*8 :: aersmom(0:MAXMOMENTS_INPUT)
    real*8 :: deltau_vert_input(MAXLAYERS), omega_total_input(MAXLAYERS)
    real*8, parameter :: pi = 4.*atan(1.0)

    allocate(V)
    allocate(greekmat_total_input(0:MAXMOMENTS_INPUT,MAXLAYERS,16))
    greekmat_total_input = 0.0d0
    rc = 0
    nfourier = 0
    V%NSTREAMS = nstreams
    V%DO_PLANE_PARALLEL = (plane_parallel /= 0)
    call VLIDORT_Init(V, km, irc)
    rc = irc
    if (rc /= 0) return

    do j = 1, nobs
      if (is_missing(solar_zenith(j),MISSING) .or. is_missing(sensor_zenith(j),MISSING) .or. &
          is_missing(relat_azymuth(j),MISSING)) then
        radiance_VL(j,:) = MISSING
        reflectance_VL(j,:) = MISSING
        cycle
      end if
      do i = 1, nch
        if (is_missing(albedo(j,i),MISSING)) then
          radiance_VL(j,i) = MISSING
          reflectance_VL(j,i) = MISSING
          cycle
        end if
        ier = 0
        if (nMom > MAXMOMENTS_INPUT) ier = 5
        if (ier == 0) then
          call set_column(V, 1, nMom, albedo(j,i), solar_zenith(j), sensor_zenith(j), &
                          relat_azymuth(j), pe(:,j), he(:,j), te(:,j), km)
          depol_ratio = depol(i)
          aersmom = 0.0
          raysmom2 = (1.0 - depol_ratio)/(2.0 + depol_ratio)
          do l = 1, km
            ray_l = ROT(l,j,i)
            alpha_l = 0.0
            tau_l = tau(l,i,j)
            ssa_l = ssa(l,i,j)
            tau_ext = alpha_l + ray_l + tau_l
            tau_scat = ray_l + ssa_l * tau_l
            ssa_tot = tau_scat / tau_ext
            if (ssa_tot > 0.99999) then
              ssa_tot = 0.99999
            end if
            deltau_vert_input(l) = tau_ext
            omega_total_input(l) = ssa_tot
            aerswt = ssa_l * tau_l / tau_scat
            rayswt = ray_l / tau_scat
            aersmom(0:nMom-1) = pmom(l,i,j,1:nMom,1)
            greekmat_total_input(0,l,1) = 1.0
            greekmat_total_input(1,l,1) = aersmom(1) * aerswt
            greekmat_total_input(2,l,1) = raysmom2 * rayswt + aersmom(2) * aerswt
            greekmat_total_input(3:nMom,l,1) = aersmom(3:nMom) * aerswt
          end do
          V%VIO%VLIDORT_FixIn%Optical%TS_DELTAU_VERT_INPUT = deltau_vert_input
          V%VIO%VLIDORT_ModIn%MOptical%TS_OMEGA_TOTAL_INPUT = omega_total_input
          V%VIO%VLIDORT_FixIn%Optical%TS_GREEKMAT_TOTAL_INPUT = greekmat_total_input
          call VLIDORT_MASTER(V%DO_DEBUG_INPUT, V%VIO%VLIDORT_FixIn, V%VIO%VLIDORT_ModIn, &
                              V%VIO%VLIDORT_Sup, V%VIO%VLIDORT_Out)
          if (V%VIO%VLIDORT_Out%Status%TS_STATUS_INPUTCHECK /= 0) ier = 3
          if (V%VIO%VLIDORT_Out%Status%TS_STATUS_CALCULATION /= 0) ier = 4
        end if
        if (ier /= 0) then
          radiance_VL(j,i) = MISSING
          reflectance_VL(j,i) = MISSING
          cycle
        end if
        radiance_VL(j,i) = V%VIO%VLIDORT_Out%Main%TS_STOKES(1,1,1,1)
        reflectance_VL(j,i) = (pi * radiance_VL(j,i)) / &
             (cos(V%VIO%VLIDORT_ModIn%MSunRays%TS_SZANGLES(1)*pi/180.0) * &
              V%VIO%VLIDORT_FixIn%SunRays%TS_FLUX_FACTOR)
        nfourier(j,i) = V%VIO%VLIDORT_Out%Main%TS_FOURIER_SAVED(1) + 1
      end do
    end do
    deallocate(greekmat_total_input)
    deallocate(V)
  end subroutine ref_scalar_lambert

  subroutine ref_vector_lambert(km, nch, nobs, channels, nstreams, plane_parallel, nMom, &
       nPol, ROT, depol, alpha, tau, ssa, pmom, pe, he, te, albedo, &
       solar_zenith, relat_azymuth, sensor_zenith, flux_factor, MISSING, &
       radiance_VL, reflectance_VL, Q, U, nfourier, rc) bind(C, name="ref_vector_lambert")
    integer(c_int), value, intent(in) :: km, nch, nobs, nstreams, plane_parallel, nMom, nPol
    real(c_double), intent(in) :: channels(nch), ROT(km,nobs,nch), depol(nch), alpha(km,nobs,nch)
    real(c_double), intent(in) :: tau(km,nch,nobs), ssa(km,nch,nobs)
    real(c_double), intent(in) :: pmom(km,nch,nobs,nMom,nPol)
    real(c_double), intent(in) :: pe(km+1,nobs), he(km+1,nobs), te(km+1,nobs)
    real(c_double), intent(in) :: albedo(nobs,nch), flux_factor(nch,nobs)
    real(c_double), intent(in) :: solar_zenith(nobs), relat_azymuth(nobs), sensor_zenith(nobs)
    real(c_double), value, intent(in) :: MISSING
    real(c_double), intent(out) :: radiance_VL(nobs,nch), reflectance_VL(nobs,nch)
    real(c_double), intent(out) :: Q(nobs,nch), U(nobs,nch)
    integer(c_int), intent(out) :: nfourier(nobs,nch), rc

    type(VLIDORT), allocatable :: V
    integer :: i, j, k, l, ier, irc
    real*8 :: ray_l, alpha_l, tau_l, ssa_l, tau_ext, tau_scat, ssa_tot
    real*8 :: raysmom2, gammamom2, alphamom2, deltamom1, aerswt, rayswt, depol_ratio
    real*8, allocatable :: greekmat_total_input(:,:,:), aervmoms(:,:,:)
    real*8 :: rayvmoms(0:2,16)
    real*8 :: deltau_vert_input(MAXLAYERS), omega_total_input(MAXLAYERS)
    real*8, parameter :: pi = 4.*atan(1.0)

    allocate(V)
    allocate(greekmat_total_input(0:MAXMOMENTS_INPUT,MAXLAYERS,16))
    allocate(aervmoms(0:MAXMOMENTS_INPUT,MAXLAYERS,16))
    greekmat_total_input = 0.0d0
    rc = 0
    nfourier = 0
    V%NSTREAMS = nstreams
    V%DO_PLANE_PARALLEL = (plane_parallel /= 0)
    V%NGREEK_MOMENTS_INPUT = nMom
    call VLIDORT_Init(V, km, irc)
    rc = irc
    if (rc /= 0) return

    do j = 1, nobs
      if (is_missing(solar_zenith(j),MISSING) .or. is_missing(sensor_zenith(j),MISSING) .or. &
          is_missing(relat_azymuth(j),MISSING)) then
        radiance_VL(j,:) = MISSING
        reflectance_VL(j,:) = MISSING
        Q(j,:) = MISSING
        U(j,:) = MISSING
        cycle
      end if
      do i = 1, nch
        V%VIO%VLIDORT_FixIn%SunRays%TS_FLUX_FACTOR = flux_factor(i,j)
        if (is_missing(albedo(j,i),MISSING)) then
          radiance_VL(j,i) = MISSING
          reflectance_VL(j,i) = MISSING
          Q(j,i) = MISSING
          U(j,i) = MISSING
          cycle
        end if
        ier = 0
        if (nMom > MAXMOMENTS_INPUT) ier = 5
        if (ier == 0) then
          call set_column(V, 3, nMom, albedo(j,i), solar_zenith(j), sensor_zenith(j), &
                          relat_azymuth(j), pe(:,j), he(:,j), te(:,j), km)
          depol_ratio = depol(i)
          rayvmoms = 0.0
          aervmoms = 0.0
          gammamom2 = -SQRT(6.) * (1 - depol_ratio) / (2 + depol_ratio)
          alphamom2 = 6 * (1 - depol_ratio) / (2 + depol_ratio)
          deltamom1 = 3 * (1 - 2 * depol_ratio) / (2 + depol_ratio)
          raysmom2 = (1.0 - depol_ratio)/(2.0 + depol_ratio)
          rayvmoms(0,1) = 1.0
          rayvmoms(2,1) = raysmom2
          rayvmoms(2,2) = gammamom2
          rayvmoms(2,5) = gammamom2
          rayvmoms(2,6) = alphamom2
          rayvmoms(1,16) = deltamom1
          do l = 1, km
            ray_l = ROT(l,j,i)
            alpha_l = alpha(l,j,i)
            tau_l = tau(l,i,j)
            ssa_l = ssa(l,i,j)
            tau_ext = alpha_l + ray_l + tau_l
            tau_scat = ray_l + ssa_l * tau_l
            ssa_tot = tau_scat / tau_ext
            if (ssa_tot > 0.99999) then
              ssa_tot = 0.99999
            end if
            deltau_vert_input(l) = tau_ext
            omega_total_input(l) = ssa_tot
            aervmoms(0:nMom-1,l,1)  = pmom(l,i,j,1:nMom,1)
            aervmoms(0:nMom-1,l,2)  = pmom(l,i,j,1:nMom,2)
            aervmoms(0:nMom-1,l,5)  = pmom(l,i,j,1:nMom,2)
            aervmoms(0:nMom-1,l,11) = pmom(l,i,j,1:nMom,3)
            aervmoms(0:nMom-1,l,12) = pmom(l,i,j,1:nMom,4)
            aervmoms(0:nMom-1,l,15) = -pmom(l,i,j,1:nMom,4)
            aervmoms(0:nMom-1,l,6)  = pmom(l,i,j,1:nMom,5)
            aervmoms(0:nMom-1,l,16) = pmom(l,i,j,1:nMom,6)
            aerswt = ssa_l * tau_l / tau_scat
            rayswt = ray_l / tau_scat
            do k = 1, 16
              greekmat_total_input(0:2,l,k) = rayvmoms(0:2,k) * rayswt + aervmoms(0:2,l,k) * aerswt
              greekmat_total_input(3:nMom,l,k) = aervmoms(3:nMom,l,k) * aerswt
            end do
            greekmat_total_input(0,l,1) = 1.0
          end do
          V%VIO%VLIDORT_FixIn%Optical%TS_DELTAU_VERT_INPUT = deltau_vert_input
          V%VIO%VLIDORT_ModIn%MOptical%TS_OMEGA_TOTAL_INPUT = omega_total_input
          V%VIO%VLIDORT_FixIn%Optical%TS_GREEKMAT_TOTAL_INPUT = greekmat_total_input
          call VLIDORT_MASTER(V%DO_DEBUG_INPUT, V%VIO%VLIDORT_FixIn, V%VIO%VLIDORT_ModIn, &
                              V%VIO%VLIDORT_Sup, V%VIO%VLIDORT_Out)
          if (V%VIO%VLIDORT_Out%Status%TS_STATUS_INPUTCHECK /= 0) ier = 3
          if (V%VIO%VLIDORT_Out%Status%TS_STATUS_CALCULATION /= 0) ier = 4
        end if
        if (ier /= 0) then
          radiance_VL(j,i) = MISSING
          reflectance_VL(j,i) = MISSING
          Q(j,i) = MISSING
          U(j,i) = MISSING
          cycle
        end if
        radiance_VL(j,i) = V%VIO%VLIDORT_Out%Main%TS_STOKES(1,1,1,1)
        Q(j,i) = V%VIO%VLIDORT_Out%Main%TS_STOKES(1,1,2,1)
        U(j,i) = V%VIO%VLIDORT_Out%Main%TS_STOKES(1,1,3,1)
        reflectance_VL(j,i) = (pi * radiance_VL(j,i)) / &
             (cos(V%VIO%VLIDORT_ModIn%MSunRays%TS_SZANGLES(1)*pi/180.0) * &
              V%VIO%VLIDORT_FixIn%SunRays%TS_FLUX_FACTOR)
        nfourier(j,i) = V%VIO%VLIDORT_Out%Main%TS_FOURIER_SAVED(1) + 1
      end do
    end do
    deallocate(greekmat_total_input, aervmoms)
    deallocate(V)
  end subroutine ref_vector_lambert

  ! VBRDF_Sup_In as VLIDORT_LANDMODIS_SingleGeom fills it
  ! (V283/VLIDORT_SurfaceMod.F90:1432-1585); fields it leaves unset are zero.
  subroutine landmodis_inputs(B, nstreams, sza, vza, raa, fiso, fgeo, fvol, p1, p2)
    type(VBRDF_Sup_Inputs), intent(out) :: B
    integer, intent(in) :: nstreams
    real*8, intent(in) :: sza, vza, raa, fiso, fgeo, fvol, p1, p2
    B%BS_DO_USER_STREAMS = .true.
    B%BS_DO_BRDF_SURFACE = .true.
    B%BS_DO_SURFACE_EMISSION = .false.
    B%BS_DO_SOLAR_SOURCES = .true.
    B%BS_DO_USER_OBSGEOMS = .true.
    B%BS_DO_DOUBLET_GEOMETRY = .false.
    B%BS_NSTOKES = 1
    B%BS_NSTREAMS = nstreams
    B%BS_NBEAMS = 1
    B%BS_BEAM_SZAS = 0.0d0
    B%BS_BEAM_SZAS(1) = sza
    B%BS_N_USER_RELAZMS = 1
    B%BS_USER_RELAZMS = 0.0d0
    B%BS_USER_RELAZMS(1) = raa
    B%BS_N_USER_STREAMS = 1
    B%BS_USER_ANGLES_INPUT = 0.0d0
    B%BS_USER_ANGLES_INPUT(1) = vza
    B%BS_N_USER_OBSGEOMS = 1
    B%BS_USER_OBSGEOMS = 0.0d0
    B%BS_USER_OBSGEOMS(1,1) = sza
    B%BS_USER_OBSGEOMS(1,2) = vza
    B%BS_USER_OBSGEOMS(1,3) = raa
    B%BS_N_USER_DOUBLETS = 0
    B%BS_USER_DOUBLETS = 0.0d0
    B%BS_N_BRDF_KERNELS = 3
    B%BS_BRDF_NAMES = '          '
    B%BS_WHICH_BRDF = 0
    B%BS_BRDF_FACTORS = 0.0d0
    B%BS_N_BRDF_PARAMETERS = 0
    B%BS_BRDF_PARAMETERS = 0.0d0
    B%BS_LAMBERTIAN_KERNEL_FLAG = .false.
    B%BS_BRDF_NAMES(1) = 'Lambertian'
    B%BS_WHICH_BRDF(1) = 1
    B%BS_BRDF_FACTORS(1) = fiso
    B%BS_LAMBERTIAN_KERNEL_FLAG(1) = .true.
    B%BS_BRDF_NAMES(2) = 'Ross-thick'
    B%BS_WHICH_BRDF(2) = 3
    B%BS_BRDF_FACTORS(2) = fvol
    B%BS_BRDF_NAMES(3) = 'Li-sparse'
    B%BS_WHICH_BRDF(3) = 4
    B%BS_BRDF_FACTORS(3) = fgeo
    B%BS_N_BRDF_PARAMETERS(3) = 2
    B%BS_BRDF_PARAMETERS(3,1) = p1
    B%BS_BRDF_PARAMETERS(3,2) = p2
    B%BS_NSTREAMS_BRDF = 50
    B%BS_DO_DIRECTBOUNCE_ONLY = .false.
    B%BS_DO_SHADOW_EFFECT = .false.
    B%BS_DO_GLITTER_MSRCORR = .false.
    B%BS_DO_GLITTER_MSRCORR_DBONLY = .false.
    B%BS_GLITTER_MSRCORR_ORDER = 0
    B%BS_GLITTER_MSRCORR_NMUQUAD = 0
    B%BS_GLITTER_MSRCORR_NPHIQUAD = 0
    B%BS_DO_NewCMGLINT = .false.
    B%BS_DO_NewGCMGLINT = .false.
    B%BS_SALINITY = 0
    B%BS_WAVELENGTH = 0
    B%BS_WINDSPEED = 0
    B%BS_WINDDIR = 0
    B%BS_DO_GlintShadow = .false.
    B%BS_DO_FoamOption = .false.
    B%BS_DO_FacetIsotropy = .false.
    B%BS_DO_WSABSA_OUTPUT = .true.
    B%BS_DO_WSA_SCALING = .false.
    B%BS_DO_BSA_SCALING = .false.
    B%BS_WSA_VALUE = 0
    B%BS_BSA_VALUE = 0
  end subroutine landmodis_inputs

  ! VBRDF_Sup_In as VLIDORT_CoxMunk_SingleGeom fills it (scalar)
  ! (V283/VLIDORT_SurfaceMod.F90:880-1000): one Cox-Munk kernel, 100 azimuths.
  subroutine coxmunk_inputs(B, nstreams, sza, vza, raa, u10m, v10m, m)
    type(VBRDF_Sup_Inputs), intent(out) :: B
    integer, intent(in) :: nstreams
    real*8, intent(in) :: sza, vza, raa, u10m, v10m, m
    real*8 :: w10m, sigma2
    call landmodis_inputs(B, nstreams, sza, vza, raa, 0.0d0, 0.0d0, 0.0d0, 0.0d0, 0.0d0)
    w10m = sqrt(u10m*u10m + v10m*v10m)
    if (w10m < 0) w10m = 0
    if (w10m > 50) w10m = 50
    sigma2 = 0.003 + 0.00512 * w10m
    B%BS_N_BRDF_KERNELS = 1
    B%BS_BRDF_NAMES = '          '
    B%BS_WHICH_BRDF = 0
    B%BS_BRDF_FACTORS = 0.0d0
    B%BS_N_BRDF_PARAMETERS = 0
    B%BS_BRDF_PARAMETERS = 0.0d0
    B%BS_LAMBERTIAN_KERNEL_FLAG = .false.
    B%BS_BRDF_NAMES(1) = 'Cox-Munk  '
    B%BS_WHICH_BRDF(1) = 9
    B%BS_BRDF_FACTORS(1) = 1.
    B%BS_N_BRDF_PARAMETERS(1) = 3
    B%BS_BRDF_PARAMETERS(1,1) = sigma2
    B%BS_BRDF_PARAMETERS(1,2) = m*m
    B%BS_BRDF_PARAMETERS(1,3) = 0
    B%BS_NSTREAMS_BRDF = 100
    B%BS_DO_SHADOW_EFFECT = .true.
  end subroutine coxmunk_inputs

  subroutine ref_vbrdf_coxmunk(nstreams, sza, vza, raa, u10m, v10m, m, &
       F, F0, UF, UF0, db, wsa, status) bind(C, name="ref_vbrdf_coxmunk")
    integer(c_int), value, intent(in) :: nstreams
    real(c_double), value, intent(in) :: sza, vza, raa, u10m, v10m, m
    real(c_double), intent(out) :: F(0:2*nstreams-1,nstreams,nstreams), F0(0:2*nstreams-1,nstreams)
    real(c_double), intent(out) :: UF(0:2*nstreams-1,nstreams), UF0(0:2*nstreams-1), db, wsa
    integer(c_int), intent(out) :: status
    type(VBRDF_Sup_Inputs), allocatable :: B
    type(VBRDF_Sup_Outputs), allocatable :: O
    type(VBRDF_Output_Exception_Handling) :: S
    integer :: k
    allocate(B, O)
    call coxmunk_inputs(B, nstreams, sza, vza, raa, u10m, v10m, m)
    call VBRDF_MAINMASTER(.false., 2*nstreams-1, B, O, S)
    do k = 0, 2*nstreams-1
      F(k,:,:) = O%BS_BRDF_F(k,1,1:nstreams,1:nstreams)
      F0(k,:) = O%BS_BRDF_F_0(k,1,1:nstreams,1)
      UF(k,:) = O%BS_USER_BRDF_F(k,1,1,1:nstreams)
      UF0(k) = O%BS_USER_BRDF_F_0(k,1,1,1)
    end do
    db = O%BS_DBOUNCE_BRDFUNC(1,1,1,1)
    wsa = O%BS_WSA_CALCULATED
    status = S%BS_STATUS_OUTPUT
    deallocate(B, O)
  end subroutine ref_vbrdf_coxmunk

  ! VBRDF_Sup_In as VLIDORT_LANDMODIS_BPDF (modis /= 0: Lambertian fiso,
  ! Ross-thick fvol, Li-sparse fgeo (h/b, b/r), BPDF-NDVI) or
  ! VLIDORT_LAMBERTIAN_BPDF (modis = 0: Lambertian albedo = fiso, BPDF-NDVI)
  ! fill it (V283/VLIDORT_SurfaceMod.F90:1822-2060, 2062-2200): NSTOKES = 3,
  ! 50 azimuth streams, BPDF factor 1.
  subroutine bpdf_inputs(B, nstreams, sza, vza, raa, modis, fiso, fgeo, fvol, p1, p2, b1, b2, b3)
    type(VBRDF_Sup_Inputs), intent(out) :: B
    integer, intent(in) :: nstreams, modis
    real*8, intent(in) :: sza, vza, raa, fiso, fgeo, fvol, p1, p2, b1, b2, b3
    integer :: k
    call landmodis_inputs(B, nstreams, sza, vza, raa, fiso, fgeo, fvol, p1, p2)
    B%BS_NSTOKES = 3
    if (modis /= 0) then
      k = 4
    else
      B%BS_BRDF_NAMES(2:3) = '          '
      B%BS_WHICH_BRDF(2:3) = 0
      B%BS_BRDF_FACTORS(2:3) = 0.0d0
      B%BS_N_BRDF_PARAMETERS(2:3) = 0
      B%BS_BRDF_PARAMETERS(2:3,:) = 0.0d0
      k = 2
    end if
    B%BS_N_BRDF_KERNELS = k
    B%BS_BRDF_NAMES(k) = 'BPDF-NDVI'
    B%BS_WHICH_BRDF(k) = 14
    B%BS_BRDF_FACTORS(k) = 1.0
    B%BS_N_BRDF_PARAMETERS(k) = 3
    B%BS_BRDF_PARAMETERS(k,1) = b1
    B%BS_BRDF_PARAMETERS(k,2) = b2
    B%BS_BRDF_PARAMETERS(k,3) = b3
    B%BS_LAMBERTIAN_KERNEL_FLAG(k) = .false.
  end subroutine bpdf_inputs

  ! The polarised supplement for one geometry, every Mueller index q = 1..16
  ! (4 (o1 - 1) + o2): F(0:2N-1,16,N,N), F0(0:2N-1,16,N), UF(0:2N-1,16,N),
  ! direct bounce DB(16), white-sky albedo, status.
  subroutine ref_vbrdf_bpdf(nstreams, sza, vza, raa, modis, fiso, fgeo, fvol, p1, p2, b1, b2, b3, &
       F, F0, UF, DB, wsa, status) bind(C, name="ref_vbrdf_bpdf")
    integer(c_int), value, intent(in) :: nstreams, modis
    real(c_double), value, intent(in) :: sza, vza, raa, fiso, fgeo, fvol, p1, p2, b1, b2, b3
    real(c_double), intent(out) :: F(0:2*nstreams-1,16,nstreams,nstreams), F0(0:2*nstreams-1,16,nstreams)
    real(c_double), intent(out) :: UF(0:2*nstreams-1,16,nstreams), DB(16), wsa
    integer(c_int), intent(out) :: status
    type(VBRDF_Sup_Inputs), allocatable :: B
    type(VBRDF_Sup_Outputs), allocatable :: O
    type(VBRDF_Output_Exception_Handling) :: S
    integer :: m
    allocate(B, O)
    call bpdf_inputs(B, nstreams, sza, vza, raa, modis, fiso, fgeo, fvol, p1, p2, b1, b2, b3)
    call VBRDF_MAINMASTER(.false., 2*nstreams-1, B, O, S)
    do m = 0, 2*nstreams-1
      F(m,:,:,:) = O%BS_BRDF_F(m,1:16,1:nstreams,1:nstreams)
      F0(m,:,:) = O%BS_BRDF_F_0(m,1:16,1:nstreams,1)
      UF(m,:,:) = O%BS_USER_BRDF_F(m,1:16,1,1:nstreams)
    end do
    DB = O%BS_DBOUNCE_BRDFUNC(1:16,1,1,1)
    wsa = O%BS_WSA_CALCULATED
    status = S%BS_STATUS_OUTPUT
    deallocate(B, O)
  end subroutine ref_vbrdf_bpdf

  subroutine ref_vector_bpdf(km, nch, nobs, channels, nstreams, plane_parallel, nMom, &
       nPol, ROT, depol, tau, ssa, pmom, pe, he, te, modis, albedo, kernel_wt, param, bpdf, &
       solar_zenith, relat_azymuth, sensor_zenith, MISSING, &
       radiance_VL, reflectance_VL, Q, U, BR, BR_Q, BR_U, nfourier, rc) bind(C, name="ref_vector_bpdf")
    integer(c_int), value, intent(in) :: km, nch, nobs, nstreams, plane_parallel, nMom, nPol, modis
    real(c_double), intent(in) :: channels(nch), ROT(km,nobs,nch), depol(nch)
    real(c_double), intent(in) :: tau(km,nch,nobs), ssa(km,nch,nobs)
    real(c_double), intent(in) :: pmom(km,nch,nobs,nMom,nPol)
    real(c_double), intent(in) :: pe(km+1,nobs), he(km+1,nobs), te(km+1,nobs)
    real(c_double), intent(in) :: albedo(nobs,nch), kernel_wt(3,nch,nobs), param(2,nch,nobs), bpdf(3,nch,nobs)
    real(c_double), intent(in) :: solar_zenith(nobs), relat_azymuth(nobs), sensor_zenith(nobs)
    real(c_double), value, intent(in) :: MISSING
    real(c_double), intent(out) :: radiance_VL(nobs,nch), reflectance_VL(nobs,nch)
    real(c_double), intent(out) :: Q(nobs,nch), U(nobs,nch), BR(nobs,nch), BR_Q(nobs,nch), BR_U(nobs,nch)
    integer(c_int), intent(out) :: nfourier(nobs,nch), rc

    type(VLIDORT), allocatable :: V
    type(VBRDF_Sup_Inputs), allocatable :: B
    integer :: i, j, k, l, ier, irc
    real*8 :: ray_l, tau_l, ssa_l, tau_ext, tau_scat, ssa_tot
    real*8 :: raysmom2, gammamom2, alphamom2, deltamom1, aerswt, rayswt, depol_ratio
    real*8, allocatable :: greekmat_total_input(:,:,:), aervmoms(:,:,:)
    real*8 :: rayvmoms(0:2,16)
    real*8 :: deltau_vert_input(MAXLAYERS), omega_total_input(MAXLAYERS)
    real*8, parameter :: pi = 4.*atan(1.0)

    allocate(V, B)
    allocate(greekmat_total_input(0:MAXMOMENTS_INPUT,MAXLAYERS,16))
    allocate(aervmoms(0:MAXMOMENTS_INPUT,MAXLAYERS,16))
    greekmat_total_input = 0.0d0
    radiance_VL = MISSING
    reflectance_VL = MISSING
    Q = MISSING
    U = MISSING
    BR = MISSING
    BR_Q = MISSING
    BR_U = MISSING
    rc = 0
    nfourier = 0
    V%NSTREAMS = nstreams
    V%DO_PLANE_PARALLEL = (plane_parallel /= 0)
    V%NGREEK_MOMENTS_INPUT = nMom
    call VLIDORT_Init(V, km, irc)
    rc = irc
    if (rc /= 0) return

    do j = 1, nobs
      if (is_missing(solar_zenith(j),MISSING) .or. is_missing(sensor_zenith(j),MISSING) .or. &
          is_missing(relat_azymuth(j),MISSING)) cycle
      do i = 1, nch
        ! (the drivers' checks of missing kernel weights and BPDF parameters
        ! leave only their inner loops, so only the albedo check skips)
        if (modis == 0) then
          if (is_missing(albedo(j,i),MISSING)) cycle
          call bpdf_inputs(B, nstreams, solar_zenith(j), sensor_zenith(j), relat_azymuth(j), 0, &
                           albedo(j,i), 0.0d0, 0.0d0, 0.0d0, 0.0d0, bpdf(1,i,j), bpdf(2,i,j), bpdf(3,i,j))
        else
          call bpdf_inputs(B, nstreams, solar_zenith(j), sensor_zenith(j), relat_azymuth(j), 1, &
                           kernel_wt(1,i,j), kernel_wt(2,i,j), kernel_wt(3,i,j), param(1,i,j), param(2,i,j), &
                           bpdf(1,i,j), bpdf(2,i,j), bpdf(3,i,j))
        end if
        call VBRDF_MAINMASTER(.false., 2*nstreams-1, B, V%VIO%VBRDF_Sup_Out, V%VIO%VBRDF_Sup_OutputStatus)
        BR(j,i) = V%VIO%VBRDF_Sup_Out%BS_DBOUNCE_BRDFUNC(1,1,1,1)
        BR_Q(j,i) = V%VIO%VBRDF_Sup_Out%BS_DBOUNCE_BRDFUNC(2,1,1,1)
        BR_U(j,i) = V%VIO%VBRDF_Sup_Out%BS_DBOUNCE_BRDFUNC(3,1,1,1)
        if (BR(j,i) < 0) then
          radiance_VL(j,i) = -500
          reflectance_VL(j,i) = -500
          BR(j,i) = -500
          BR_Q(j,i) = -500
          BR_U(j,i) = -500
          Q(j,i) = -500
          U(j,i) = -500
          cycle
        end if
        ier = 0
        call set_column(V, 3, nMom, 0.0d0, solar_zenith(j), sensor_zenith(j), &
                        relat_azymuth(j), pe(:,j), he(:,j), te(:,j), km)
        V%VIO%VLIDORT_FixIn%Bool%TS_DO_LAMBERTIAN_SURFACE = .false.
        V%VIO%VLIDORT_FixIn%Optical%TS_LAMBERTIAN_ALBEDO = 0.0
        V%VIO%VLIDORT_Sup%BRDF%TS_BRDF_F_0        = V%VIO%VBRDF_Sup_Out%BS_BRDF_F_0
        V%VIO%VLIDORT_Sup%BRDF%TS_BRDF_F          = V%VIO%VBRDF_Sup_Out%BS_BRDF_F
        V%VIO%VLIDORT_Sup%BRDF%TS_USER_BRDF_F_0   = V%VIO%VBRDF_Sup_Out%BS_USER_BRDF_F_0
        V%VIO%VLIDORT_Sup%BRDF%TS_USER_BRDF_F     = V%VIO%VBRDF_Sup_Out%BS_USER_BRDF_F
        V%VIO%VLIDORT_Sup%BRDF%TS_EXACTDB_BRDFUNC = V%VIO%VBRDF_Sup_Out%BS_DBOUNCE_BRDFUNC
        V%VIO%VLIDORT_Sup%BRDF%TS_EMISSIVITY      = V%VIO%VBRDF_Sup_Out%BS_EMISSIVITY
        V%VIO%VLIDORT_Sup%BRDF%TS_USER_EMISSIVITY = V%VIO%VBRDF_Sup_Out%BS_USER_EMISSIVITY
        if (nMom > MAXMOMENTS_INPUT) ier = 5
        if (ier == 0) then
          depol_ratio = depol(i)
          rayvmoms = 0.0
          aervmoms = 0.0
          gammamom2 = -SQRT(6.) * (1 - depol_ratio) / (2 + depol_ratio)
          alphamom2 = 6 * (1 - depol_ratio) / (2 + depol_ratio)
          deltamom1 = 3 * (1 - 2 * depol_ratio) / (2 + depol_ratio)
          raysmom2 = (1.0 - depol_ratio)/(2.0 + depol_ratio)
          rayvmoms(0,1) = 1.0
          rayvmoms(2,1) = raysmom2
          rayvmoms(2,2) = gammamom2
          rayvmoms(2,5) = gammamom2
          rayvmoms(2,6) = alphamom2
          rayvmoms(1,16) = deltamom1
          do l = 1, km
            ray_l = ROT(l,j,i)
            tau_l = tau(l,i,j)
            ssa_l = ssa(l,i,j)
            tau_ext = ray_l + tau_l
            tau_scat = ray_l + ssa_l * tau_l
            ssa_tot = tau_scat / tau_ext
            if (ssa_tot > 0.99999) then
              ssa_tot = 0.99999
            end if
            deltau_vert_input(l) = tau_ext
            omega_total_input(l) = ssa_tot
            aervmoms(0:nMom-1,l,1)  = pmom(l,i,j,1:nMom,1)
            aervmoms(0:nMom-1,l,2)  = pmom(l,i,j,1:nMom,2)
            aervmoms(0:nMom-1,l,5)  = pmom(l,i,j,1:nMom,2)
            aervmoms(0:nMom-1,l,11) = pmom(l,i,j,1:nMom,3)
            aervmoms(0:nMom-1,l,12) = pmom(l,i,j,1:nMom,4)
            aervmoms(0:nMom-1,l,15) = -pmom(l,i,j,1:nMom,4)
            aervmoms(0:nMom-1,l,6)  = pmom(l,i,j,1:nMom,5)
            aervmoms(0:nMom-1,l,16) = pmom(l,i,j,1:nMom,6)
            aerswt = ssa_l * tau_l / tau_scat
            rayswt = ray_l / tau_scat
            do k = 1, 16
              greekmat_total_input(0:2,l,k) = rayvmoms(0:2,k) * rayswt + aervmoms(0:2,l,k) * aerswt
              greekmat_total_input(3:nMom,l,k) = aervmoms(3:nMom,l,k) * aerswt
            end do
            greekmat_total_input(0,l,1) = 1.0
          end do
          V%VIO%VLIDORT_FixIn%Optical%TS_DELTAU_VERT_INPUT = deltau_vert_input
          V%VIO%VLIDORT_ModIn%MOptical%TS_OMEGA_TOTAL_INPUT = omega_total_input
          V%VIO%VLIDORT_FixIn%Optical%TS_GREEKMAT_TOTAL_INPUT = greekmat_total_input
          call VLIDORT_MASTER(V%DO_DEBUG_INPUT, V%VIO%VLIDORT_FixIn, V%VIO%VLIDORT_ModIn, &
                              V%VIO%VLIDORT_Sup, V%VIO%VLIDORT_Out)
          if (V%VIO%VLIDORT_Out%Status%TS_STATUS_INPUTCHECK /= 0) ier = 3
          if (V%VIO%VLIDORT_Out%Status%TS_STATUS_CALCULATION /= 0) ier = 4
        end if
        if (ier /= 0) then
          radiance_VL(j,i) = MISSING
          reflectance_VL(j,i) = MISSING
          cycle
        end if
        radiance_VL(j,i) = V%VIO%VLIDORT_Out%Main%TS_STOKES(1,1,1,1)
        Q(j,i) = V%VIO%VLIDORT_Out%Main%TS_STOKES(1,1,2,1)
        U(j,i) = V%VIO%VLIDORT_Out%Main%TS_STOKES(1,1,3,1)
        reflectance_VL(j,i) = (pi * radiance_VL(j,i)) / &
             (cos(V%VIO%VLIDORT_ModIn%MSunRays%TS_SZANGLES(1)*pi/180.0) * &
              V%VIO%VLIDORT_FixIn%SunRays%TS_FLUX_FACTOR)
        nfourier(j,i) = V%VIO%VLIDORT_Out%Main%TS_FOURIER_SAVED(1) + 1
      end do
    end do
    deallocate(greekmat_total_input, aervmoms)
    deallocate(V, B)
  end subroutine ref_vector_bpdf

  ! The BRDF supplement for one geometry: F(0:2N-1,N,N) (incident j -> i),
  ! F0(0:2N-1,N), UF(0:2N-1,N), UF0(0:2N-1), exact direct bounce, total
  ! white-sky albedo and the supplement's output status.
  subroutine ref_vbrdf_modis(nstreams, sza, vza, raa, fiso, fgeo, fvol, p1, p2, &
       F, F0, UF, UF0, db, wsa, status) bind(C, name="ref_vbrdf_modis")
    integer(c_int), value, intent(in) :: nstreams
    real(c_double), value, intent(in) :: sza, vza, raa, fiso, fgeo, fvol, p1, p2
    real(c_double), intent(out) :: F(0:2*nstreams-1,nstreams,nstreams), F0(0:2*nstreams-1,nstreams)
    real(c_double), intent(out) :: UF(0:2*nstreams-1,nstreams), UF0(0:2*nstreams-1), db, wsa
    integer(c_int), intent(out) :: status
    type(VBRDF_Sup_Inputs), allocatable :: B
    type(VBRDF_Sup_Outputs), allocatable :: O
    type(VBRDF_Output_Exception_Handling) :: S
    integer :: m
    allocate(B, O)
    call landmodis_inputs(B, nstreams, sza, vza, raa, fiso, fgeo, fvol, p1, p2)
    call VBRDF_MAINMASTER(.false., 2*nstreams-1, B, O, S)
    do m = 0, 2*nstreams-1
      F(m,:,:) = O%BS_BRDF_F(m,1,1:nstreams,1:nstreams)
      F0(m,:) = O%BS_BRDF_F_0(m,1,1:nstreams,1)
      UF(m,:) = O%BS_USER_BRDF_F(m,1,1,1:nstreams)
      UF0(m) = O%BS_USER_BRDF_F_0(m,1,1,1)
    end do
    db = O%BS_DBOUNCE_BRDFUNC(1,1,1,1)
    wsa = O%BS_WSA_CALCULATED
    status = S%BS_STATUS_OUTPUT
    deallocate(B, O)
  end subroutine ref_vbrdf_modis

  subroutine ref_scalar_landmodis(km, nch, nobs, channels, nstreams, plane_parallel, nMom, &
       nPol, ROT, depol, tau, ssa, g, pmom, pe, he, te, kernel_wt, param, &
       solar_zenith, relat_azymuth, sensor_zenith, MISSING, &
       radiance_VL, reflectance_VL, BR, nfourier, rc) bind(C, name="ref_scalar_landmodis")
    integer(c_int), value, intent(in) :: km, nch, nobs, nstreams, plane_parallel, nMom, nPol
    real(c_double), intent(in) :: channels(nch), ROT(km,nobs,nch), depol(nch)
    real(c_double), intent(in) :: tau(km,nch,nobs), ssa(km,nch,nobs), g(km,nch,nobs)
    real(c_double), intent(in) :: pmom(km,nch,nobs,nMom,nPol)
    real(c_double), intent(in) :: pe(km+1,nobs), he(km+1,nobs), te(km+1,nobs)
    real(c_double), intent(in) :: kernel_wt(3,nch,nobs), param(2,nch,nobs)
    real(c_double), intent(in) :: solar_zenith(nobs), relat_azymuth(nobs), sensor_zenith(nobs)
    real(c_double), value, intent(in) :: MISSING
    real(c_double), intent(out) :: radiance_VL(nobs,nch), reflectance_VL(nobs,nch), BR(nobs,nch)
    integer(c_int), intent(out) :: nfourier(nobs,nch), rc

    type(VLIDORT), allocatable :: V
    type(VBRDF_Sup_Inputs), allocatable :: B
    integer :: i, j, l, n, p, ier, irc
    real*8 :: ray_l, alpha_l, tau_l, ssa_l, tau_ext, tau_scat, ssa_tot
    real*8 :: raysmom2, aerswt, rayswt, depol_ratio
    real*8, allocatable :: greekmat_total_input(:,:,:)
    real*8 :: aersmom(0:MAXMOMENTS_INPUT)
    real*8 :: deltau_vert_input(MAXLAYERS), omega_total_input(MAXLAYERS)
    real*8, parameter :: pi = 4.*atan(1.0)

    allocate(V, B)
    allocate(greekmat_total_input(0:MAXMOMENTS_INPUT,MAXLAYERS,16))
    greekmat_total_input = 0.0d0
    rc = 0
    nfourier = 0
    V%NSTREAMS = nstreams
    V%DO_PLANE_PARALLEL = (plane_parallel /= 0)
    V%NGREEK_MOMENTS_INPUT = nMom
    call VLIDORT_Init(V, km, irc)
    rc = irc
    if (rc /= 0) return

    do j = 1, nobs
      if (is_missing(solar_zenith(j),MISSING) .or. is_missing(sensor_zenith(j),MISSING) .or. &
          is_missing(relat_azymuth(j),MISSING)) then
        radiance_VL(j,:) = MISSING
        reflectance_VL(j,:) = MISSING
        BR(j,:) = MISSING
        cycle
      end if
      do i = 1, nch
        ! the reference's kernel-weight checks: their CYCLE only leaves the
        ! inner loop, so the channel is still computed afterwards
        do n = 1, 3
          if (is_missing(kernel_wt(n,i,j),MISSING)) then
            radiance_VL(j,i) = MISSING
            reflectance_VL(j,i) = MISSING
            cycle
          end if
        end do
        do p = 1, 2
          if (is_missing(param(p,i,j),MISSING)) then
            radiance_VL(j,i) = MISSING
            reflectance_VL(j,i) = MISSING
            cycle
          end if
        end do
        call landmodis_inputs(B, nstreams, solar_zenith(j), sensor_zenith(j), relat_azymuth(j), &
                              kernel_wt(1,i,j), kernel_wt(2,i,j), kernel_wt(3,i,j), &
                              param(1,i,j), param(2,i,j))
        call VBRDF_MAINMASTER(.false., 2*nstreams-1, B, V%VIO%VBRDF_Sup_Out, V%VIO%VBRDF_Sup_OutputStatus)
        BR(j,i) = V%VIO%VBRDF_Sup_Out%BS_DBOUNCE_BRDFUNC(1,1,1,1)
        if (BR(j,i) < 0) then
          radiance_VL(j,i) = -500
          reflectance_VL(j,i) = -500
          BR(j,i) = -500
          cycle
        end if
        ier = 0
        call set_column(V, 1, nMom, 0.0d0, solar_zenith(j), sensor_zenith(j), &
                        relat_azymuth(j), pe(:,j), he(:,j), te(:,j), km)
        V%VIO%VLIDORT_FixIn%Bool%TS_DO_LAMBERTIAN_SURFACE = .false.
        V%VIO%VLIDORT_FixIn%Optical%TS_LAMBERTIAN_ALBEDO = 0.0
        V%VIO%VLIDORT_Sup%BRDF%TS_BRDF_F_0        = V%VIO%VBRDF_Sup_Out%BS_BRDF_F_0
        V%VIO%VLIDORT_Sup%BRDF%TS_BRDF_F          = V%VIO%VBRDF_Sup_Out%BS_BRDF_F
        V%VIO%VLIDORT_Sup%BRDF%TS_USER_BRDF_F_0   = V%VIO%VBRDF_Sup_Out%BS_USER_BRDF_F_0
        V%VIO%VLIDORT_Sup%BRDF%TS_USER_BRDF_F     = V%VIO%VBRDF_Sup_Out%BS_USER_BRDF_F
        V%VIO%VLIDORT_Sup%BRDF%TS_EXACTDB_BRDFUNC = V%VIO%VBRDF_Sup_Out%BS_DBOUNCE_BRDFUNC
        V%VIO%VLIDORT_Sup%BRDF%TS_EMISSIVITY      = V%VIO%VBRDF_Sup_Out%BS_EMISSIVITY
        V%VIO%VLIDORT_Sup%BRDF%TS_USER_EMISSIVITY = V%VIO%VBRDF_Sup_Out%BS_USER_EMISSIVITY
        depol_ratio = depol(i)
        aersmom = 0.0
        raysmom2 = (1.0 - depol_ratio)/(2.0 + depol_ratio)
        do l = 1, km
          ray_l = ROT(l,j,i)
          alpha_l = 0.0
          tau_l = tau(l,i,j)
          ssa_l = ssa(l,i,j)
          tau_ext = alpha_l + ray_l + tau_l
          tau_scat = ray_l + ssa_l * tau_l
          ssa_tot = tau_scat / tau_ext
          if (ssa_tot > 0.99999) then
            ssa_tot = 0.99999
          end if
          deltau_vert_input(l) = tau_ext
          omega_total_input(l) = ssa_tot
          aerswt = ssa_l * tau_l / tau_scat
          rayswt = ray_l / tau_scat
          aersmom(0:nMom-1) = pmom(l,i,j,1:nMom,1)
          greekmat_total_input(0,l,1) = 1.0
          greekmat_total_input(1,l,1) = aersmom(1) * aerswt
          greekmat_total_input(2,l,1) = raysmom2 * rayswt + aersmom(2) * aerswt
          greekmat_total_input(3:nMom,l,1) = aersmom(3:nMom) * aerswt
        end do
        V%VIO%VLIDORT_FixIn%Optical%TS_DELTAU_VERT_INPUT = deltau_vert_input
        V%VIO%VLIDORT_ModIn%MOptical%TS_OMEGA_TOTAL_INPUT = omega_total_input
        V%VIO%VLIDORT_FixIn%Optical%TS_GREEKMAT_TOTAL_INPUT = greekmat_total_input
        call VLIDORT_MASTER(V%DO_DEBUG_INPUT, V%VIO%VLIDORT_FixIn, V%VIO%VLIDORT_ModIn, &
                            V%VIO%VLIDORT_Sup, V%VIO%VLIDORT_Out)
        if (V%VIO%VLIDORT_Out%Status%TS_STATUS_INPUTCHECK /= 0) ier = 3
        if (V%VIO%VLIDORT_Out%Status%TS_STATUS_CALCULATION /= 0) ier = 4
        radiance_VL(j,i) = V%VIO%VLIDORT_Out%Main%TS_STOKES(1,1,1,1)
        reflectance_VL(j,i) = (pi * radiance_VL(j,i)) / &
             (cos(V%VIO%VLIDORT_ModIn%MSunRays%TS_SZANGLES(1)*pi/180.0) * &
              V%VIO%VLIDORT_FixIn%SunRays%TS_FLUX_FACTOR)
        if (ier /= 0) then
          radiance_VL(j,i) = MISSING
          reflectance_VL(j,i) = MISSING
          cycle
        end if
        nfourier(j,i) = V%VIO%VLIDORT_Out%Main%TS_FOURIER_SAVED(1) + 1
      end do
    end do
    deallocate(greekmat_total_input)
    deallocate(V, B)
  end subroutine ref_scalar_landmodis

  subroutine ref_scalar_cx(km, nch, nobs, channels, nstreams, plane_parallel, nMom, &
       nPol, ROT, depol, tau, ssa, g, pmom, pe, he, te, u10m, v10m, mr, &
       solar_zenith, relat_azymuth, sensor_zenith, MISSING, &
       radiance_VL, reflectance_VL, BR, nfourier, rc) bind(C, name="ref_scalar_cx")
    integer(c_int), value, intent(in) :: km, nch, nobs, nstreams, plane_parallel, nMom, nPol
    real(c_double), intent(in) :: channels(nch), ROT(km,nobs,nch), depol(nch)
    real(c_double), intent(in) :: tau(km,nch,nobs), ssa(km,nch,nobs), g(km,nch,nobs)
    real(c_double), intent(in) :: pmom(km,nch,nobs,nMom,nPol)
    real(c_double), intent(in) :: pe(km+1,nobs), he(km+1,nobs), te(km+1,nobs)
    real(c_double), intent(in) :: u10m(nobs), v10m(nobs), mr(nch)
    real(c_double), intent(in) :: solar_zenith(nobs), relat_azymuth(nobs), sensor_zenith(nobs)
    real(c_double), value, intent(in) :: MISSING
    real(c_double), intent(out) :: radiance_VL(nobs,nch), reflectance_VL(nobs,nch), BR(nobs,nch)
    integer(c_int), intent(out) :: nfourier(nobs,nch), rc

    type(VLIDORT), allocatable :: V
    type(VBRDF_Sup_Inputs), allocatable :: B
    integer :: i, j, l, n, p, ier, irc
    real*8 :: ray_l, alpha_l, tau_l, ssa_l, tau_ext, tau_scat, ssa_tot
    real*8 :: raysmom2, aerswt, rayswt, depol_ratio
    real*8, allocatable :: greekmat_total_input(:,:,:)
    real*8 :: aersmom(0:MAXMOMENTS_INPUT)
    real*8 :: deltau_vert_input(MAXLAYERS), omega_total_input(MAXLAYERS)
    real*8, parameter :: pi = 4.*atan(1.0)

    allocate(V, B)
    allocate(greekmat_total_input(0:MAXMOMENTS_INPUT,MAXLAYERS,16))
    greekmat_total_input = 0.0d0
    rc = 0
    nfourier = 0
    V%NSTREAMS = nstreams
    V%DO_PLANE_PARALLEL = (plane_parallel /= 0)
    V%NGREEK_MOMENTS_INPUT = nMom
    call VLIDORT_Init(V, km, irc)
    rc = irc
    if (rc /= 0) return

    do j = 1, nobs
      if (is_missing(solar_zenith(j),MISSING) .or. is_missing(sensor_zenith(j),MISSING) .or. &
          is_missing(relat_azymuth(j),MISSING)) then
        radiance_VL(j,:) = MISSING
        reflectance_VL(j,:) = MISSING
        BR(j,:) = MISSING
        cycle
      end if
      do i = 1, nch
        if (is_missing(u10m(j),MISSING) .or. is_missing(v10m(j),MISSING) .or. is_missing(mr(i),MISSING)) then
          radiance_VL(j,i) = MISSING
          reflectance_VL(j,i) = MISSING
          BR(j,i) = MISSING
          cycle
        end if
        call coxmunk_inputs(B, nstreams, solar_zenith(j), sensor_zenith(j), relat_azymuth(j), &
                            u10m(j), v10m(j), mr(i))
        call VBRDF_MAINMASTER(.false., 2*nstreams-1, B, V%VIO%VBRDF_Sup_Out, V%VIO%VBRDF_Sup_OutputStatus)
        BR(j,i) = V%VIO%VBRDF_Sup_Out%BS_DBOUNCE_BRDFUNC(1,1,1,1)
        ier = 0
        call set_column(V, 1, nMom, 0.0d0, solar_zenith(j), sensor_zenith(j), &
                        relat_azymuth(j), pe(:,j), he(:,j), te(:,j), km)
        V%VIO%VLIDORT_FixIn%Bool%TS_DO_LAMBERTIAN_SURFACE = .false.
        V%VIO%VLIDORT_FixIn%Optical%TS_LAMBERTIAN_ALBEDO = 0.0
        V%VIO%VLIDORT_Sup%BRDF%TS_BRDF_F_0        = V%VIO%VBRDF_Sup_Out%BS_BRDF_F_0
        V%VIO%VLIDORT_Sup%BRDF%TS_BRDF_F          = V%VIO%VBRDF_Sup_Out%BS_BRDF_F
        V%VIO%VLIDORT_Sup%BRDF%TS_USER_BRDF_F_0   = V%VIO%VBRDF_Sup_Out%BS_USER_BRDF_F_0
        V%VIO%VLIDORT_Sup%BRDF%TS_USER_BRDF_F     = V%VIO%VBRDF_Sup_Out%BS_USER_BRDF_F
        V%VIO%VLIDORT_Sup%BRDF%TS_EXACTDB_BRDFUNC = V%VIO%VBRDF_Sup_Out%BS_DBOUNCE_BRDFUNC
        V%VIO%VLIDORT_Sup%BRDF%TS_EMISSIVITY      = V%VIO%VBRDF_Sup_Out%BS_EMISSIVITY
        V%VIO%VLIDORT_Sup%BRDF%TS_USER_EMISSIVITY = V%VIO%VBRDF_Sup_Out%BS_USER_EMISSIVITY
        depol_ratio = depol(i)
        aersmom = 0.0
        raysmom2 = (1.0 - depol_ratio)/(2.0 + depol_ratio)
        do l = 1, km
          ray_l = ROT(l,j,i)
          alpha_l = 0.0
          tau_l = tau(l,i,j)
          ssa_l = ssa(l,i,j)
          tau_ext = alpha_l + ray_l + tau_l
          tau_scat = ray_l + ssa_l * tau_l
          ssa_tot = tau_scat / tau_ext
          if (ssa_tot > 0.99999) then
            ssa_tot = 0.99999
          end if
          deltau_vert_input(l) = tau_ext
          omega_total_input(l) = ssa_tot
          aerswt = ssa_l * tau_l / tau_scat
          rayswt = ray_l / tau_scat
          aersmom(0:nMom-1) = pmom(l,i,j,1:nMom,1)
          greekmat_total_input(0,l,1) = 1.0
          greekmat_total_input(1,l,1) = aersmom(1) * aerswt
          greekmat_total_input(2,l,1) = raysmom2 * rayswt + aersmom(2) * aerswt
          greekmat_total_input(3:nMom,l,1) = aersmom(3:nMom) * aerswt
        end do
        V%VIO%VLIDORT_FixIn%Optical%TS_DELTAU_VERT_INPUT = deltau_vert_input
        V%VIO%VLIDORT_ModIn%MOptical%TS_OMEGA_TOTAL_INPUT = omega_total_input
        V%VIO%VLIDORT_FixIn%Optical%TS_GREEKMAT_TOTAL_INPUT = greekmat_total_input
        call VLIDORT_MASTER(V%DO_DEBUG_INPUT, V%VIO%VLIDORT_FixIn, V%VIO%VLIDORT_ModIn, &
                            V%VIO%VLIDORT_Sup, V%VIO%VLIDORT_Out)
        if (V%VIO%VLIDORT_Out%Status%TS_STATUS_INPUTCHECK /= 0) ier = 3
        if (V%VIO%VLIDORT_Out%Status%TS_STATUS_CALCULATION /= 0) ier = 4
        radiance_VL(j,i) = V%VIO%VLIDORT_Out%Main%TS_STOKES(1,1,1,1)
        reflectance_VL(j,i) = (pi * radiance_VL(j,i)) / &
             (cos(V%VIO%VLIDORT_ModIn%MSunRays%TS_SZANGLES(1)*pi/180.0) * &
              V%VIO%VLIDORT_FixIn%SunRays%TS_FLUX_FACTOR)
        if (ier /= 0) then
          radiance_VL(j,i) = MISSING
          reflectance_VL(j,i) = MISSING
          cycle
        end if
        nfourier(j,i) = V%VIO%VLIDORT_Out%Main%TS_FOURIER_SAVED(1) + 1
      end do
    end do
    deallocate(greekmat_total_input)
    deallocate(V, B)
  end subroutine ref_scalar_cx

  subroutine ref_vector_cx(km, nch, nobs, channels, nstreams, plane_parallel, nMom, &
       nPol, ROT, depol, tau, ssa, pmom, pe, he, te, u10m, v10m, mr, &
       solar_zenith, relat_azymuth, sensor_zenith, MISSING, &
       radiance_VL, reflectance_VL, Q, U, BR, BR_Q, BR_U, nfourier, rc) bind(C, name="ref_vector_cx")
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
    integer(c_int), intent(out) :: nfourier(nobs,nch), rc

    type(VLIDORT), allocatable :: V
    type(VBRDF_Sup_Inputs), allocatable :: B
    integer :: i, j, k, l, ier, irc
    real*8 :: ray_l, tau_l, ssa_l, tau_ext, tau_scat, ssa_tot
    real*8 :: raysmom2, gammamom2, alphamom2, deltamom1, aerswt, rayswt, depol_ratio
    real*8, allocatable :: greekmat_total_input(:,:,:), aervmoms(:,:,:)
    real*8 :: rayvmoms(0:2,16)
    real*8 :: deltau_vert_input(MAXLAYERS), omega_total_input(MAXLAYERS)
    real*8, parameter :: pi = 4.*atan(1.0)

    allocate(V, B)
    allocate(greekmat_total_input(0:MAXMOMENTS_INPUT,MAXLAYERS,16))
    allocate(aervmoms(0:MAXMOMENTS_INPUT,MAXLAYERS,16))
    greekmat_total_input = 0.0d0
    radiance_VL = MISSING
    reflectance_VL = MISSING
    Q = MISSING
    U = MISSING
    BR = MISSING
    BR_Q = MISSING
    BR_U = MISSING
    rc = 0
    nfourier = 0
    V%NSTREAMS = nstreams
    V%DO_PLANE_PARALLEL = (plane_parallel /= 0)
    V%NGREEK_MOMENTS_INPUT = nMom
    call VLIDORT_Init(V, km, irc)
    rc = irc
    if (rc /= 0) return

    do j = 1, nobs
      if (is_missing(solar_zenith(j),MISSING) .or. is_missing(sensor_zenith(j),MISSING) .or. &
          is_missing(relat_azymuth(j),MISSING)) cycle
      do i = 1, nch
        if (is_missing(u10m(j),MISSING) .or. is_missing(v10m(j),MISSING) .or. is_missing(mr(i),MISSING)) cycle
        call coxmunk_inputs(B, nstreams, solar_zenith(j), sensor_zenith(j), relat_azymuth(j), &
                            u10m(j), v10m(j), mr(i))
        B%BS_NSTOKES = 3
        call VBRDF_MAINMASTER(.false., 2*nstreams-1, B, V%VIO%VBRDF_Sup_Out, V%VIO%VBRDF_Sup_OutputStatus)
        BR(j,i) = V%VIO%VBRDF_Sup_Out%BS_DBOUNCE_BRDFUNC(1,1,1,1)
        BR_Q(j,i) = 0
        BR_U(j,i) = 0
        ier = 0
        call set_column(V, 3, nMom, 0.0d0, solar_zenith(j), sensor_zenith(j), &
                        relat_azymuth(j), pe(:,j), he(:,j), te(:,j), km)
        V%VIO%VLIDORT_FixIn%Bool%TS_DO_LAMBERTIAN_SURFACE = .false.
        V%VIO%VLIDORT_FixIn%Optical%TS_LAMBERTIAN_ALBEDO = 0.0
        V%VIO%VLIDORT_Sup%BRDF%TS_BRDF_F_0        = V%VIO%VBRDF_Sup_Out%BS_BRDF_F_0
        V%VIO%VLIDORT_Sup%BRDF%TS_BRDF_F          = V%VIO%VBRDF_Sup_Out%BS_BRDF_F
        V%VIO%VLIDORT_Sup%BRDF%TS_USER_BRDF_F_0   = V%VIO%VBRDF_Sup_Out%BS_USER_BRDF_F_0
        V%VIO%VLIDORT_Sup%BRDF%TS_USER_BRDF_F     = V%VIO%VBRDF_Sup_Out%BS_USER_BRDF_F
        V%VIO%VLIDORT_Sup%BRDF%TS_EXACTDB_BRDFUNC = V%VIO%VBRDF_Sup_Out%BS_DBOUNCE_BRDFUNC
        V%VIO%VLIDORT_Sup%BRDF%TS_EMISSIVITY      = V%VIO%VBRDF_Sup_Out%BS_EMISSIVITY
        V%VIO%VLIDORT_Sup%BRDF%TS_USER_EMISSIVITY = V%VIO%VBRDF_Sup_Out%BS_USER_EMISSIVITY
        if (nMom > MAXMOMENTS_INPUT) ier = 5
        if (ier == 0) then
          depol_ratio = depol(i)
          rayvmoms = 0.0
          aervmoms = 0.0
          gammamom2 = -SQRT(6.) * (1 - depol_ratio) / (2 + depol_ratio)
          alphamom2 = 6 * (1 - depol_ratio) / (2 + depol_ratio)
          deltamom1 = 3 * (1 - 2 * depol_ratio) / (2 + depol_ratio)
          raysmom2 = (1.0 - depol_ratio)/(2.0 + depol_ratio)
          rayvmoms(0,1) = 1.0
          rayvmoms(2,1) = raysmom2
          rayvmoms(2,2) = gammamom2
          rayvmoms(2,5) = gammamom2
          rayvmoms(2,6) = alphamom2
          rayvmoms(1,16) = deltamom1
          do l = 1, km
            ray_l = ROT(l,j,i)
            tau_l = tau(l,i,j)
            ssa_l = ssa(l,i,j)
            tau_ext = ray_l + tau_l
            tau_scat = ray_l + ssa_l * tau_l
            ssa_tot = tau_scat / tau_ext
            if (ssa_tot > 0.99999) then
              ssa_tot = 0.99999
            end if
            deltau_vert_input(l) = tau_ext
            omega_total_input(l) = ssa_tot
            aervmoms(0:nMom-1,l,1)  = pmom(l,i,j,1:nMom,1)
            aervmoms(0:nMom-1,l,2)  = pmom(l,i,j,1:nMom,2)
            aervmoms(0:nMom-1,l,5)  = pmom(l,i,j,1:nMom,2)
            aervmoms(0:nMom-1,l,11) = pmom(l,i,j,1:nMom,3)
            aervmoms(0:nMom-1,l,12) = pmom(l,i,j,1:nMom,4)
            aervmoms(0:nMom-1,l,15) = -pmom(l,i,j,1:nMom,4)
            aervmoms(0:nMom-1,l,6)  = pmom(l,i,j,1:nMom,5)
            aervmoms(0:nMom-1,l,16) = pmom(l,i,j,1:nMom,6)
            aerswt = ssa_l * tau_l / tau_scat
            rayswt = ray_l / tau_scat
            do k = 1, 16
              greekmat_total_input(0:2,l,k) = rayvmoms(0:2,k) * rayswt + aervmoms(0:2,l,k) * aerswt
              greekmat_total_input(3:nMom,l,k) = aervmoms(3:nMom,l,k) * aerswt
            end do
            greekmat_total_input(0,l,1) = 1.0
          end do
          V%VIO%VLIDORT_FixIn%Optical%TS_DELTAU_VERT_INPUT = deltau_vert_input
          V%VIO%VLIDORT_ModIn%MOptical%TS_OMEGA_TOTAL_INPUT = omega_total_input
          V%VIO%VLIDORT_FixIn%Optical%TS_GREEKMAT_TOTAL_INPUT = greekmat_total_input
          call VLIDORT_MASTER(V%DO_DEBUG_INPUT, V%VIO%VLIDORT_FixIn, V%VIO%VLIDORT_ModIn, &
                              V%VIO%VLIDORT_Sup, V%VIO%VLIDORT_Out)
          if (V%VIO%VLIDORT_Out%Status%TS_STATUS_INPUTCHECK /= 0) ier = 3
          if (V%VIO%VLIDORT_Out%Status%TS_STATUS_CALCULATION /= 0) ier = 4
        end if
        if (ier /= 0) then
          radiance_VL(j,i) = MISSING
          reflectance_VL(j,i) = MISSING
          cycle
        end if
        radiance_VL(j,i) = V%VIO%VLIDORT_Out%Main%TS_STOKES(1,1,1,1)
        Q(j,i) = V%VIO%VLIDORT_Out%Main%TS_STOKES(1,1,2,1)
        U(j,i) = V%VIO%VLIDORT_Out%Main%TS_STOKES(1,1,3,1)
        reflectance_VL(j,i) = (pi * radiance_VL(j,i)) / &
             (cos(V%VIO%VLIDORT_ModIn%MSunRays%TS_SZANGLES(1)*pi/180.0) * &
              V%VIO%VLIDORT_FixIn%SunRays%TS_FLUX_FACTOR)
        nfourier(j,i) = V%VIO%VLIDORT_Out%Main%TS_FOURIER_SAVED(1) + 1
      end do
    end do
    deallocate(greekmat_total_input, aervmoms)
    deallocate(V, B)
  end subroutine ref_vector_cx

end module ref_driver_m
