! Test driver of the Fortran drop-in shim: reads a raw little-endian input
! file written by tests/test_gpu_fortran.py, calls the reference-named
! procedures (VLIDORT_Scalar_Lambert / VLIDORT_Vector_Lambert from module
! VLIDORT_LAMBERT, VLIDORT_Scalar_LandMODIS from module VLIDORT_BRDF_MODIS,
! VLIDORT_Scalar_CX from module VLIDORT_BRDF_CX,
! VLIDORT_getAOPvector with a Chem_Mie object) and writes
! their outputs to a raw file.
program shim_driver
  use Chem_MieMod
  use VLIDORT_LAMBERT
  use VLIDORT_BRDF_MODIS
  use VLIDORT_BRDF_CX
  implicit none
  character(len=512) :: fin, fout
  integer :: mode, km, nch, nobs, nstreams, nmom, npol, ipp, rc, u, i
  real*8 :: missing
  real*8, allocatable :: channels(:), rot(:,:,:), depol(:), alpha(:,:,:), tau(:,:,:), ssa(:,:,:), g(:,:,:)
  real*8, allocatable :: pmom(:,:,:,:,:), pe(:,:), he(:,:), te(:,:), albedo(:,:), sza(:), raa(:), vza(:)
  real*8, allocatable :: ff(:,:), rad(:,:), refl(:,:), q(:,:), uu(:,:), kw(:,:,:), kp(:,:,:), br(:,:)
  real*8, allocatable :: u10(:), v10(:), mr(:), brx(:,:,:)
  ! Mie mode
  integer :: nq, ntab, nrh, nbin, t, b
  type(Chem_Mie) :: mie
  real, allocatable :: chm(:), chnm(:), rh1(:), tb(:,:,:,:), tp(:,:,:,:,:), qm(:,:,:), rhc(:,:)
  real, allocatable :: otau(:,:,:), ossa(:,:,:), og(:,:,:), opm(:,:,:,:,:)
  character(len=16), allocatable :: names(:)
  character, allocatable :: vname(:,:)

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
  read(u) mode
  if (mode == 5) then
    ! VLIDORT_Scalar_CX
    read(u) km, nch, nobs, nstreams, nmom, npol, ipp
    read(u) missing
    allocate(channels(nch), rot(km,nobs,nch), depol(nch), tau(km,nch,nobs), ssa(km,nch,nobs), g(km,nch,nobs))
    allocate(pmom(km,nch,nobs,nmom,npol), pe(km+1,nobs), he(km+1,nobs), te(km+1,nobs))
    allocate(u10(nobs), v10(nobs), mr(nch), sza(nobs), raa(nobs), vza(nobs))
    allocate(rad(nobs,nch), refl(nobs,nch), brx(1,nobs,nch))
    read(u) channels, rot, depol, tau, ssa, g, pmom, pe, he, te, u10, v10, mr, sza, raa, vza
    close(u)
    call VLIDORT_Scalar_CX(km, nch, nobs, channels, nstreams, ipp /= 0, nmom, npol, rot, depol, tau, ssa, &
                           g, pmom, pe, he, te, u10, v10, mr, sza, raa, vza, missing, 0, rad, refl, brx, rc)
    open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
    write(u) rc, rad, refl, brx
    close(u)
  else if (mode == 4) then
    ! VLIDORT_Scalar_LandMODIS
    read(u) km, nch, nobs, nstreams, nmom, npol, ipp
    read(u) missing
    allocate(channels(nch), rot(km,nobs,nch), depol(nch), tau(km,nch,nobs), ssa(km,nch,nobs), g(km,nch,nobs))
    allocate(pmom(km,nch,nobs,nmom,npol), pe(km+1,nobs), he(km+1,nobs), te(km+1,nobs))
    allocate(kw(3,nch,nobs), kp(2,nch,nobs), sza(nobs), raa(nobs), vza(nobs))
    allocate(rad(nobs,nch), refl(nobs,nch), br(nobs,nch))
    read(u) channels, rot, depol, tau, ssa, g, pmom, pe, he, te, kw, kp, sza, raa, vza
    close(u)
    call VLIDORT_Scalar_LandMODIS(km, nch, nobs, channels, nstreams, ipp /= 0, nmom, npol, rot, depol, tau, ssa, &
                                  g, pmom, pe, he, te, kw, kp, sza, raa, vza, missing, 0, rad, refl, br, rc)
    open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
    write(u) rc, rad, refl, br
    close(u)
  else if (mode == 1 .or. mode == 2) then
    read(u) km, nch, nobs, nstreams, nmom, npol, ipp
    read(u) missing
    allocate(channels(nch), rot(km,nobs,nch), depol(nch), alpha(km,nobs,nch), tau(km,nch,nobs), ssa(km,nch,nobs))
    allocate(g(km,nch,nobs), pmom(km,nch,nobs,nmom,npol), pe(km+1,nobs), he(km+1,nobs), te(km+1,nobs))
    allocate(albedo(nobs,nch), sza(nobs), raa(nobs), vza(nobs), ff(nch,nobs))
    allocate(rad(nobs,nch), refl(nobs,nch), q(nobs,nch), uu(nobs,nch))
    read(u) channels, rot, depol, tau, ssa, g, pmom, pe, he, te, albedo, sza, raa, vza, ff
    close(u)
    alpha = 0.0d0
    if (mode == 1) then
      call VLIDORT_Scalar_Lambert(km, nch, nobs, channels, nstreams, ipp /= 0, nmom, npol, rot, depol, tau, ssa, &
                                  g, pmom, pe, he, te, albedo, sza, raa, vza, missing, 0, rad, refl, rc)
      q = 0.0d0
      uu = 0.0d0
    else
      call VLIDORT_Vector_Lambert(km, nch, nobs, channels, nstreams, ipp /= 0, nmom, npol, rot, depol, alpha, tau, &
                                  ssa, pmom, pe, he, te, albedo, sza, raa, vza, ff, missing, 0, rad, refl, q, uu, rc)
    end if
    open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
    write(u) rc, rad, refl, q, uu
    close(u)
  else
    ! Chem_Mie tables + VLIDORT_getAOPvector
    read(u) nch, ntab, km, nobs, nq, nmom, npol
    allocate(chm(nch), chnm(nch))
    read(u) chm, chnm
    mie = Chem_MieCreate(nch, chm, rc)
    do t = 1, ntab
      read(u) nrh, nbin
      allocate(rh1(nrh), tb(nrh,nch,nbin,6), tp(nrh,nch,nbin,nmom,npol), names(nbin))
      read(u) rh1, tb, tp
      do b = 1, nbin
        read(u) names(b)
      end do
      call Chem_MieAddTable(mie, nrh, nbin, nmom, npol, rh1, tb(:,:,:,1), tb(:,:,:,2), tb(:,:,:,3), tb(:,:,:,4), &
                            tb(:,:,:,5), tb(:,:,:,6), tp, names, rc)
      deallocate(rh1, tb, tp, names)
    end do
    allocate(vname(nq,16), qm(km,nq,nobs), rhc(km,nobs))
    read(u) vname, qm, rhc
    close(u)
    allocate(otau(km,nch,nobs), ossa(km,nch,nobs), og(km,nch,nobs), opm(km,nch,nobs,nmom,npol))
    call VLIDORT_getAOPvector(mie, km, nobs, nch, nq, chnm, vname, 0, qm, rhc, nmom, npol, otau, ossa, og, opm, rc)
    i = Chem_MieQueryIdx(mie, 'SS::ss003', rc)
    open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
    write(u) rc, i, otau, ossa, og, opm
    close(u)
    call Chem_MieDestroy(mie, rc)
  end if
end program shim_driver
