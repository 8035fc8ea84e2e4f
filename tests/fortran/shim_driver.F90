! Test driver of the Fortran drop-in shim: reads a raw little-endian input
! file written by tests/test_gpu_fortran.py, calls the reference-named
! procedures (VLIDORT_Scalar_Lambert / VLIDORT_Vector_Lambert from module
! VLIDORT_LAMBERT, single- and multi-geometry, with DO_2OS_CORRECTION,
! VLIDORT_Scalar_LandMODIS from module VLIDORT_BRDF_MODIS,
! VLIDORT_Scalar_CX and VLIDORT_Vector_CX from module VLIDORT_BRDF_CX,
! VLIDORT_getAOPvector with a Chem_Mie object, Chem_MieCreate from a
! resource file and the Chem_MieQuery family) and writes their outputs to a
! raw file.

! Optics-table reader for Chem_MieCreate(rcfile): the files are raw
! little-endian streams (int32 nch, nrh, nbin, nmom, npol, then float64
! lambda, rh, reff, bext, bsca, bbck, g, pback, pmom, gf, rhop, refr, refi in
! their Fortran layouts).  The arrays live until the next call, which is
! after the library has taken its copy.
module raw_table_reader
  use iso_c_binding
  use Chem_MieMod, only: aa_mie_table_file
  implicit none
  real(c_double), allocatable, target, save :: lam(:), rh(:), reff(:,:), bext(:,:,:), bsca(:,:,:), bbck(:,:,:)
  real(c_double), allocatable, target, save :: g(:,:,:), pback(:,:,:,:), pmom(:,:,:,:,:), gf(:,:), rhop(:,:)
  real(c_double), allocatable, target, save :: refr(:,:,:), refi(:,:,:)
  integer, save :: ncalls = 0
contains
  integer(c_int) function read_raw_table(user, path, t) bind(C) result(rc)
    type(c_ptr), value :: user
    type(c_ptr), value :: path
    type(aa_mie_table_file) :: t
    character(kind=c_char), pointer :: cp(:)
    character(len=512) :: fname
    integer :: n, u, ios, nch, nrh, nbin, nmom, npol
    ncalls = ncalls + 1
    call c_f_pointer(path, cp, (/512/))
    fname = ' '
    do n = 1, 512
      if (cp(n) == c_null_char) exit
      fname(n:n) = cp(n)
    end do
    open(newunit=u, file=trim(fname), access='stream', form='unformatted', status='old', iostat=ios)
    if (ios /= 0) then
      rc = 1
      return
    end if
    read(u) nch, nrh, nbin, nmom, npol
    if (allocated(lam)) deallocate(lam, rh, reff, bext, bsca, bbck, g, pback, pmom, gf, rhop, refr, refi)
    allocate(lam(nch), rh(nrh), reff(nrh,nbin), bext(nch,nrh,nbin), bsca(nch,nrh,nbin), bbck(nch,nrh,nbin))
    allocate(g(nch,nrh,nbin), pback(nch,nrh,nbin,npol), pmom(nch,nrh,nbin,nmom,npol), gf(nrh,nbin))
    allocate(rhop(nrh,nbin), refr(nch,nrh,nbin), refi(nch,nrh,nbin))
    read(u) lam, rh, reff, bext, bsca, bbck, g, pback, pmom, gf, rhop, refr, refi
    close(u)
    t%nch = nch
    t%nrh = nrh
    t%nbin = nbin
    t%nmom = nmom
    t%npol = npol
    t%lambda = c_loc(lam)
    t%rh = c_loc(rh)
    t%reff = c_loc(reff)
    t%bext = c_loc(bext)
    t%bsca = c_loc(bsca)
    t%bbck = c_loc(bbck)
    t%g = c_loc(g)
    t%pback = c_loc(pback)
    t%pmom = c_loc(pmom)
    t%gf = c_loc(gf)
    t%rhop = c_loc(rhop)
    t%refr = c_loc(refr)
    t%refi = c_loc(refi)
    rc = 0
  end function read_raw_table
end module raw_table_reader

program shim_driver
  use Chem_MieMod
  use VLIDORT_LAMBERT
  use VLIDORT_BRDF_MODIS
  use VLIDORT_BRDF_CX
  use VLIDORT_BRDF_MODIS_BPDF
  use raw_table_reader, only: read_raw_table, ncalls
  implicit none
  character(len=512) :: fin, fout
  integer :: mode, km, nch, nobs, nstreams, nmom, npol, ipp, rc, u, i
  real*8 :: missing
  real*8, allocatable :: channels(:), rot(:,:,:), depol(:), alpha(:,:,:), tau(:,:,:), ssa(:,:,:), g(:,:,:)
  real*8, allocatable :: pmom(:,:,:,:,:), pe(:,:), he(:,:), te(:,:), albedo(:,:), sza(:), raa(:), vza(:)
  real*8, allocatable :: ff(:,:), rad(:,:), refl(:,:), q(:,:), uu(:,:), kw(:,:,:), kp(:,:,:), br(:,:)
  real*8, allocatable :: u10(:), v10(:), mr(:), brx(:,:,:), brq(:,:), bru(:,:), bp(:,:,:)
  ! Mie mode
  integer :: nq, ntab, nrh, nbin, t, b
  type(Chem_Mie) :: mie
  real, allocatable :: chm(:), chnm(:), rh1(:), tb(:,:,:,:), tp(:,:,:,:,:), qm(:,:,:), rhc(:,:)
  real, allocatable :: otau(:,:,:), ossa(:,:,:), og(:,:,:), opm(:,:,:,:,:)
  character(len=16), allocatable :: names(:)
  character, allocatable :: vname(:,:)
  ! MultiGeom mode
  integer :: ngeom
  real*8, allocatable :: sza2(:,:), raa2(:,:), vza2(:,:), alb3(:,:,:), rad3(:,:,:), refl3(:,:,:), q3(:,:,:), u3(:,:,:)
  ! Chem_Mie resource-file mode
  integer :: nobsq, idx, idx2, rcs(8)
  real, allocatable :: qv(:), rhv(:), outs(:,:), pmq(:,:,:), tl(:), q2(:,:), rh2(:,:), t3(:,:,:), s3(:,:,:), a3(:,:,:)
  real, allocatable :: bc1(:)

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
  read(u) mode
  if (mode == 8) then
    ! Chem_MieCreate(rcfile) with the registry Chem_MieRegistry.rc in the working
    ! directory, then the queries of bin idx of the table Chem_MieQueryIdx selects
    read(u) nobsq
    allocate(qv(nobsq), rhv(nobsq))
    read(u) qv, rhv
    close(u)
    call Chem_MieSetTableReader(read_raw_table)
    mie = Chem_MieCreate('Chem_Mie.rc', rc)
    rcs = 0
    rcs(1) = rc
    if (rc /= 0) then
      open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
      write(u) rcs, ncalls
      close(u)
      stop
    end if
    idx = Chem_MieQueryIdx(mie, 'SS::ss002', rcs(2))
    allocate(outs(nobsq, 16), pmq(mie%nMom, mie%nPol, nobsq))
    outs = -1.0
    do i = 1, nobsq
      ! channel 2; every optional output
      call Chem_MieQuery(mie, idx, 2.0, qv(i), rhv(i), tau=outs(i,1), ssa=outs(i,2), gasym=outs(i,3), &
                         bext=outs(i,4), bsca=outs(i,5), bbck=outs(i,6), reff=outs(i,7), pmom=pmq(:,:,i), &
                         p11=outs(i,8), p22=outs(i,9), gf=outs(i,10), rhod=outs(i,11), vol=outs(i,12), &
                         area=outs(i,13), refr=outs(i,14), refi=outs(i,15), rc=rcs(3))
    end do
    ! by name: registry position of du001 (1) as a bin of the selected table
    call Chem_MieQuery(mie, 'GOCART::du001', 1.0, qv(1), rhv(1), tau=outs(1,16), rc=rcs(4))
    allocate(tl(nobsq))
    call Chem_MieQueryTauList(mie, idx, 1.0, qv, rhv, tl, rcs(5))
    allocate(q2(2, nobsq / 2), rh2(2, nobsq / 2), t3(2, nobsq / 2, 2), s3(2, nobsq / 2, 2), a3(2, nobsq / 2, 2))
    q2 = reshape(qv(1:nobsq / 2 * 2), (/2, nobsq / 2/))
    rh2 = reshape(rhv(1:nobsq / 2 * 2), (/2, nobsq / 2/))
    call Chem_MieQueryAllBand3D(mie, idx, 2, 0, q2, rh2, t3, s3, a3, rcs(6))
    ! a tracer of another table, one the registry lists that has no bin (DMS)
    idx2 = Chem_MieQueryIdx(mie, 'BCphilic', rcs(7))
    allocate(bc1(2))
    call Chem_MieQuery(mie, idx2, 1.0, qv(1), rhv(1), tau=bc1(1), ssa=bc1(2))
    i = Chem_MieQueryIdx(mie, 'DMS', rcs(8))
    open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
    write(u) rcs, ncalls, idx, idx2, i, mie%nch, mie%nMom, mie%nPol, mie%nq
    write(u) mie%channels, outs, pmq, tl, t3, s3, a3, bc1
    close(u)
    call Chem_MieDestroy(mie, rc)
  else if (mode == 6 .or. mode == 7) then
    ! VLIDORT_Vector_Lambert: MultiGeom (6) or DO_2OS_CORRECTION (7)
    read(u) km, nch, nobs, nstreams, nmom, npol, ipp, ngeom
    read(u) missing
    allocate(channels(nch), rot(km,nobs,nch), depol(nch), alpha(km,nobs,nch), tau(km,nch,nobs), ssa(km,nch,nobs))
    allocate(pmom(km,nch,nobs,nmom,npol), pe(km+1,nobs), he(km+1,nobs), te(km+1,nobs), ff(nch,nobs))
    allocate(sza2(nobs,ngeom), raa2(nobs,ngeom), vza2(nobs,ngeom), alb3(nobs,nch,ngeom))
    allocate(rad3(nobs,nch,ngeom), refl3(nobs,nch,ngeom), q3(nobs,nch,ngeom), u3(nobs,nch,ngeom))
    read(u) channels, rot, depol, tau, ssa, pmom, pe, he, te, alb3, sza2, raa2, vza2, ff
    close(u)
    alpha = 0.0d0
    if (mode == 6) then
      call VLIDORT_Vector_Lambert(km, nch, nobs, ngeom, channels, nstreams, ipp /= 0, nmom, npol, rot, depol, &
                                  alpha, tau, ssa, pmom, pe, he, te, alb3, sza2, raa2, vza2, ff, missing, 0, &
                                  rad3, refl3, q3, u3, rc)
    else
      call VLIDORT_Vector_Lambert(km, nch, nobs, channels, nstreams, ipp /= 0, nmom, npol, rot, depol, alpha, &
                                  tau, ssa, pmom, pe, he, te, alb3(:,:,1), sza2(:,1), raa2(:,1), vza2(:,1), ff, &
                                  missing, 0, rad3(:,:,1), refl3(:,:,1), q3(:,:,1), u3(:,:,1), rc, &
                                  DO_2OS_CORRECTION=.true.)
    end if
    open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
    write(u) rc, rad3, refl3, q3, u3
    close(u)
  else if (mode == 10) then
    ! VLIDORT_Vector_LandMODIS_BPDF
    read(u) km, nch, nobs, nstreams, nmom, npol, ipp
    read(u) missing
    allocate(channels(nch), rot(km,nobs,nch), depol(nch), tau(km,nch,nobs), ssa(km,nch,nobs))
    allocate(pmom(km,nch,nobs,nmom,npol), pe(km+1,nobs), he(km+1,nobs), te(km+1,nobs))
    allocate(kw(3,nch,nobs), kp(2,nch,nobs), bp(3,nch,nobs), sza(nobs), raa(nobs), vza(nobs))
    allocate(rad(nobs,nch), refl(nobs,nch), q(nobs,nch), uu(nobs,nch), br(nobs,nch), brq(nobs,nch), bru(nobs,nch))
    read(u) channels, rot, depol, tau, ssa, pmom, pe, he, te, kw, kp, bp, sza, raa, vza
    close(u)
    call VLIDORT_Vector_LandMODIS_BPDF(km, nch, nobs, channels, nstreams, ipp /= 0, nmom, npol, rot, depol, &
                                       tau, ssa, pmom, pe, he, te, kw, kp, bp, sza, raa, vza, missing, 0, &
                                       rad, refl, br, q, uu, brq, bru, rc)
    open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
    write(u) rc, rad, refl, q, uu, br, brq, bru
    close(u)
  else if (mode == 9) then
    ! VLIDORT_Vector_CX (generic -> SingleGeom)
    read(u) km, nch, nobs, nstreams, nmom, npol, ipp
    read(u) missing
    allocate(channels(nch), rot(km,nobs,nch), depol(nch), tau(km,nch,nobs), ssa(km,nch,nobs))
    allocate(pmom(km,nch,nobs,nmom,npol), pe(km+1,nobs), he(km+1,nobs), te(km+1,nobs))
    allocate(u10(nobs), v10(nobs), mr(nch), sza(nobs), raa(nobs), vza(nobs))
    allocate(rad(nobs,nch), refl(nobs,nch), q(nobs,nch), uu(nobs,nch), br(nobs,nch), brq(nobs,nch), bru(nobs,nch))
    read(u) channels, rot, depol, tau, ssa, pmom, pe, he, te, u10, v10, mr, sza, raa, vza
    close(u)
    call VLIDORT_Vector_CX(km, nch, nobs, channels, nstreams, ipp /= 0, nmom, npol, rot, depol, tau, ssa, &
                           pmom, pe, he, te, u10, v10, mr, sza, raa, vza, missing, 0, rad, refl, q, uu, &
                           br, brq, bru, rc)
    open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
    write(u) rc, rad, refl, q, uu, br, brq, bru
    close(u)
  else if (mode == 5) then
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
