! Module Chem_MieMod (src/Shared/GMAO_Shared/Chem_Base/Chem_MieMod.F90:16-1262)
! over the GPU-resident table set of the C-ABI: type(Chem_Mie) with its public
! components (nch, nMom, nPol, channels, rcfile, nq, vname, vtableUse), the
! constructor from resource files, Chem_MieQueryIdx, the Chem_MieQuery
! generic (by tracer index or name, every output optional),
! Chem_MieQueryTauList, Chem_MieQueryAllBand3D/4D and Chem_MieDestroy.
!
! Chem_MieCreate(rcfile, rc[, chemReg]) reads the resource file and the
! chemistry registry as the reference does (Chem_MieCreateFromRC, :124-411;
! without chemReg the registry file is Chem_MieRegistry.rc, else
! Chem_Registry.rc, in the working directory).  The seven optics-table files
! it names are read by the reader registered with Chem_MieSetTableReader
! (the reference reads them as netCDF; this shim has no file format of its
! own): an interoperable function filling an aa_mie_table_file with the
! file's arrays.  Without a registered reader Chem_MieCreate returns rc = 2.
! Extensions: Chem_MieCreate(nch, channels_m, rc) and Chem_MieAddTable build
! a table set from tables already on the channels.
!
! The reference stops with die() on resource-file errors; this module returns
! them in rc instead (1 resource file, 2 table reader, 3 registry).
module Chem_MieMod
  use iso_c_binding
  use aeroapps_amd_capi
  use Chem_RegistryMod
  implicit none
  private
  public Chem_Mie, Chem_MieTable, Chem_MieCreate, Chem_MieDestroy, Chem_MieQuery, Chem_MieQueryTauList
  public Chem_MieQueryAllBand3D, Chem_MieQueryAllBand4D, Chem_MieQueryIdx
  public Chem_MieAddTable, Chem_MieSetTableReader, aa_mie_table_file

  ! What vtableUse exposes of a table (Chem_MieTable, CB/Chem_MieTableMod.F90:55-83):
  ! its dimensions; tr0 is the library's index of its first bin.
  type Chem_MieTable
    integer :: nch = 0, nrh = 0, nbin = 0, nMom = 0, nPol = 0
    integer :: tr0 = -1
  end type Chem_MieTable

  type Chem_Mie
    integer :: nch = 0
    integer :: nMom = 0
    integer :: nPol = 0
    real, pointer :: channels(:) => null()
    character(len=255) :: rcfile = ' '
    integer :: nq = 0
    character(len=255), pointer :: vname(:) => null()
    type(Chem_MieTable), pointer :: vtableUse => null()
    type(c_ptr) :: handle = c_null_ptr
  end type Chem_Mie

  ! The arrays of one optics-table file (include/aeroapps_amd.h), float64,
  ! Fortran layouts: lambda(nch) [m], rh(nrh), reff/gf/rhop(nrh,nbin),
  ! bext/bsca/bbck/g/refr/refi(nch,nrh,nbin), pback(nch,nrh,nbin,npol),
  ! pmom(nch,nrh,nbin,nmom,npol); absent ones c_null_ptr.
  type, bind(C) :: aa_mie_table_file
    integer(c_int) :: nch, nrh, nbin, nmom, npol
    type(c_ptr) :: lambda, rh, reff, bext, bsca, bbck, g, pback, pmom, gf, rhop, refr, refi
  end type aa_mie_table_file

  abstract interface
    integer(c_int) function mie_table_reader(user, path, table) bind(C)
      import :: c_int, c_ptr, aa_mie_table_file
      type(c_ptr), value :: user
      type(c_ptr), value :: path            ! NUL-terminated file name
      type(aa_mie_table_file) :: table
    end function
  end interface

  interface Chem_MieCreate
    module procedure Chem_MieCreateFromRC
    module procedure Chem_MieCreateChannels
  end interface

  interface Chem_MieQuery
    module procedure Chem_MieQueryByInt
    module procedure Chem_MieQueryByChar
  end interface

  type(c_funptr), save :: g_reader = c_null_funptr
  type(c_ptr), save :: g_user = c_null_ptr

contains

  ! Registers the optics-table file reader used by Chem_MieCreate(rcfile, ...).
  subroutine Chem_MieSetTableReader(reader, user)
    procedure(mie_table_reader) :: reader
    type(c_ptr), optional, intent(in) :: user
    g_reader = c_funloc(reader)
    g_user = c_null_ptr
    if (present(user)) g_user = user
  end subroutine Chem_MieSetTableReader

  function cstr(s) result(c)
    character(len=*), intent(in) :: s
    character(kind=c_char) :: c(len_trim(s) + 1)
    integer :: i
    do i = 1, len_trim(s)
      c(i) = s(i:i)
    end do
    c(len_trim(s) + 1) = c_null_char
  end function cstr

  ! Chem_MieCreateFromRC (CB/Chem_MieMod.F90:124-411).
  type(Chem_Mie) function Chem_MieCreateFromRC(rcfile, rc, chemReg) result(this)
    character(len=*), intent(in) :: rcfile
    integer, intent(out) :: rc
    type(Chem_Registry), target, optional, intent(in) :: chemReg
    type(Chem_Registry), pointer :: reg
    type(Chem_Registry), target :: own
    logical :: fexists
    character(kind=c_char), allocatable :: names(:)
    integer(c_int), allocatable :: sp(:)
    integer(c_int) :: e, crc, nch, nmom, npol, ntr, tb, bin, nbin, nrh, tm, tp
    integer :: iq, c
    rc = 0
    if (present(chemReg)) then
      reg => chemReg
    else
      inquire(file='Chem_MieRegistry.rc', exist=fexists)
      if (fexists) then
        own = Chem_RegistryCreate(rc, rcfile='Chem_MieRegistry.rc')
      else
        own = Chem_RegistryCreate(rc, rcfile='Chem_Registry.rc')
      end if
      if (rc /= 0) then
        rc = 3
        return
      end if
      reg => own
    end if
    if (.not. c_associated(g_reader)) then
      rc = 2
      return
    end if
    allocate(names(16 * max(reg%nq, 1)), sp(max(reg%nq, 1)))
    do iq = 1, reg%nq
      do c = 1, 16
        names((iq - 1) * 16 + c) = reg%vname(iq)(c:c)
      end do
      sp(iq) = reg%species(iq)
    end do
    e = aa_mie_create_rc(this%handle, cstr(rcfile), reg%nq, names, sp, g_reader, g_user, crc)
    rc = crc
    if (e /= 0) rc = 1000 - e
    this%nq = reg%nq
    if (.not. present(chemReg)) call Chem_RegistryDestroy(own, crc)
    if (rc /= 0) return
    this%rcfile = rcfile
    e = aa_mie_dims(this%handle, nch, nmom, npol, ntr)
    this%nch = nch
    this%nMom = nmom
    allocate(this%channels(nch))
    e = aa_mie_channels(this%handle, this%channels)
    ! nPol of the dust table (:310)
    e = aa_mie_tracer_table(this%handle, 0, tb, bin, nbin, nrh, tm, tp)
    this%nPol = tp
    allocate(this%vname(max(this%nq, 1)))
    do iq = 1, this%nq
      this%vname(iq) = ' '
      do c = 1, 16
        this%vname(iq)(c:c) = names((iq - 1) * 16 + c)
      end do
    end do
  end function Chem_MieCreateFromRC

  ! Extension: empty table set on nch channels given in metres.
  type(Chem_Mie) function Chem_MieCreateChannels(nch, channels_m, rc) result(this)
    integer, intent(in) :: nch
    real, intent(in) :: channels_m(nch)
    integer, intent(out) :: rc
    real(c_float), allocatable :: ch(:)
    allocate(ch(nch))
    ch = channels_m
    rc = aa_mie_create(this%handle, nch, ch)
    this%nch = nch
    allocate(this%channels(nch))
    this%channels = channels_m
  end function Chem_MieCreateChannels

  ! Extension: one species table already on the channels, bext..refi(nrh,nch,nbin),
  ! pmom(nrh,nch,nbin,nmom,npol), bin_names(nbin) the tracer names it serves.
  subroutine Chem_MieAddTable(this, nrh, nbin, nmom, npol, rh, bext, bsca, bbck, g, refr, refi, pmom, &
                              bin_names, rc)
    type(Chem_Mie), intent(inout) :: this
    integer, intent(in) :: nrh, nbin, nmom, npol
    real, intent(in) :: rh(nrh), bext(nrh,this%nch,nbin), bsca(nrh,this%nch,nbin), bbck(nrh,this%nch,nbin)
    real, intent(in) :: g(nrh,this%nch,nbin), refr(nrh,this%nch,nbin), refi(nrh,this%nch,nbin)
    real, intent(in) :: pmom(nrh,this%nch,nbin,nmom,npol)
    character(len=16), intent(in) :: bin_names(nbin)
    integer, intent(out) :: rc
    character(kind=c_char) :: names(16*nbin)
    integer :: b, c
    do b = 1, nbin
      do c = 1, 16
        names((b-1)*16 + c) = bin_names(b)(c:c)
      end do
    end do
    rc = aa_mie_add_table(this%handle, nrh, nbin, nmom, npol, rh, bext, bsca, bbck, g, refr, refi, pmom, names)
    if (this%nMom == 0 .or. nmom < this%nMom) this%nMom = nmom
    if (npol > this%nPol) this%nPol = npol
  end subroutine Chem_MieAddTable

  ! Chem_MieQueryIdx (CB/Chem_MieMod.F90:720-777): the tracer's bin index in
  ! its table (vindex), -1 when the name is not served; a found name makes
  ! that table this%vtableUse, which the queries below read.
  integer function Chem_MieQueryIdx(this, qname, rc) result(idx)
    type(Chem_Mie), intent(inout) :: this
    character(len=*), intent(in) :: qname
    integer, optional, intent(out) :: rc
    integer(c_int) :: k, tb, bin, nbin, nrh, nmom, npol, e
    k = aa_mie_query_idx(this%handle, cstr(qname))
    idx = -1
    if (k >= 0) then
      e = aa_mie_tracer_table(this%handle, k, tb, bin, nbin, nrh, nmom, npol)
      if (e == 0) then
        if (.not. associated(this%vtableUse)) allocate(this%vtableUse)
        this%vtableUse%nch = this%nch
        this%vtableUse%nrh = nrh
        this%vtableUse%nbin = nbin
        this%vtableUse%nMom = nmom
        this%vtableUse%nPol = npol
        this%vtableUse%tr0 = k - bin
        idx = bin + 1
      end if
    end if
    if (present(rc)) then
      rc = 0
      if (idx == -1) rc = 1
    end if
  end function Chem_MieQueryIdx

  ! Library index of bin idx of vtableUse, -1 when there is none.
  integer function lib_index(this, idx) result(k)
    type(Chem_Mie), intent(in) :: this
    integer, intent(in) :: idx
    k = -1
    if (.not. associated(this%vtableUse)) return
    if (idx < 1 .or. idx > this%vtableUse%nbin) return
    k = this%vtableUse%tr0 + idx - 1
  end function lib_index

  ! Chem_MieQueryByInt (CB/Chem_MieMod.F90:789-965): bin idx of vtableUse.
  ! rhop is never assigned by the reference and is left untouched here.
  subroutine Chem_MieQueryByInt(this, idx, channel, q_mass, rh, tau, ssa, gasym, bext, bsca, bbck, &
                                reff, pmom, p11, p22, gf, rhop, rhod, vol, area, refr, refi, rc)
    type(Chem_Mie), target, intent(in) :: this
    integer, intent(in) :: idx
    real, intent(in) :: channel, q_mass, rh
    real, optional, target, intent(out) :: tau, ssa, gasym, bext, bsca, bbck, reff
    real, optional, intent(out) :: pmom(:,:)
    real, optional, target, intent(out) :: p11, p22, gf, rhop, rhod, vol, area, refr, refi
    integer, optional, intent(out) :: rc
    real(c_float), target :: q1(1), r1(1)
    real(c_float), allocatable, target :: pm(:,:)
    integer(c_int) :: k, e, nm, np
    type(c_ptr) :: ppm
    if (present(rc)) rc = 0
    k = lib_index(this, idx)
    if (k < 0) then
      if (present(rc)) rc = 1
      return
    end if
    q1(1) = q_mass
    r1(1) = rh
    nm = 0
    np = 1
    ppm = c_null_ptr
    if (present(pmom)) then
      nm = size(pmom, 1)
      np = size(pmom, 2)
      allocate(pm(nm, np))
      ppm = c_loc(pm)
    end if
    e = aa_mie_query(this%handle, k, channel, 1, q1, r1, opt(tau), opt(ssa), opt(gasym), opt(bext), opt(bsca), &
                     opt(bbck), opt(refr), opt(refi), nm, np, ppm)
    if (e == 0 .and. present(pmom)) pmom = pm
    if (e == 0 .and. (present(reff) .or. present(p11) .or. present(p22) .or. present(gf) .or. present(rhod) &
                      .or. present(vol) .or. present(area))) &
      e = aa_mie_query_extra(this%handle, k, channel, 1, r1, opt(reff), opt(p11), opt(p22), opt(gf), opt(rhod), &
                             opt(vol), opt(area))
    if (e /= 0 .and. present(rc)) rc = 1000 - e
  contains
    type(c_ptr) function opt(x)
      real, optional, target, intent(in) :: x
      opt = c_null_ptr
      if (present(x)) opt = c_loc(x)
    end function opt
  end subroutine Chem_MieQueryByInt

  ! Chem_MieQueryByChar (CB/Chem_MieMod.F90:1021-1069) by tracer name, as the
  ! reference does it: the name (qualifier "XX::" removed, case ignored) is
  ! found among the registry tracers this%vname, and its position iq there is
  ! queried as bin iq of the table vtableUse last selected by
  ! Chem_MieQueryIdx.  rc = 1 where there is no such bin (no table selected,
  ! or iq beyond its bins: the reference reads outside the table there).
  subroutine Chem_MieQueryByChar(this, idx, channel, q_mass, rh, tau, ssa, gasym, bext, bsca, bbck, &
                                 reff, pmom, p11, p22, rc)
    type(Chem_Mie), target, intent(inout) :: this
    character(len=*), intent(in) :: idx
    real, intent(in) :: channel, q_mass, rh
    real, optional, intent(out) :: tau, ssa, gasym, bext, bsca, bbck, reff
    real, optional, intent(out) :: pmom(:,:)
    real, optional, intent(out) :: p11, p22
    integer, optional, intent(out) :: rc
    character(len=255) :: name
    integer :: iq, i
    if (present(rc)) rc = 0
    name = idx
    i = index(name, '::')
    if (i > 0) name = name(i + 2:)
    do iq = 1, this%nq
      if (upper_(trim(name)) == upper_(trim(this%vname(iq)))) then
        call Chem_MieQueryByInt(this, iq, channel, q_mass, rh, tau=tau, ssa=ssa, gasym=gasym, bext=bext, &
                                bsca=bsca, bbck=bbck, reff=reff, pmom=pmom, p11=p11, p22=p22, rc=rc)
        if (present(rc)) then
          if (rc /= 0) return
        end if
      end if
    end do
  contains
    function upper_(s) result(u)
      character(len=*), intent(in) :: s
      character(len=len(s)) :: u
      integer :: k
      u = s
      do k = 1, len(s)
        if (u(k:k) >= 'a' .and. u(k:k) <= 'z') u(k:k) = achar(iachar(u(k:k)) - 32)
      end do
    end function upper_
  end subroutine Chem_MieQueryByChar

  ! Chem_MieQueryTauList (CB/Chem_MieMod.F90:968-1018): tau(i) of bin idx of
  ! vtableUse for every (q_mass(i), rh(i)).
  subroutine Chem_MieQueryTauList(this, idx, channel, q_mass, rh, tau, rc)
    type(Chem_Mie), target, intent(in) :: this
    integer, intent(in) :: idx
    real, intent(in) :: channel
    real, intent(in) :: q_mass(:), rh(:)
    real, intent(out) :: tau(:)
    integer, optional, intent(out) :: rc
    real(c_float), allocatable, target :: t(:)
    integer(c_int) :: k, e
    if (present(rc)) rc = 0
    k = lib_index(this, idx)
    if (k < 0) then
      if (present(rc)) rc = 1
      return
    end if
    allocate(t(size(tau)))
    e = aa_mie_query(this%handle, k, channel, size(tau), q_mass, rh, c_loc(t), c_null_ptr, c_null_ptr, &
                     c_null_ptr, c_null_ptr, c_null_ptr, c_null_ptr, c_null_ptr, 0, 1, c_null_ptr)
    tau = t
    if (e /= 0 .and. present(rc)) rc = 1000 - e
  end subroutine Chem_MieQueryTauList

  ! One band of Chem_MieQueryAllBand3D/4D: tau, ssa, asymmetry of n elements.
  subroutine all_band_(this, k, ich, n, q, r, tau, ssa, asy, e)
    type(Chem_Mie), intent(in) :: this
    integer(c_int), intent(in) :: k, n
    integer, intent(in) :: ich
    real(c_float), intent(in) :: q(n), r(n)
    real(c_float), target, intent(out) :: tau(n), ssa(n), asy(n)
    integer(c_int), intent(out) :: e
    e = aa_mie_query(this%handle, k, real(ich), n, q, r, c_loc(tau), c_loc(ssa), c_loc(asy), c_null_ptr, &
                     c_null_ptr, c_null_ptr, c_null_ptr, c_null_ptr, 0, 1, c_null_ptr)
  end subroutine all_band_

  ! Chem_MieQueryAllBand3D (CB/Chem_MieMod.F90:1071-1164): channels
  ! offset+1 .. offset+nbands of bin idx of vtableUse.
  subroutine Chem_MieQueryAllBand3D(this, idx, nbands, offset, q_mass, rh, tau, ssa, asy, rc)
    type(Chem_Mie), target, intent(in) :: this
    integer, intent(in) :: idx, nbands, offset
    real, intent(in) :: q_mass(:,:), rh(:,:)
    real, intent(out) :: tau(:,:,:), ssa(:,:,:), asy(:,:,:)
    integer, optional, intent(out) :: rc
    real(c_float), allocatable :: q(:), r(:), t(:), s(:), a(:)
    integer(c_int) :: k, e, n
    integer :: ib
    if (present(rc)) rc = 0
    k = lib_index(this, idx)
    if (k < 0) then
      if (present(rc)) rc = 1
      return
    end if
    n = size(rh)
    q = reshape(q_mass, (/n/))
    r = reshape(rh, (/n/))
    allocate(t(n), s(n), a(n))
    do ib = 1, nbands
      call all_band_(this, k, ib + offset, n, q, r, t, s, a, e)
      if (e /= 0) then
        if (present(rc)) rc = 1000 - e
        return
      end if
      tau(:,:,ib) = reshape(t, shape(rh))
      ssa(:,:,ib) = reshape(s, shape(rh))
      asy(:,:,ib) = reshape(a, shape(rh))
    end do
  end subroutine Chem_MieQueryAllBand3D

  ! Chem_MieQueryAllBand4D (CB/Chem_MieMod.F90:1166-1262).
  subroutine Chem_MieQueryAllBand4D(this, idx, nbands, offset, q_mass, rh, tau, ssa, asy, rc)
    type(Chem_Mie), target, intent(in) :: this
    integer, intent(in) :: idx, nbands, offset
    real, intent(in) :: q_mass(:,:,:), rh(:,:,:)
    real, intent(out) :: tau(:,:,:,:), ssa(:,:,:,:), asy(:,:,:,:)
    integer, optional, intent(out) :: rc
    real(c_float), allocatable :: q(:), r(:), t(:), s(:), a(:)
    integer(c_int) :: k, e, n
    integer :: ib
    if (present(rc)) rc = 0
    k = lib_index(this, idx)
    if (k < 0) then
      if (present(rc)) rc = 1
      return
    end if
    n = size(rh)
    q = reshape(q_mass, (/n/))
    r = reshape(rh, (/n/))
    allocate(t(n), s(n), a(n))
    do ib = 1, nbands
      call all_band_(this, k, ib + offset, n, q, r, t, s, a, e)
      if (e /= 0) then
        if (present(rc)) rc = 1000 - e
        return
      end if
      tau(:,:,:,ib) = reshape(t, shape(rh))
      ssa(:,:,:,ib) = reshape(s, shape(rh))
      asy(:,:,:,ib) = reshape(a, shape(rh))
    end do
  end subroutine Chem_MieQueryAllBand4D

  subroutine Chem_MieDestroy(this, rc)
    type(Chem_Mie), intent(inout) :: this
    integer, intent(out) :: rc
    rc = aa_mie_destroy(this%handle)
    this%handle = c_null_ptr
    if (associated(this%channels)) deallocate(this%channels)
    if (associated(this%vname)) deallocate(this%vname)
    if (associated(this%vtableUse)) deallocate(this%vtableUse)
  end subroutine Chem_MieDestroy

end module Chem_MieMod

! VLIDORT_getAOPscalar / VLIDORT_getAOPvector
! (src/Components/rtms/geo_vlidort/VLIDORT_MieMod.F90:10-336): same dummy
! arguments; the tracer sum runs on the GPU.
subroutine VLIDORT_getAOPscalar(mieTables, km, nobs, nch, nq, channels, vname, verbose, qm, rh, tau, ssa, g, rc)
  use iso_c_binding
  use aeroapps_amd_capi
  use Chem_MieMod
  implicit none
  type(Chem_Mie), intent(inout) :: mieTables
  integer, intent(in) :: km, nobs, nch, nq, verbose
  real, intent(in) :: channels(nch)
  character, intent(in) :: vname(nq,16)
  real, intent(in) :: qm(km,nq,nobs), rh(km,nobs)
  real, intent(out) :: tau(km,nch,nobs), ssa(km,nch,nobs), g(km,nch,nobs)
  integer, intent(out) :: rc
  integer(c_int) :: e, crc
  e = aa_get_aop_scalar(mieTables%handle, km, nobs, nch, nq, channels, vname, verbose, qm, rh, tau, ssa, g, crc)
  rc = crc
  if (rc == 0 .and. e /= 0) rc = 1000 - e
end subroutine VLIDORT_getAOPscalar

subroutine VLIDORT_getAOPvector(mieTables, km, nobs, nch, nq, channels, vname, verbose, qm, rh, nMom, nPol, &
                                tau, ssa, g, pmom, rc)
  use iso_c_binding
  use aeroapps_amd_capi
  use Chem_MieMod
  implicit none
  type(Chem_Mie), intent(inout) :: mieTables
  integer, intent(in) :: km, nobs, nch, nq, verbose, nMom, nPol
  real, intent(in) :: channels(nch)
  character, intent(in) :: vname(nq,16)
  real, intent(in) :: qm(km,nq,nobs), rh(km,nobs)
  real, intent(out) :: tau(km,nch,nobs), ssa(km,nch,nobs), g(km,nch,nobs), pmom(km,nch,nobs,nMom,nPol)
  integer, intent(out) :: rc
  integer(c_int) :: e, crc
  e = aa_get_aop_vector(mieTables%handle, km, nobs, nch, nq, channels, vname, verbose, qm, rh, nMom, nPol, &
                        tau, ssa, g, pmom, crc)
  rc = crc
  if (rc == 0 .and. e /= 0) rc = 1000 - e
end subroutine VLIDORT_getAOPvector
