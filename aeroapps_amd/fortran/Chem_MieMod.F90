! Device-backed subset of module Chem_MieMod
! (src/Shared/GMAO_Shared/Chem_Base/Chem_MieMod.F90:16-109) for the batch
! optics path: type(Chem_Mie) with its public components (nch, nMom, nPol,
! channels) holding a handle to the GPU-resident tables, Chem_MieQueryIdx and
! Chem_MieDestroy.  The netCDF/RC-file reader of Chem_MieCreate is not on the
! accelerated path; tables interpolated to the channels are added with
! Chem_MieAddTable (the result of Chem_MieTableRead's polint step,
! src/Shared/GMAO_Shared/Chem_Base/Chem_MieTableMod.F90:560-589).
module Chem_MieMod
  use iso_c_binding
  use aeroapps_amd_capi
  implicit none
  private
  public Chem_Mie, Chem_MieCreate, Chem_MieAddTable, Chem_MieQueryIdx, Chem_MieDestroy

  type Chem_Mie
    integer :: nch = 0
    integer :: nMom = 0
    integer :: nPol = 0
    real, pointer :: channels(:) => null()
    type(c_ptr) :: handle = c_null_ptr
  end type Chem_Mie

contains

  ! Empty table set on nch channels given in metres.
  type(Chem_Mie) function Chem_MieCreate(nch, channels_m, rc) result(this)
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
  end function Chem_MieCreate

  ! One species table: bext..refi(nrh,nch,nbin), pmom(nrh,nch,nbin,nmom,npol),
  ! bin_names(nbin) tracer names served by the table.
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

  ! Chem_MieQueryIdx (src/Shared/GMAO_Shared/Chem_Base/Chem_MieMod.F90:720-777):
  ! 1-based index, -1 when the name is not served.
  integer function Chem_MieQueryIdx(this, qname, rc) result(idx)
    type(Chem_Mie), intent(inout) :: this
    character(len=*), intent(in) :: qname
    integer, intent(out) :: rc
    character(kind=c_char) :: buf(len_trim(qname)+1)
    integer :: i, k
    do i = 1, len_trim(qname)
      buf(i) = qname(i:i)
    end do
    buf(len_trim(qname)+1) = c_null_char
    k = aa_mie_query_idx(this%handle, buf)
    rc = 0
    if (k < 0) then
      idx = -1
      rc = 1
    else
      idx = k + 1
    end if
  end function Chem_MieQueryIdx

  subroutine Chem_MieDestroy(this, rc)
    type(Chem_Mie), intent(inout) :: this
    integer, intent(out) :: rc
    rc = aa_mie_destroy(this%handle)
    this%handle = c_null_ptr
    if (associated(this%channels)) deallocate(this%channels)
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
