! Subset of module Chem_RegistryMod
! (src/Shared/GMAO_Shared/Chem_Base/Chem_RegistryMod.F90) that Chem_MieCreate
! needs: type(Chem_Registry) with the tracer count and names and, for the
! seven aerosol species with optics tables, doing_XX and the tracer range
! i_XX..j_XX; Chem_RegistryCreate parses the i90 registry file through the
! C-ABI (aa_mie_registry), Chem_RegistryDestroy frees it.
module Chem_RegistryMod
  use iso_c_binding
  use aeroapps_amd_capi
  implicit none
  private
  public Chem_Registry, Chem_RegistryCreate, Chem_RegistryDestroy

  integer, parameter, public :: CHEM_REG_NCH = 16   ! characters of a tracer name kept

  type Chem_Registry
    integer :: nq = 0
    character(len=CHEM_REG_NCH), pointer :: vname(:) => null()
    integer, pointer :: species(:) => null()      ! AA_SP_* code of each tracer, -1 without optics table
    logical :: doing_DU = .false., doing_SS = .false., doing_SU = .false., doing_OC = .false.
    logical :: doing_BC = .false., doing_NI = .false., doing_BRC = .false.
    integer :: i_DU = -1, j_DU = -2, i_SS = -1, j_SS = -2, i_SU = -1, j_SU = -2, i_OC = -1, j_OC = -2
    integer :: i_BC = -1, j_BC = -2, i_NI = -1, j_NI = -2, i_BRC = -1, j_BRC = -2
  end type Chem_Registry

contains

  ! Chem_RegistryCreate(rc, rcfile): rcfile defaults to 'Chem_Registry.rc'.
  ! rc: 0, 1 unreadable file, 2 a species on without nbins or units, 3 too few names
  ! in a variable table (the reference stops in these cases).
  type(Chem_Registry) function Chem_RegistryCreate(rc, rcfile) result(this)
    integer, intent(out) :: rc
    character(len=*), optional, intent(in) :: rcfile
    character(len=:), allocatable :: path
    character(kind=c_char), allocatable :: cpath(:), names(:)
    integer(c_int), allocatable :: sp(:)
    integer(c_int) :: nq, crc, e
    integer :: i, c
    if (present(rcfile)) then
      path = trim(rcfile)
    else
      path = 'Chem_Registry.rc'
    end if
    allocate(cpath(len(path) + 1))
    do i = 1, len(path)
      cpath(i) = path(i:i)
    end do
    cpath(len(path) + 1) = c_null_char
    allocate(names(16), sp(1))
    e = aa_mie_registry(cpath, 0, nq, names, sp, crc)
    rc = crc
    if (e /= 0) rc = 1000 - e
    if (rc /= 0) return
    deallocate(names, sp)
    allocate(names(16 * max(nq, 1)), sp(max(nq, 1)))
    e = aa_mie_registry(cpath, nq, nq, names, sp, crc)
    rc = crc
    if (e /= 0) rc = 1000 - e
    if (rc /= 0) return
    this%nq = nq
    allocate(this%vname(nq), this%species(nq))
    do i = 1, nq
      this%vname(i) = ' '
      do c = 1, 16
        this%vname(i)(c:c) = names((i - 1) * 16 + c)
      end do
      this%species(i) = sp(i)
    end do
    call range_(0, this%doing_DU, this%i_DU, this%j_DU)
    call range_(1, this%doing_SS, this%i_SS, this%j_SS)
    call range_(2, this%doing_SU, this%i_SU, this%j_SU)
    call range_(3, this%doing_OC, this%i_OC, this%j_OC)
    call range_(4, this%doing_BC, this%i_BC, this%j_BC)
    call range_(5, this%doing_NI, this%i_NI, this%j_NI)
    call range_(6, this%doing_BRC, this%i_BRC, this%j_BRC)
  contains
    subroutine range_(code, doing, i0, i1)
      integer, intent(in) :: code
      logical, intent(out) :: doing
      integer, intent(out) :: i0, i1
      integer :: k
      doing = .false.
      i0 = -1
      i1 = -2
      do k = 1, this%nq
        if (this%species(k) == code) then
          if (.not. doing) i0 = k
          doing = .true.
          i1 = k
        end if
      end do
    end subroutine range_
  end function Chem_RegistryCreate

  subroutine Chem_RegistryDestroy(this, rc)
    type(Chem_Registry), intent(inout) :: this
    integer, intent(out) :: rc
    if (associated(this%vname)) deallocate(this%vname)
    if (associated(this%species)) deallocate(this%species)
    this%nq = 0
    rc = 0
  end subroutine Chem_RegistryDestroy

end module Chem_RegistryMod
