! ISO_C_BINDING interfaces of the C-ABI declared in include/aeroapps_amd.h.
! Arrays are passed by address in their Fortran (column-major) layout.
module aeroapps_amd_capi
  use iso_c_binding
  implicit none

  interface
    integer(c_int) function aa_mie_create(t, nch, channels_m) bind(C, name="aa_mie_create")
      import :: c_int, c_float, c_ptr
      type(c_ptr), intent(out) :: t
      integer(c_int), value :: nch
      real(c_float), intent(in) :: channels_m(*)
    end function

    integer(c_int) function aa_mie_add_table(t, nrh, nbin, nmom, npol, rh, bext, bsca, bbck, g, refr, refi, &
                                             pmom, bin_names) bind(C, name="aa_mie_add_table")
      import :: c_int, c_float, c_ptr, c_char
      type(c_ptr), value :: t
      integer(c_int), value :: nrh, nbin, nmom, npol
      real(c_float), intent(in) :: rh(*), bext(*), bsca(*), bbck(*), g(*), refr(*), refi(*), pmom(*)
      character(kind=c_char), intent(in) :: bin_names(*)
    end function

    integer(c_int) function aa_mie_query_idx(t, qname) bind(C, name="aa_mie_query_idx")
      import :: c_int, c_ptr, c_char
      type(c_ptr), value :: t
      character(kind=c_char), intent(in) :: qname(*)
    end function

    integer(c_int) function aa_mie_destroy(t) bind(C, name="aa_mie_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: t
    end function

    integer(c_int) function aa_get_aop_vector(t, km, nobs, nch, nq, channels, vname, verbose, qm, rh, nmom, npol, &
                                              tau, ssa, g, pmom, rc) bind(C, name="aa_get_aop_vector")
      import :: c_int, c_float, c_ptr, c_char
      type(c_ptr), value :: t
      integer(c_int), value :: km, nobs, nch, nq, verbose, nmom, npol
      real(c_float), intent(in) :: channels(*), qm(*), rh(*)
      character(kind=c_char), intent(in) :: vname(*)
      real(c_float), intent(out) :: tau(*), ssa(*), g(*), pmom(*)
      integer(c_int), intent(out) :: rc
    end function

    integer(c_int) function aa_get_aop_scalar(t, km, nobs, nch, nq, channels, vname, verbose, qm, rh, &
                                              tau, ssa, g, rc) bind(C, name="aa_get_aop_scalar")
      import :: c_int, c_float, c_ptr, c_char
      type(c_ptr), value :: t
      integer(c_int), value :: km, nobs, nch, nq, verbose
      real(c_float), intent(in) :: channels(*), qm(*), rh(*)
      character(kind=c_char), intent(in) :: vname(*)
      real(c_float), intent(out) :: tau(*), ssa(*), g(*)
      integer(c_int), intent(out) :: rc
    end function

    integer(c_int) function aa_vlidort_scalar_lambert(km, nch, nobs, channels, nstreams, plane_parallel, nmom, &
        npol, rot, depol, tau, ssa, g, pmom, pe, he, te, albedo, sza, raa, vza, missing, verbose, precision, &
        radiance, reflectance, nfourier, rc) bind(C, name="aa_vlidort_scalar_lambert")
      import :: c_int, c_double, c_ptr
      integer(c_int), value :: km, nch, nobs, nstreams, plane_parallel, nmom, npol
      real(c_double), intent(in) :: channels(*), rot(*), depol(*), tau(*), ssa(*), g(*), pmom(*)
      real(c_double), intent(in) :: pe(*), he(*), te(*), albedo(*), sza(*), raa(*), vza(*)
      real(c_double), value :: missing
      integer(c_int), value :: verbose, precision
      real(c_double), intent(out) :: radiance(*), reflectance(*)
      type(c_ptr), value :: nfourier
      integer(c_int), intent(out) :: rc
    end function

    integer(c_int) function aa_vlidort_scalar_landmodis(km, nch, nobs, channels, nstreams, plane_parallel, nmom, &
        npol, rot, depol, tau, ssa, g, pmom, pe, he, te, kernel_wt, param, sza, raa, vza, missing, verbose, &
        precision, radiance, reflectance, br, nfourier, rc) bind(C, name="aa_vlidort_scalar_landmodis")
      import :: c_int, c_double, c_ptr
      integer(c_int), value :: km, nch, nobs, nstreams, plane_parallel, nmom, npol
      real(c_double), intent(in) :: channels(*), rot(*), depol(*), tau(*), ssa(*), g(*), pmom(*)
      real(c_double), intent(in) :: pe(*), he(*), te(*), kernel_wt(*), param(*), sza(*), raa(*), vza(*)
      real(c_double), value :: missing
      integer(c_int), value :: verbose, precision
      real(c_double), intent(out) :: radiance(*), reflectance(*), br(*)
      type(c_ptr), value :: nfourier
      integer(c_int), intent(out) :: rc
    end function

    integer(c_int) function aa_vlidort_scalar_cx(km, nch, nobs, channels, nstreams, plane_parallel, nmom, &
        npol, rot, depol, tau, ssa, g, pmom, pe, he, te, u10m, v10m, mr, sza, raa, vza, missing, verbose, &
        precision, radiance, reflectance, brdf, nfourier, rc) bind(C, name="aa_vlidort_scalar_cx")
      import :: c_int, c_double, c_ptr
      integer(c_int), value :: km, nch, nobs, nstreams, plane_parallel, nmom, npol
      real(c_double), intent(in) :: channels(*), rot(*), depol(*), tau(*), ssa(*), g(*), pmom(*)
      real(c_double), intent(in) :: pe(*), he(*), te(*), u10m(*), v10m(*), mr(*), sza(*), raa(*), vza(*)
      real(c_double), value :: missing
      integer(c_int), value :: verbose, precision
      real(c_double), intent(out) :: radiance(*), reflectance(*), brdf(*)
      type(c_ptr), value :: nfourier
      integer(c_int), intent(out) :: rc
    end function

    integer(c_int) function aa_vlidort_vector_lambert(km, nch, nobs, channels, nstreams, plane_parallel, nmom, &
        npol, rot, depol, alpha, tau, ssa, pmom, pe, he, te, albedo, sza, raa, vza, flux_factor, missing, verbose, &
        precision, radiance, reflectance, q, u, nfourier, rc) bind(C, name="aa_vlidort_vector_lambert")
      import :: c_int, c_double, c_ptr
      integer(c_int), value :: km, nch, nobs, nstreams, plane_parallel, nmom, npol
      real(c_double), intent(in) :: channels(*), rot(*), depol(*), alpha(*), tau(*), ssa(*), pmom(*)
      real(c_double), intent(in) :: pe(*), he(*), te(*), albedo(*), sza(*), raa(*), vza(*), flux_factor(*)
      real(c_double), value :: missing
      integer(c_int), value :: verbose, precision
      real(c_double), intent(out) :: radiance(*), reflectance(*), q(*), u(*)
      type(c_ptr), value :: nfourier
      integer(c_int), intent(out) :: rc
    end function
  end interface
end module aeroapps_amd_capi
