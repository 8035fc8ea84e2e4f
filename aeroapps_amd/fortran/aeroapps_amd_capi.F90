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

    integer(c_int) function aa_mie_query(t, idx, channel, n, q_mass, rh, tau, ssa, gasym, bext, bsca, bbck, &
                                         refr, refi, nmom, npol, pmom) bind(C, name="aa_mie_query")
      import :: c_int, c_float, c_ptr
      type(c_ptr), value :: t
      integer(c_int), value :: idx, n, nmom, npol
      real(c_float), value :: channel
      real(c_float), intent(in) :: q_mass(*), rh(*)
      type(c_ptr), value :: tau, ssa, gasym, bext, bsca, bbck, refr, refi, pmom
    end function

    integer(c_int) function aa_mie_query_extra(t, idx, channel, n, rh, reff, p11, p22, gf, rhod, vol, area) &
        bind(C, name="aa_mie_query_extra")
      import :: c_int, c_float, c_ptr
      type(c_ptr), value :: t
      integer(c_int), value :: idx, n
      real(c_float), value :: channel
      real(c_float), intent(in) :: rh(*)
      type(c_ptr), value :: reff, p11, p22, gf, rhod, vol, area
    end function

    integer(c_int) function aa_mie_dims(t, nch, nmom_min, npol_max, ntracer) bind(C, name="aa_mie_dims")
      import :: c_int, c_ptr
      type(c_ptr), value :: t
      integer(c_int), intent(out) :: nch, nmom_min, npol_max, ntracer
    end function

    integer(c_int) function aa_mie_channels(t, channels_m) bind(C, name="aa_mie_channels")
      import :: c_int, c_float, c_ptr
      type(c_ptr), value :: t
      real(c_float), intent(out) :: channels_m(*)
    end function

    integer(c_int) function aa_mie_tracer_table(t, idx, table, bin, nbin, nrh, nmom, npol) &
        bind(C, name="aa_mie_tracer_table")
      import :: c_int, c_ptr
      type(c_ptr), value :: t
      integer(c_int), value :: idx
      integer(c_int), intent(out) :: table, bin, nbin, nrh, nmom, npol
    end function

    integer(c_int) function aa_mie_registry(path, nq_max, nq, names, species, rc) bind(C, name="aa_mie_registry")
      import :: c_int, c_char
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int), value :: nq_max
      integer(c_int), intent(out) :: nq
      character(kind=c_char), intent(out) :: names(*)
      integer(c_int), intent(out) :: species(*)
      integer(c_int), intent(out) :: rc
    end function

    integer(c_int) function aa_mie_create_rc(t, rcfile, nq, names, species, reader, user, rc) &
        bind(C, name="aa_mie_create_rc")
      import :: c_int, c_char, c_ptr, c_funptr
      type(c_ptr), intent(out) :: t
      character(kind=c_char), intent(in) :: rcfile(*)
      integer(c_int), value :: nq
      character(kind=c_char), intent(in) :: names(*)
      integer(c_int), intent(in) :: species(*)
      type(c_funptr), value :: reader
      type(c_ptr), value :: user
      integer(c_int), intent(out) :: rc
    end function

    integer(c_int) function aa_vlidort_vector_lambert_opt(km, nch, nobs, channels, nstreams, plane_parallel, &
        nmom, npol, rot, depol, alpha, tau, ssa, pmom, pe, he, te, albedo, sza, raa, vza, flux_factor, missing, &
        verbose, precision, do_2os, do_boa, radiance, reflectance, q, u, nfourier, rc) &
        bind(C, name="aa_vlidort_vector_lambert_opt")
      import :: c_int, c_double, c_ptr
      integer(c_int), value :: km, nch, nobs, nstreams, plane_parallel, nmom, npol
      real(c_double), intent(in) :: channels(*), rot(*), depol(*), alpha(*), tau(*), ssa(*), pmom(*)
      real(c_double), intent(in) :: pe(*), he(*), te(*), albedo(*), sza(*), raa(*), vza(*), flux_factor(*)
      real(c_double), value :: missing
      integer(c_int), value :: verbose, precision, do_2os, do_boa
      real(c_double), intent(out) :: radiance(*), reflectance(*), q(*), u(*)
      type(c_ptr), value :: nfourier
      integer(c_int), intent(out) :: rc
    end function

    integer(c_int) function aa_vlidort_vector_lambert_multigeom(km, nch, nobs, ngeom, channels, nstreams, &
        plane_parallel, nmom, npol, rot, depol, alpha, tau, ssa, pmom, pe, he, te, albedo, sza, raa, vza, &
        flux_factor, missing, verbose, precision, do_2os, do_boa, radiance, reflectance, q, u, nfourier, rc) &
        bind(C, name="aa_vlidort_vector_lambert_multigeom")
      import :: c_int, c_double, c_ptr
      integer(c_int), value :: km, nch, nobs, ngeom, nstreams, plane_parallel, nmom, npol
      real(c_double), intent(in) :: channels(*), rot(*), depol(*), alpha(*), tau(*), ssa(*), pmom(*)
      real(c_double), intent(in) :: pe(*), he(*), te(*), albedo(*), sza(*), raa(*), vza(*), flux_factor(*)
      real(c_double), value :: missing
      integer(c_int), value :: verbose, precision, do_2os, do_boa
      real(c_double), intent(out) :: radiance(*), reflectance(*), q(*), u(*)
      type(c_ptr), value :: nfourier
      integer(c_int), intent(out) :: rc
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

    integer(c_int) function aa_vlidort_vector_cx(km, nch, nobs, channels, nstreams, plane_parallel, nmom, &
        npol, rot, depol, tau, ssa, pmom, pe, he, te, u10m, v10m, mr, sza, raa, vza, missing, verbose, &
        precision, radiance, reflectance, q, u, brdf, brdf_q, brdf_u, nfourier, rc) &
        bind(C, name="aa_vlidort_vector_cx")
      import :: c_int, c_double, c_ptr
      integer(c_int), value :: km, nch, nobs, nstreams, plane_parallel, nmom, npol
      real(c_double), intent(in) :: channels(*), rot(*), depol(*), tau(*), ssa(*), pmom(*)
      real(c_double), intent(in) :: pe(*), he(*), te(*), u10m(*), v10m(*), mr(*), sza(*), raa(*), vza(*)
      real(c_double), value :: missing
      integer(c_int), value :: verbose, precision
      real(c_double), intent(out) :: radiance(*), reflectance(*), q(*), u(*), brdf(*), brdf_q(*), brdf_u(*)
      type(c_ptr), value :: nfourier
      integer(c_int), intent(out) :: rc
    end function

    integer(c_int) function aa_vlidort_vector_lambert_bpdf(km, nch, nobs, channels, nstreams, plane_parallel, &
        nmom, npol, rot, depol, tau, ssa, pmom, pe, he, te, albedo, bpdf, sza, raa, vza, missing, verbose, &
        precision, radiance, reflectance, br, q, u, br_q, br_u, nfourier, rc) &
        bind(C, name="aa_vlidort_vector_lambert_bpdf")
      import :: c_int, c_double, c_ptr
      integer(c_int), value :: km, nch, nobs, nstreams, plane_parallel, nmom, npol
      real(c_double), intent(in) :: channels(*), rot(*), depol(*), tau(*), ssa(*), pmom(*)
      real(c_double), intent(in) :: pe(*), he(*), te(*), albedo(*), bpdf(*), sza(*), raa(*), vza(*)
      real(c_double), value :: missing
      integer(c_int), value :: verbose, precision
      real(c_double), intent(out) :: radiance(*), reflectance(*), br(*), q(*), u(*), br_q(*), br_u(*)
      type(c_ptr), value :: nfourier
      integer(c_int), intent(out) :: rc
    end function

    integer(c_int) function aa_vlidort_vector_landmodis_bpdf(km, nch, nobs, channels, nstreams, plane_parallel, &
        nmom, npol, rot, depol, tau, ssa, pmom, pe, he, te, kernel_wt, param, bpdf, sza, raa, vza, missing, &
        verbose, precision, radiance, reflectance, br, q, u, br_q, br_u, nfourier, rc) &
        bind(C, name="aa_vlidort_vector_landmodis_bpdf")
      import :: c_int, c_double, c_ptr
      integer(c_int), value :: km, nch, nobs, nstreams, plane_parallel, nmom, npol
      real(c_double), intent(in) :: channels(*), rot(*), depol(*), tau(*), ssa(*), pmom(*)
      real(c_double), intent(in) :: pe(*), he(*), te(*), kernel_wt(*), param(*), bpdf(*), sza(*), raa(*), vza(*)
      real(c_double), value :: missing
      integer(c_int), value :: verbose, precision
      real(c_double), intent(out) :: radiance(*), reflectance(*), br(*), q(*), u(*), br_q(*), br_u(*)
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
