! ISO_C_BINDING interface to librl.so (include/radlite.h) for a Fortran
! host.  The calls replace the line loop of RADLite's main program
! (main.F:1043-1049, one calc_line_spectrum per line) and the ray set-up
! before it (main.F:788-795).  Arrays are passed in the reference's own
! storage order: a Fortran array a(it, ir) with it fastest is the C
! array [nr][nth_half] the library expects.
module radlite_gpu
    use iso_c_binding
    implicit none

    integer(c_int), parameter :: RL_FP64 = 0, RL_FP32 = 1
    integer(c_int), parameter :: RL_SKIP_REFERENCE = 0, RL_SKIP_NONE = 1
    integer(c_int), parameter :: RL_WARN_MASER = 1, RL_WARN_OVERSHOOT = 2

    type, bind(C) :: rl_item
        integer(c_int) :: line       ! 0-based line index
        integer(c_int) :: incl       ! 0-based inclination index
        integer(c_int) :: nfr        ! channels
        real(c_double) :: vmax_kms   ! passband half width
        integer(c_int) :: skip_mode
    end type rl_item

    interface
        integer(c_int) function rl_create(ctx, device, precision) bind(C, name="rl_create")
            import :: c_ptr, c_int
            type(c_ptr), intent(out) :: ctx
            integer(c_int), value :: device, precision
        end function
        subroutine rl_destroy(ctx) bind(C, name="rl_destroy")
            import :: c_ptr
            type(c_ptr), value :: ctx
        end subroutine
        type(c_ptr) function rl_last_error(ctx) bind(C, name="rl_last_error")
            import :: c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function rl_set_grid(ctx, nr, r_cm, nth_half, theta_half) &
                bind(C, name="rl_set_grid")
            import :: c_ptr, c_int, c_double
            type(c_ptr), value :: ctx
            integer(c_int), value :: nr, nth_half
            real(c_double), intent(in) :: r_cm(*), theta_half(*)
        end function
        integer(c_int) function rl_set_medium(ctx, rho, abund, vel_cms, lw_kms, umass_av, &
                aksmax_kms) bind(C, name="rl_set_medium")
            import :: c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), intent(in) :: rho(*), abund(*), vel_cms(*), lw_kms(*)
            real(c_double), value :: umass_av, aksmax_kms
        end function
        integer(c_int) function rl_set_levels(ctx, nlev, pops) bind(C, name="rl_set_levels")
            import :: c_ptr, c_int, c_double
            type(c_ptr), value :: ctx
            integer(c_int), value :: nlev
            real(c_double), intent(in) :: pops(*)
        end function
        integer(c_int) function rl_set_lines(ctx, nlines, nu0, a, gu, gl, lev_up, lev_low) &
                bind(C, name="rl_set_lines")
            import :: c_ptr, c_int, c_double
            type(c_ptr), value :: ctx
            integer(c_int), value :: nlines
            real(c_double), intent(in) :: nu0(*), a(*), gu(*), gl(*)
            integer(c_int), intent(in) :: lev_up(*), lev_low(*)
        end function
        integer(c_int) function rl_set_dust_line(ctx, jd, ad) bind(C, name="rl_set_dust_line")
            import :: c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), intent(in) :: jd(*), ad(*)
        end function
        integer(c_int) function rl_set_star(ctx, rstar_cm, ncont, cont_nu, istar) &
                bind(C, name="rl_set_star")
            import :: c_ptr, c_int, c_double
            type(c_ptr), value :: ctx
            real(c_double), value :: rstar_cm
            integer(c_int), value :: ncont
            real(c_double), intent(in) :: cont_nu(*), istar(*)
        end function
        integer(c_int) function rl_set_outer_bc(ctx, itype) bind(C, name="rl_set_outer_bc")
            import :: c_ptr, c_int
            type(c_ptr), value :: ctx
            integer(c_int), value :: itype
        end function
        integer(c_int) function rl_set_interstellar_field(ctx, ncont, cont_nu, isrf) &
                bind(C, name="rl_set_interstellar_field")
            import :: c_ptr, c_int, c_double
            type(c_ptr), value :: ctx
            integer(c_int), value :: ncont
            real(c_double), intent(in) :: cont_nu(*), isrf(*)
        end function
        integer(c_int) function rl_setup_rays_circular(ctx, nincl, incl_rad, nphi, nrext, dbdr, &
                n_rays, n_rings, ring_r, ring_edge) bind(C, name="rl_setup_rays_circular")
            import :: c_ptr, c_int, c_double
            type(c_ptr), value :: ctx
            integer(c_int), value :: nincl, nphi, nrext, dbdr
            real(c_double), intent(in) :: incl_rad(*)
            integer(c_int), intent(out) :: n_rays(*), n_rings(*)
            type(c_ptr), value :: ring_r, ring_edge
        end function
        integer(c_int) function rl_trace_lines_forward(ctx, nitems, items, flux, image, warn_bits) &
                bind(C, name="rl_trace_lines_forward")
            import :: c_ptr, c_int, c_double, rl_item
            type(c_ptr), value :: ctx
            integer(c_int), value :: nitems
            type(rl_item), intent(in) :: items(*)
            real(c_double), intent(out) :: flux(*)
            type(c_ptr), value :: image
            integer(c_int), intent(out) :: warn_bits(*)
        end function
        integer(c_int) function rl_setup_rays_rectang(ctx, nincl, incl_rad, nx, ny, size_x, &
                size_y, phioff, xoff, yoff, addstar, n_rays) bind(C, name="rl_setup_rays_rectang")
            import :: c_ptr, c_int, c_double
            type(c_ptr), value :: ctx
            integer(c_int), value :: nincl, nx, ny, addstar
            real(c_double), intent(in) :: incl_rad(*)
            real(c_double), value :: size_x, size_y, phioff, xoff, yoff
            integer(c_int), intent(out) :: n_rays(*)
        end function
        integer(c_int) function rl_trace_lines_posvel(ctx, nitems, items, image, tau, warn_bits) &
                bind(C, name="rl_trace_lines_posvel")
            import :: c_ptr, c_int, c_double, rl_item
            type(c_ptr), value :: ctx
            integer(c_int), value :: nitems
            type(rl_item), intent(in) :: items(*)
            real(c_double), intent(out) :: image(*), tau(*)
            integer(c_int), intent(out) :: warn_bits(*)
        end function
        integer(c_int) function rl_resample_image(ctx, nrays, nfr, image, npix, nv, vtx, w, out) &
                bind(C, name="rl_resample_image")
            import :: c_ptr, c_int, c_double
            type(c_ptr), value :: ctx
            integer(c_int), value :: nrays, nfr, npix, nv
            real(c_double), intent(in) :: image(*), w(*)
            integer(c_int), intent(in) :: vtx(*)
            real(c_double), intent(out) :: out(*)
        end function
        integer(c_int) function rl_sync(ctx) bind(C, name="rl_sync")
            import :: c_ptr, c_int
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function rl_set_memory_budget(ctx, bytes) &
                bind(C, name="rl_set_memory_budget")
            import :: c_ptr, c_int, c_double
            type(c_ptr), value :: ctx
            real(c_double), value :: bytes
        end function
        integer(c_int) function rl_set_option(ctx, option, val) bind(C, name="rl_set_option")
            import :: c_ptr, c_int
            type(c_ptr), value :: ctx
            integer(c_int), value :: option, val
        end function
        integer(c_int) function rl_plan(ctx, nitems, items, out) bind(C, name="rl_plan")
            import :: c_ptr, c_int, c_double, rl_item
            type(c_ptr), value :: ctx
            integer(c_int), value :: nitems
            type(rl_item), intent(in) :: items(*)
            real(c_double), intent(out) :: out(8)
        end function
        integer(c_int) function rl_memory_info(ctx, out) bind(C, name="rl_memory_info")
            import :: c_ptr, c_int, c_double
            type(c_ptr), value :: ctx
            real(c_double), intent(out) :: out(5)
        end function
        ! host pre-pass with the reference's arithmetic (include/radlite.h)
        integer(c_int) function rl_prepare_line_freqs(nl, e_cm, lev_up, lev_low, nu0) &
                bind(C, name="rl_prepare_line_freqs")
            import :: c_int, c_double
            integer(c_int), value :: nl
            real(c_double), intent(in) :: e_cm(*)
            integer(c_int), intent(in) :: lev_up(*), lev_low(*)
            real(c_double), intent(out) :: nu0(*)
        end function
        integer(c_int) function rl_prepare_line_dust(nr, nth_half, nspec, rho_d, t_d, nfreq, &
                cont_nu, kabs, kscat, nlines, nu0, jd, ad) bind(C, name="rl_prepare_line_dust")
            import :: c_int, c_double
            integer(c_int), value :: nr, nth_half, nspec, nfreq, nlines
            real(c_double), intent(in) :: rho_d(*), t_d(*), cont_nu(*), kabs(*), kscat(*), nu0(*)
            real(c_double), intent(out) :: jd(*), ad(*)
        end function
        integer(c_int) function rl_prepare_line_dust_scat(nr, nth_half, nspec, rho_d, t_d, &
                nfreq, cont_nu, kabs, kscat, scat, nlines, nu0, jd, ad) &
                bind(C, name="rl_prepare_line_dust_scat")
            import :: c_int, c_double, c_ptr
            integer(c_int), value :: nr, nth_half, nspec, nfreq, nlines
            real(c_double), intent(in) :: rho_d(*), t_d(*), cont_nu(*), kabs(*), kscat(*), nu0(*)
            type(c_ptr), value :: scat
            real(c_double), intent(out) :: jd(*), ad(*)
        end function
        integer(c_int) function rl_prepare_star(nfreq, fnu_1pc, rstar, istar) &
                bind(C, name="rl_prepare_star")
            import :: c_int, c_double
            integer(c_int), value :: nfreq
            real(c_double), intent(in) :: fnu_1pc(*)
            real(c_double), value :: rstar
            real(c_double), intent(out) :: istar(*)
        end function
        real(c_double) function rl_prepare_aksmax(n, lw_kms, fix_lw) &
                bind(C, name="rl_prepare_aksmax")
            import :: c_int, c_double
            integer(c_int), value :: n
            real(c_double), intent(in) :: lw_kms(*)
            real(c_double), value :: fix_lw
        end function
        integer(c_int) function rl_passband(vmax, dv, nfr, vmax_out) bind(C, name="rl_passband")
            import :: c_int, c_double
            real(c_double), value :: vmax, dv
            integer(c_int), intent(out) :: nfr
            real(c_double), intent(out) :: vmax_out
        end function
    end interface

contains

    ! rl_last_error as a Fortran string
    function rl_error_text(ctx) result(s)
        type(c_ptr), intent(in) :: ctx
        character(len=256) :: s
        character(kind=c_char), pointer :: p(:)
        integer :: i
        s = ' '
        call c_f_pointer(rl_last_error(ctx), p, [256])
        do i = 1, 256
            if (p(i) == c_null_char) exit
            s(i:i) = p(i)
        end do
    end function

end module radlite_gpu
