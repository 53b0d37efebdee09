! Fortran host for the MI355X raytracer: the RADlite executable's line
! spectrum job for one run directory, with the raytrace done by librl.so
! through ISO_C_BINDING (module radlite_gpu).
!
!   rl_host [run directory] [fp64|fp32]
!
! Reads the run directory's L4 files as RADLite's readers do (radlite.inp
! main.F:535-597 and tools.F:6-19; linespectrum.inp telescope.F:104-238;
! line.inp line.F:168-279; moldata line.F:1865-1935; levelpop
! line.F:842-919; fields setup.F:754-911,1453-1485; dust
! dust.F:92-520 and setup.F:159-250,701-725; star star.F:449-620),
! prepares the line tables with the library's host pre-pass (the
! reference's arithmetic), traces every requested line at the
! linespectrum.inp inclination (calc_line_spectrum, telescope.F:1465) and
! writes linespectrum_<molecule>.dat (write_line_spectrum,
! telescope.F:1662-1804); when linespectrum.inp asks for position-velocity
! cubes (command 2, telescope.F:202-230) it renders them instead on the
! rectangular image and writes lineposvel_<molecule>_<line>.dat
! (calc_write_line_posvel, telescope.F:1828-2031).  A library error stops
! with the reference's stop code.
program rl_host
    use iso_c_binding
    use radlite_gpu
    implicit none
    character(len=512) :: dir, arg, molfile, molname, line
    integer :: nr, nth, nf, nspec, nsp_t, nlev, nlin, i, k, ir, it, u, idum, isp, nsize
    integer :: nphi, dbdr, b_extra, outer_bc, nlines_run, istart, iprec, ios, nrows
    integer :: lcmd, imx, imy, isz, iaddstar
    real(c_double) :: szx, szy, phioff, xoff, yoff
    real(c_double), allocatable :: image(:), tau(:)
    real(c_double) :: umass, fix_lw, vmax, dv, incl_deg, vrad, dist, rstar, aks, vmax2, dummy
    real(c_double), allocatable :: r(:), th(:), rho(:,:), abund(:,:), vel(:,:,:), lw(:,:)
    real(c_double), allocatable :: rhod(:,:,:), td(:,:,:), cnu(:), kabs(:,:), kscat(:,:)
    real(c_double), allocatable :: fnu(:), istar(:), pops(:,:,:), E(:), g(:), A(:), nu0(:)
    real(c_double), allocatable :: gu(:), gl(:), jd(:,:,:), ad(:,:,:), flux(:), work(:)
    real(c_double), allocatable, target :: scat(:,:,:)
    real(c_double), allocatable :: isrf(:)
    logical :: have_scat, fex
    integer(c_int), allocatable :: lup(:), llo(:), warn(:), nrays(:), nrings(:)
    type(rl_item), allocatable :: items(:)
    character(len=16), allocatable :: rows(:)
    type(c_ptr) :: ctx
    integer(c_int) :: ierr, nfr
    real(c_double) :: inc(1)

    dir = '.'
    iprec = RL_FP64
    if (command_argument_count() >= 1) call get_command_argument(1, dir)
    if (command_argument_count() >= 2) then
        call get_command_argument(2, arg)
        if (trim(arg) == 'fp32') iprec = RL_FP32
    end if

    ! radlite.inp: first 16 columns of the non-comment lines
    allocate(rows(200))
    nrows = 0
    open(newunit=u, file=path('radlite.inp'), status='old', action='read')
    do
        read(u, '(a)', iostat=ios) line
        if (ios /= 0) exit
        if (len_trim(line) == 0) cycle
        if (index('#;%=', line(1:1)) > 0) cycle
        nrows = nrows + 1
        rows(nrows) = line(1:16)
    end do
    close(u)
    read(rows(14), *) outer_bc
    read(rows(18), *) nphi
    read(rows(19), *) dbdr
    read(rows(20), *) b_extra

    ! grid and fields (it fastest, then ir)
    open(newunit=u, file=path('radius.inp'), status='old', action='read')
    read(u, *) nr
    allocate(r(nr))
    read(u, *) r
    close(u)
    open(newunit=u, file=path('theta.inp'), status='old', action='read')
    read(u, *) nth, idum
    allocate(th(nth))
    read(u, *) th
    close(u)
    allocate(rho(nth, nr), abund(nth, nr), vel(3, nth, nr), lw(nth, nr))
    open(newunit=u, file=path('density.inp'), status='old', action='read')
    read(u, *) idum, idum, idum
    read(u, *) rho
    close(u)
    open(newunit=u, file=path('abundance.inp'), status='old', action='read')
    read(u, *) idum, idum
    do ir = 1, nr
        do it = 1, nth
            read(u, *) abund(it, ir), dummy
        end do
    end do
    close(u)
    open(newunit=u, file=path('velocity.inp'), status='old', action='read')
    read(u, *) idum, idum
    read(u, *) vel
    close(u)
    open(newunit=u, file=path('turbulence.inp'), status='old', action='read')
    read(u, *) idum
    read(u, *) idum, idum
    read(u, *) lw
    close(u)

    ! dust density, temperature, frequencies, opacities
    open(newunit=u, file=path('dustdens.inp'), status='old', action='read')
    read(u, *) nspec, idum, idum, idum
    allocate(rhod(nth, nr, nspec), td(nth, nr, nspec))
    read(u, *) rhod
    close(u)
    open(newunit=u, file=path('dusttemp_final.dat'), status='old', action='read')
    read(u, *) nsp_t, idum, idum, idum
    do isp = 1, nsp_t
        read(u, *) nsize
        read(u, *) td(:, :, isp)
    end do
    close(u)
    open(newunit=u, file=path('frequency.inp'), status='old', action='read')
    read(u, *) nf
    allocate(cnu(nf), kabs(nf, nspec), kscat(nf, nspec), fnu(nf), istar(nf), work(2 * nf))
    read(u, *) cnu
    close(u)
    open(newunit=u, file=path('dustopac.inp'), status='old', action='read')
    read(u, *) idum
    read(u, *) nspec
    read(u, '(a)') line
    do isp = 1, nspec
        read(u, *) idum
        read(u, *) idum
        read(u, *) k
        read(u, '(a)') line
        call read_opac(k, isp)
    end do
    close(u)
    open(newunit=u, file=path('starinfo.inp'), status='old', action='read')
    read(u, *) idum
    read(u, *) rstar
    close(u)
    open(newunit=u, file=path('starspectrum.inp'), status='old', action='read')
    read(u, *) idum
    read(u, *) work
    close(u)
    fnu = work(2:2 * nf:2)

    ! line.inp: mean molecular weight and the fixed line width fall-back
    open(newunit=u, file=path('line.inp'), status='old', action='read')
    read(u, *) idum
    read(u, *) idum
    read(u, *) umass
    read(u, '(a)') line
    read(u, '(a)') line
    read(u, *) idum
    read(u, *) idum
    read(u, *) idum
    read(u, *) dummy
    read(u, *) idum
    fix_lw = 0.0d0
    do
        read(u, *) k
        if (k == 0) exit
        read(u, *) fix_lw
    end do
    close(u)

    ! linespectrum.inp
    open(newunit=u, file=path('linespectrum.inp'), status='old', action='read')
    read(u, *) idum
    read(u, *) idum
    read(u, '(a)') line
    read(u, *) idum
    read(u, '(a)') line
    read(u, *) vmax
    read(u, *) dv
    read(u, '(a)') line
    read(u, *) idum
    read(u, '(a)') molfile
    read(u, *) lcmd
    read(u, *) dist
    read(u, *) incl_deg
    read(u, *) vrad
    read(u, *) nlines_run
    read(u, *) istart
    if (lcmd == 2) then
        read(u, *) imx
        read(u, *) imy
        read(u, *) isz
        read(u, *) szx
        read(u, *) szy
        read(u, *) phioff
        read(u, *) xoff
        read(u, *) yoff
        read(u, *) iaddstar
        if (isz /= 1) then
            write(*, '(a,i0)') 'Dont recognize size spec ', isz
            stop 13
        end if
    else if (lcmd /= 0) then
        write(*, '(a,i0)') 'Dont recognize command no. ', lcmd
        stop 13
    end if
    close(u)
    molfile = adjustl(molfile)
    molname = molfile(1:index(molfile, '.dat', back=.true.) - 1)

    ! molecule: levels (I5,F12.4,F7.1), transitions (3I5,E12.3)
    open(newunit=u, file=path(trim(molfile)), status='old', action='read')
    do i = 1, 5
        read(u, '(a)') line
    end do
    read(u, '(I6)') nlev
    read(u, '(a)') line
    allocate(E(nlev), g(nlev))
    do i = 1, nlev
        read(u, '(I5,F12.4,F7.1)') idum, E(i), g(i)
    end do
    read(u, '(a)') line
    read(u, '(I6)') nlin
    read(u, '(a)') line
    allocate(lup(nlin), llo(nlin), A(nlin), nu0(nlin), gu(nlin), gl(nlin))
    do i = 1, nlin
        read(u, '(3I5,E12.3)') idum, lup(i), llo(i), A(i)
    end do
    close(u)
    lup = lup - 1
    llo = llo - 1
    do i = 1, nlin
        gu(i) = g(lup(i) + 1)
        gl(i) = g(llo(i) + 1)
    end do

    ! level populations
    open(newunit=u, file=path('levelpop_' // trim(molname) // '.dat'), status='old', &
         action='read')
    read(u, *) idum, idum, k, idum
    allocate(pops(k, nth, nr))
    call skip_values(u, 2 * k)          ! level energies and weights
    read(u, *) pops
    close(u)
    nlev = k

    ! the dust scattering/PAH source (read_scatsource, source.F:494-546):
    ! scatsource.dat and quantsource.dat, each optional, summed; one value
    ! per record, frequency outer, theta, radius inner
    allocate(scat(nth, nr, nf))
    scat = 0.0d0
    have_scat = .false.
    call read_scat('scatsource.dat')
    call read_scat('quantsource.dat')
    ! the interstellar field for outer boundary type 3 (read_interstellfield,
    ! star.F:675-720)
    if (outer_bc == 3) then
        inquire(file=path('interstellfield.inp'), exist=fex)
        if (.not. fex) then
            write(*, '(a)') 'ERROR: outer BC type 3 without interstellfield.inp'
            stop 13
        end if
        open(newunit=u, file=path('interstellfield.inp'), status='old', action='read')
        read(u, *) k
        if (k /= nf) then
            write(*, '(a)') 'ERROR: interstellfield.inp inconsistent nr of frequencies'
            stop 13
        end if
        allocate(isrf(nf))
        do i = 1, nf
            read(u, *) isrf(i)
        end do
        close(u)
    end if

    ! host pre-pass with the reference's arithmetic
    allocate(jd(nlin, nth, nr), ad(nlin, nth, nr))
    call chk(rl_prepare_line_freqs(nlin, E, lup, llo, nu0), 'rl_prepare_line_freqs')
    if (have_scat) then
        call chk(rl_prepare_line_dust_scat(nr, nth, nspec, rhod, td, nf, cnu, kabs, kscat, &
                                           c_loc(scat), nlin, nu0, jd, ad), &
                 'rl_prepare_line_dust_scat')
    else
        call chk(rl_prepare_line_dust_scat(nr, nth, nspec, rhod, td, nf, cnu, kabs, kscat, &
                                           c_null_ptr, nlin, nu0, jd, ad), &
                 'rl_prepare_line_dust_scat')
    end if
    call chk(rl_prepare_star(nf, fnu, rstar, istar), 'rl_prepare_star')
    aks = rl_prepare_aksmax(nr * nth, lw, fix_lw)
    call chk(rl_passband(vmax, dv, nfr, vmax2), 'rl_passband')

    ! the trace
    ierr = rl_create(ctx, 0, iprec)
    if (ierr /= 0) then
        write(*, '(a,i0)') 'rl_create failed: ', ierr
        stop 900
    end if
    call chkc(rl_set_grid(ctx, nr, r, nth, th))
    call chkc(rl_set_star(ctx, rstar, nf, cnu, istar))
    call chkc(rl_set_medium(ctx, rho, abund, vel, lw, umass, aks))
    call chkc(rl_set_levels(ctx, nlev, pops))
    call chkc(rl_set_lines(ctx, nlin, nu0, A, gu, gl, lup, llo))
    call chkc(rl_set_dust_line(ctx, jd, ad))
    if (outer_bc == 3) call chkc(rl_set_interstellar_field(ctx, nf, cnu, isrf))
    call chkc(rl_set_outer_bc(ctx, outer_bc))
    inc(1) = incl_deg * 0.0174532925199d0
    allocate(nrays(1), nrings(1))
    if (lcmd == 2) then
        call chkc(rl_setup_rays_rectang(ctx, 1, inc, imx, imy, szx, szy, phioff, xoff, yoff, &
                                        iaddstar, nrays))
        allocate(items(nlines_run), warn(nlines_run))
        allocate(image(nfr * imx * imy * nlines_run), tau(nfr * imx * imy * nlines_run))
        do i = 1, nlines_run
            items(i) = rl_item(istart - 2 + i, 0, nfr, vmax2, RL_SKIP_NONE)
        end do
        call chkc(rl_trace_lines_posvel(ctx, nlines_run, items, image, tau, warn))
        call rl_destroy(ctx)
        call write_posvel()
        stop
    end if
    call chkc(rl_setup_rays_circular(ctx, 1, inc, nphi, b_extra, dbdr, nrays, nrings, &
                                     c_null_ptr, c_null_ptr))
    allocate(items(nlines_run), flux(nlines_run * nfr), warn(nlines_run))
    do i = 1, nlines_run
        items(i) = rl_item(istart - 2 + i, 0, nfr, vmax2, RL_SKIP_REFERENCE)
    end do
    call chkc(rl_trace_lines_forward(ctx, nlines_run, items, flux, c_null_ptr, warn))
    call rl_destroy(ctx)
    call write_spectrum()

contains

    subroutine read_scat(name)
        character(len=*), intent(in) :: name
        integer :: v, ifr, jt, jr, h(4)
        real(c_double) :: val
        inquire(file=path(name), exist=fex)
        if (.not. fex) return
        have_scat = .true.
        open(newunit=v, file=path(name), status='old', action='read')
        read(v, *) h
        do ifr = 1, nf
            do jt = 1, nth
                do jr = 1, nr
                    read(v, *) val
                    scat(jt, jr, ifr) = scat(jt, jr, ifr) + val
                end do
            end do
        end do
        close(v)
    end subroutine

    function path(name) result(p)
        character(len=*), intent(in) :: name
        character(len=1024) :: p
        p = trim(dir) // '/' // name
    end function

    subroutine read_opac(idx, isp)
        integer, intent(in) :: idx, isp
        integer :: v, nfo, nso
        character(len=32) :: num
        write(num, '(i0)') idx
        open(newunit=v, file=path('dustopac_' // trim(num) // '.inp'), status='old', &
             action='read')
        read(v, *) nfo, nso
        read(v, *) kabs(:, isp)
        read(v, *) kscat(:, isp)
        close(v)
    end subroutine

    ! skip n list-directed values (the energy and weight lines of levelpop)
    subroutine skip_values(v, n)
        integer, intent(in) :: v, n
        real(c_double) :: tmp(n)
        read(v, *) tmp
    end subroutine

    subroutine chk(ierr, what)
        integer(c_int), intent(in) :: ierr
        character(len=*), intent(in) :: what
        if (ierr /= 0) then
            write(*, '(a,a,i0)') what, ': stop ', ierr
            stop 13
        end if
    end subroutine

    subroutine chkc(ierr)
        integer(c_int), intent(in) :: ierr
        if (ierr /= 0) then
            write(*, '(a)') trim(rl_error_text(ctx))
            call rl_destroy(ctx)
            select case (ierr)
            case (13)
                stop 13
            case (124)
                stop 124
            case (749)
                stop 749
            case default
                stop 1
            end select
        end if
    end subroutine

    ! linespectrum_<molecule>.dat as header_line_spectrum and
    ! write_line_spectrum write it (telescope.F:1662-1804)
    subroutine write_spectrum()
        integer :: v, il, inu, ln
        real(c_double) :: nuf, passb, dnu, freq, velo
        character(len=200) :: name80, file80
        name80 = molname
        file80 = './' // trim(molfile)
        open(newunit=v, file=path('linespectrum_' // trim(molname) // '.dat'), &
             status='replace', action='write')
        write(v, '(i2)') 1
        write(v, '(i2)') 1
        write(v, '(a80)') name80
        write(v, '(a80)') file80
        write(v, '(i10)') nlines_run, nfr
        write(v, '(e12.4,e12.4,f7.3)') 1.0d0, vrad, inc(1) * 57.2957795132d0
        do il = 1, nlines_run
            ln = istart - 1 + il
            nuf = nu0(ln)
            passb = 3.33567d-6 * abs(nuf) * vmax2
            dnu = 2.0d0 * passb / (nfr - 1.0d0)
            write(v, *)
            write(v, '(i5,i5)') lup(ln) + 1, llo(ln) + 1
            write(v, '(e14.9)') nuf
            write(v, '(e10.5)') 0.0d0
            write(v, '(i5)') nfr
            write(v, *) ' '
            do inu = nfr, 1, -1
                freq = nuf + ((0.0d0 - passb) + (inu - 1) * dnu)
                velo = freq - nuf
                velo = -2.99792458d5 * velo / nuf
                velo = velo + vrad
                write(v, '(E13.6,1X,E13.6)') velo, flux((il - 1) * nfr + inu)
            end do
        end do
        close(v)
    end subroutine

    ! lineposvel_<molecule>_<line>.dat as calc_write_line_posvel writes it
    ! (telescope.F:1940-2019): the same records and edit descriptors; the
    ! cube of line il is image((il-1)*nfr*imx*imy + ((iy-1)*imx + ix-1)*nfr
    ! + inu), imrec_int(inu, ix, iy)'s order
    subroutine write_posvel()
        integer :: v, il, ln, inu, ix, iy
        integer(c_int64_t) :: o
        real(c_double) :: nuf, passb, dnu, temp, spx, spy
        real(c_double), allocatable :: velo(:)
        character(len=12) :: ch
        character(len=512) :: fname
        character(len=200) :: name80, file80
        name80 = molname
        file80 = './' // trim(molfile)
        spx = (szx * 0.5d0) / (imx * 0.5d0)
        spy = (szy * 0.5d0) / (imy * 0.5d0)
        allocate(velo(nfr))
        do il = 1, nlines_run
            ln = istart - 1 + il
            write(ch, '(i2)') ln
            if (ln < 10) then
                fname = 'lineposvel_' // trim(molname) // '_' // ch(2:2) // '.dat'
            else if (ln < 100) then
                fname = 'lineposvel_' // trim(molname) // '_' // ch(2:3) // '.dat'
            else
                stop 177
            end if
            nuf = nu0(ln)
            passb = 3.33567d-6 * abs(nuf) * vmax2
            dnu = 2.0d0 * passb / (nfr - 1.0d0)
            open(newunit=v, file=path(fname(1:len_trim(fname))), status='replace', &
                 action='write')
            write(v, *)
            write(v, '(I2)') 1
            write(v, '(A80)') name80
            write(v, '(A80)') file80
            write(v, '(ES12.4,ES12.4,F7.3)') 1.0d0, vrad, inc(1) * 57.2957795132d0
            write(v, '(I5,I5)') lup(ln) + 1, llo(ln) + 1
            write(v, '(ES12.4)') nuf
            write(v, '(I5)') nfr
            write(v, '(I5,I5,ES12.4,ES12.4,F7.3,ES12.4,ES12.4)') imx, imy, spx, spy, phioff, &
                xoff, yoff
            write(v, *) ' '
            do inu = 1, nfr
                velo(inu) = (abs(nuf) + ((0.0d0 - passb) + (inu - 1) * dnu)) - nuf
                velo(inu) = -2.99792458d5 * velo(inu) / nuf
                velo(inu) = velo(inu) + vrad
            end do
            write(v, '(E10.3,1X)') (velo(inu), inu = 1, nfr)
            write(v, *) ' '
            do inu = 1, nfr
                do iy = 1, imy
                    do ix = 1, imx
                        o = int(il - 1, c_int64_t) * nfr * imx * imy + &
                            (int(iy - 1, c_int64_t) * imx + (ix - 1)) * nfr + inu
                        temp = 3.25465503368d36 / nuf**2
                        temp = temp * image(o)
                        write(v, '(ES12.5,1X,ES12.5,1X)') temp, tau(o)
                    end do
                end do
                write(v, *) ' '
            end do
            close(v)
        end do
    end subroutine

end program rl_host
