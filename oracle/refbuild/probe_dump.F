c     ------------------------------------------------------------------
c     Full-precision probe for the reference RADLite build (oracle test
c     infrastructure; never part of the product library).
c
c     The reference prints its spectrum only as E13.6
c     (telescope.F:1802).  The link step wraps write_line_spectrum
c     (telescope.F:1703, called once per line from main.F:1047), so this
c     routine runs right after each calc_line_spectrum (telescope.F:1465)
c     and reads the reference's own COMMON blocks, declared by including
c     the reference headers in place.  It appends one binary record per
c     line to probe_out.bin:
c
c       int32  iline, nfr, rays_amount, nrr, nphi, maserflag, overshoot
c       f64    spec_flux_observer(1:nfr)          telescope.F:1435
c       f64    line_dnu(1:nfr,iline), line_nu0(iline)
c       f64    imcir_r(0:nrr), imcir_ri(0:nrr+1)  telescope.F:434-488
c       f64    imcir_int(1:nfr,1:nphi,0:nrr)      telescope.F:498-621
c       int32  imcir_cmask(1:nfr,1:nphi,1:nrr)
c       f32    minvel(1:rays_amount), maxvel(1:rays_amount)
c       f64    rp_x0, rp_z0 (1:rays_amount), rp_theta0(1), aksmax
c
c     If probe_rays.inp exists (a count, then ray indices), the probe
c     also re-runs make_trajectory_c (telescope.F:2787) for those rays
c     and appends, per ray, to probe_traj.bin:
c       int32  iray, tr_amount
c       f64    tr_s, tr_radius, tr_theta, tr_mu, tr_phi (1:tr_amount)
c       int32  tr_icross, tr_iradius, tr_itheta (1:tr_amount)
c     and then calls the real write_line_spectrum.
c     ------------------------------------------------------------------
#include "main.h"
      subroutine probe_write_line_spectrum(istyle)
     %     bind(C, name='__wrap_write_line_spectrum_')
      implicit none
      integer istyle
c
#include "common_grid.h"
#include "common_telescope.h"
#include "common_lines.h"
#include "common_radfield.h"
c
      interface
         subroutine real_write_line_spectrum(istyle)
     %        bind(C, name='__real_write_line_spectrum_')
         integer istyle
         end subroutine
      end interface
c
      integer nfr,nrr,nphi,nray,il,inu,ip,ir,i,k,nsel,iray
      integer hdr(7)
      logical fex
c
      il   = ispec_currentline
      nfr  = freq_nr
      nrr  = imcir_nrr
      nphi = imcir_nrphi
      nray = rays_amount
      hdr(1) = il
      hdr(2) = nfr
      hdr(3) = nray
      hdr(4) = nrr
      hdr(5) = nphi
      hdr(6) = maserflag
      hdr(7) = transerr_overshoot
      open(unit=77,file='probe_out.bin',access='stream',
     %     form='unformatted',position='append',status='unknown')
      write(77) (hdr(i),i=1,7)
      write(77) (spec_flux_observer(inu),inu=1,nfr)
      write(77) (line_dnu(inu,il),inu=1,nfr), line_nu0(il)
      write(77) (imcir_r(ir),ir=0,nrr)
      write(77) (imcir_ri(ir),ir=0,nrr+1)
      write(77) (((imcir_int(inu,ip,ir),inu=1,nfr),ip=1,nphi),
     %     ir=0,nrr)
      write(77) (((imcir_cmask(inu,ip,ir),inu=1,nfr),ip=1,nphi),
     %     ir=1,nrr)
      write(77) (minvel(i),i=1,nray), (maxvel(i),i=1,nray)
      write(77) (rp_x0(i),i=1,nray), (rp_z0(i),i=1,nray),
     %     rp_theta0(1), aksmax
      close(77)
c
      inquire(file='probe_rays.inp',exist=fex)
      if(fex) then
         open(unit=78,file='probe_rays.inp',status='old')
         read(78,*) nsel
         open(unit=79,file='probe_traj.bin',access='stream',
     %        form='unformatted',position='append',status='unknown')
         do k=1,nsel
            read(78,*) iray
            call make_trajectory_c(rp_x0(iray),rp_z0(iray),
     %           rp_theta0(iray),rp_s0(iray),0)
            write(79) iray, tr_amount
            write(79) (tr_s(i),i=1,tr_amount),
     %           (tr_radius(i),i=1,tr_amount),
     %           (tr_theta(i),i=1,tr_amount),
     %           (tr_mu(i),i=1,tr_amount),
     %           (tr_phi(i),i=1,tr_amount)
            write(79) (tr_icross(i),i=1,tr_amount),
     %           (tr_iradius(i),i=1,tr_amount),
     %           (tr_itheta(i),i=1,tr_amount)
         enddo
         close(79)
         close(78)
      endif
c
      call real_write_line_spectrum(istyle)
      end

c     ------------------------------------------------------------------
c     Position-velocity cubes: the link step also wraps
c     calc_write_line_posvel (telescope.F:1828, called per line from
c     main.F:1071).  After the real routine (which renders the
c     rectangular image, make_freq_image_rectang telescope.F:2061, and
c     writes lineposvel_*.dat at 6 digits) this appends, per line, to
c     probe_posvel.bin:
c       int32  iline, nfr, nrx, nry, rays_amount, imrec_starunres,
c              maserflag, overshoot
c       f64    line_nu0(iline), line_dnu(1:nfr,iline)
c       f64    imrec_int(1:nfr,1:nrx,1:nry), imrec_tau(1:nfr,1:nrx,1:nry)
c       f64    rp_x0, rp_z0, rp_b (1:rays_amount), rp_theta0(1)
c     ------------------------------------------------------------------
      subroutine probe_calc_write_line_posvel(iline,distance,
     %     passband,nfr) bind(C, name='__wrap_calc_write_line_posvel_')
      implicit none
      integer iline,nfr
      doubleprecision distance,passband
c
#include "common_grid.h"
#include "common_telescope.h"
#include "common_lines.h"
#include "common_radfield.h"
c
      interface
         subroutine real_calc_write_line_posvel(iline,distance,
     %        passband,nfr)
     %        bind(C, name='__real_calc_write_line_posvel_')
         integer iline,nfr
         doubleprecision distance,passband
         end subroutine
      end interface
c
      integer hdr(8),i,inu,ix,iy,nf,nray
c
      call real_calc_write_line_posvel(iline,distance,passband,nfr)
      nf   = freq_nr
      nray = rays_amount
      hdr(1) = iline
      hdr(2) = nf
      hdr(3) = imrec_nrx
      hdr(4) = imrec_nry
      hdr(5) = nray
      hdr(6) = imrec_starunres
      hdr(7) = maserflag
      hdr(8) = transerr_overshoot
      open(unit=77,file='probe_posvel.bin',access='stream',
     %     form='unformatted',position='append',status='unknown')
      write(77) (hdr(i),i=1,8)
      write(77) line_nu0(iline), (line_dnu(inu,iline),inu=1,nf)
      write(77) (((imrec_int(inu,ix,iy),inu=1,nf),ix=1,imrec_nrx),
     %     iy=1,imrec_nry)
      write(77) (((imrec_tau(inu,ix,iy),inu=1,nf),ix=1,imrec_nrx),
     %     iy=1,imrec_nry)
      write(77) (rp_x0(i),i=1,nray), (rp_z0(i),i=1,nray),
     %     (rp_b(i),i=1,nray), rp_theta0(1)
      close(77)
      end

c     ------------------------------------------------------------------
c     Timing: the link step also wraps calc_line_spectrum
c     (telescope.F:1465, the per-line trace, called from main.F:1047);
c     the wall time of the real routine is appended, per line, to
c     probe_time.txt as "iline seconds", so that the reference's trace
c     can be timed apart from its set-up and file output.
c     ------------------------------------------------------------------
      subroutine probe_calc_line_spectrum(iline,distance,passband,nfr)
     %     bind(C, name='__wrap_calc_line_spectrum_')
      implicit none
      integer iline,nfr
      doubleprecision distance,passband
      integer*8 c0,c1,crate
      interface
         subroutine real_calc_line(iline,distance,passband,nfr)
     %        bind(C, name='__real_calc_line_spectrum_')
         integer iline,nfr
         doubleprecision distance,passband
         end subroutine
      end interface
      call system_clock(c0,crate)
      call real_calc_line(iline,distance,passband,nfr)
      call system_clock(c1)
      open(unit=93,file='probe_time.txt',status='unknown',
     %     position='append')
      write(93,'(I6,1X,ES24.16)') iline, dble(c1-c0)/dble(crate)
      close(93)
      return
      end
