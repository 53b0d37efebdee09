"""ctypes binding of the CPU oracle (oracle/liborc.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker / timed CPU baseline.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

dp = C.POINTER(C.c_double)
ip = C.POINTER(C.c_int)
fp = C.POINTER(C.c_float)


class OrcModel(C.Structure):
    _fields_ = [
        ("nr", C.c_int), ("nth_half", C.c_int),
        ("r", dp), ("theta_half", dp),
        ("rho", dp), ("abund", dp), ("vel", dp), ("lw", dp),
        ("nlev", C.c_int), ("pops", dp), ("umass_av", C.c_double),
        ("nlines", C.c_int), ("nu0", dp), ("A", dp), ("gu", dp), ("gl", dp),
        ("lev_up", ip), ("lev_low", ip), ("jd", dp), ("ad", dp),
        ("aksmax", C.c_double), ("rstar", C.c_double), ("ncont", C.c_int),
        ("cont_nu", dp), ("cont_istar", dp), ("outer_bc", C.c_int),
        ("cont_isrf", dp),
    ]


class OrcRays(C.Structure):
    _fields_ = [("nrays", C.c_int), ("nrings", C.c_int), ("nphi", C.c_int),
                ("theta0", C.c_double), ("x0", dp), ("z0", dp),
                ("ring_r", dp), ("ring_ri", dp)]


class OrcTraj(C.Structure):
    _fields_ = [("n", C.c_int), ("s", dp), ("r", dp), ("theta", dp),
                ("mu", dp), ("phi", dp), ("icross", ip), ("ir", ip),
                ("it", ip), ("b", C.c_double)]


class OrcGrid(C.Structure):
    _fields_ = [("nr", C.c_int), ("nth", C.c_int), ("rc", dp), ("tc", dp)]


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liborc.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liborc.so is not built (make -C oracle)")
        L = C.CDLL(path)
        L.orc_line_spectrum.restype = C.c_int
        L.orc_bplanck.restype = C.c_double
        L.orc_voigt.restype = C.c_double
        L.orc_voigt.argtypes = [C.c_double] * 3
        L.orc_bplanck.argtypes = [C.c_double] * 2
        L.orc_qdr_src_2.restype = C.c_double
        L.orc_qdr_src_2.argtypes = [C.c_double] * 6 + [C.POINTER(C.c_long)]
        L.orc_aksmax.restype = C.c_double
        L.orc_aksmax.argtypes = [C.c_int, dp, C.c_double]
        L.orc_hunt.restype = C.c_int
        L.orc_hunt.argtypes = [dp, C.c_int, C.c_double, C.c_int]
        L.orc_setup_rays.argtypes = [C.POINTER(OrcRays), C.POINTER(OrcGrid),
                                     C.c_double, C.c_int, C.c_int, C.c_int,
                                     C.c_double]
        L.orc_trajectory.argtypes = [C.POINTER(OrcTraj), C.POINTER(OrcGrid),
                                     C.c_double, C.c_double, C.c_double,
                                     C.c_double, C.c_double]
        L.orc_setup_rays_rect.argtypes = [C.POINTER(OrcRays), dp, C.c_double, C.c_int, C.c_int,
                                          C.c_double, C.c_double, C.c_double, C.c_double,
                                          C.c_double, C.c_int, C.c_double, dp, dp,
                                          C.POINTER(C.c_int)]
        L.orc_posvel.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(OrcRays), dp, C.c_int,
                                 C.c_int, C.c_double, C.c_double, C.c_int, C.c_int,
                                 C.c_double, C.c_int, C.c_int, dp, dp, C.POINTER(C.c_int),
                                 C.POINTER(C.c_long)]
        L.orc_passband_nfr.restype = C.c_int
        L.orc_passband_nfr.argtypes = [C.c_double, C.c_double, dp]
        _LIB = L
    return _LIB


def _p(a, t=dp):
    return a.ctypes.data_as(t)


class Prepared:
    """Engine inputs (the C-ABI tables) computed from a run directory
    with the reference's host arithmetic (R13), in C."""

    def __init__(self, ri):
        L = lib()
        self.ri = ri
        self.nr, self.nth = ri.rho_gas.shape
        nlines = len(ri.A)
        self.lev_up = np.ascontiguousarray(ri.lev_up - 1, dtype=np.int32)
        self.lev_low = np.ascontiguousarray(ri.lev_low - 1, dtype=np.int32)
        self.nu0 = np.zeros(nlines)
        E = np.ascontiguousarray(ri.E_cm, dtype=np.float64)
        L.orc_line_consts(C.c_int(len(E)), _p(E), C.c_int(nlines),
                          _p(self.lev_up, ip), _p(self.lev_low, ip), _p(self.nu0))
        self.A = np.ascontiguousarray(ri.A, dtype=np.float64)
        self.gu = np.ascontiguousarray(ri.g[ri.lev_up - 1], dtype=np.float64)
        self.gl = np.ascontiguousarray(ri.g[ri.lev_low - 1], dtype=np.float64)
        nsp = ri.rho_dust.shape[0]
        self.jd = np.zeros((self.nr, self.nth, nlines))
        self.ad = np.zeros((self.nr, self.nth, nlines))
        rd = np.ascontiguousarray(ri.rho_dust); td = np.ascontiguousarray(ri.t_dust)
        nu = np.ascontiguousarray(ri.cont_nu); ka = np.ascontiguousarray(ri.kappa_abs)
        ks = np.ascontiguousarray(ri.kappa_scat)
        sc = getattr(ri, "scat_src", None)
        self.scat = None if sc is None else np.ascontiguousarray(sc, dtype=np.float64)
        L.orc_line_dust(C.c_int(self.nr), C.c_int(self.nth), C.c_int(nsp), _p(rd), _p(td),
                        C.c_int(len(nu)), _p(nu), _p(ka), _p(ks),
                        _p(self.scat) if self.scat is not None else None, C.c_int(nlines),
                        _p(self.nu0), _p(self.jd), _p(self.ad))
        isrf = getattr(ri, "isrf", None)
        self.isrf = None if isrf is None else np.ascontiguousarray(isrf, dtype=np.float64)
        self.istar = np.zeros(len(nu))
        fnu = np.ascontiguousarray(ri.star_fnu)
        L.orc_star_intensity(C.c_int(len(nu)), _p(fnu), C.c_double(ri.rstar), _p(self.istar))
        lw = np.ascontiguousarray(ri.lw_kms)
        self.aksmax = L.orc_aksmax(C.c_int(lw.size), _p(lw), C.c_double(ri.fix_lw))
        vm = C.c_double(0.0)
        self.nfr = L.orc_passband_nfr(C.c_double(ri.vmax), C.c_double(ri.dv), C.byref(vm))
        self.vmax = vm.value
        self.incl_rad = ri.incl_deg * 0.0174532925199
        self.r = np.ascontiguousarray(ri.r)
        self.theta = np.ascontiguousarray(ri.theta_half)
        self.rho = np.ascontiguousarray(ri.rho_gas)
        self.abund = np.ascontiguousarray(ri.abund)
        self.vel = np.ascontiguousarray(ri.vel)
        self.lw = lw
        self.pops = np.ascontiguousarray(ri.pops)
        self.cont_nu = nu


class Oracle:
    def __init__(self, prep: Prepared):
        self.p = prep
        L = lib()
        p = prep
        self.m = OrcModel(
            nr=p.nr, nth_half=p.nth, r=_p(p.r), theta_half=_p(p.theta),
            rho=_p(p.rho), abund=_p(p.abund), vel=_p(p.vel), lw=_p(p.lw),
            nlev=p.pops.shape[2], pops=_p(p.pops), umass_av=p.ri.umass_av,
            nlines=len(p.A), nu0=_p(p.nu0), A=_p(p.A), gu=_p(p.gu), gl=_p(p.gl),
            lev_up=_p(p.lev_up, ip), lev_low=_p(p.lev_low, ip), jd=_p(p.jd),
            ad=_p(p.ad), aksmax=p.aksmax, rstar=p.ri.rstar, ncont=len(p.cont_nu),
            cont_nu=_p(p.cont_nu), cont_istar=_p(p.istar), outer_bc=p.ri.outer_bc,
            cont_isrf=_p(p.isrf) if p.isrf is not None else None)
        self.g = OrcGrid()
        rc = L.orc_grid_build(C.byref(self.g), C.c_int(p.nr), _p(p.r), C.c_int(p.nth), _p(p.theta))
        if rc:
            raise RuntimeError(f"orc_grid_build: {rc}")
        self._rays = {}

    def rays(self, incl_rad, nphi, nrext, dbdr):
        key = (incl_rad, nphi, nrext, dbdr)
        if key not in self._rays:
            r = OrcRays()
            rc = lib().orc_setup_rays(C.byref(r), C.byref(self.g), incl_rad, nphi, nrext,
                                      dbdr, self.p.ri.rstar)
            if rc:
                raise RuntimeError(f"orc_setup_rays: {rc}")
            self._rays[key] = r
        return self._rays[key]

    def trajectory(self, rays, iray, rs0_fill=0.0):
        t = OrcTraj()
        rc = lib().orc_trajectory(C.byref(t), C.byref(self.g), self.p.ri.rstar,
                                  rays.x0[iray], rays.z0[iray], rays.theta0, rs0_fill)
        if rc:
            raise RuntimeError(f"orc_trajectory: {rc}")
        n = t.n
        out = dict(n=n, b=t.b)
        for k in ("s", "r", "theta", "mu", "phi"):
            out[k] = np.ctypeslib.as_array(getattr(t, k), shape=(n,)).copy()
        for k in ("icross", "ir", "it"):
            out[k] = np.ctypeslib.as_array(getattr(t, k), shape=(n,)).copy()
        lib().orc_traj_free(C.byref(t))
        return out

    def posvel(self, iline, pv, incl_rad=None, nfr=None, vmax=None, nthreads=0):
        """Position-velocity cube of a line (telescope.F:1828-2183): dict
        with image and tau [ny, nx, nfr], the rays (x0, z0, b) and the
        pixel sizes.  pv: nx, ny, size_x, size_y, phioff, xoff, yoff,
        addstar (linespectrum.inp, telescope.F:196-224)."""
        p = self.p
        incl_rad = p.incl_rad if incl_rad is None else incl_rad
        nfr = p.nfr if nfr is None else nfr
        vmax = p.vmax if vmax is None else vmax
        nx, ny = int(pv["nx"]), int(pv["ny"])
        n = nx * ny + (1 if pv["addstar"] > 0 else 0)
        rays = OrcRays()
        b = np.zeros(n)
        spx, spy, unres = C.c_double(0), C.c_double(0), C.c_int(0)
        rc = lib().orc_setup_rays_rect(C.byref(rays), _p(b), incl_rad, nx, ny, pv["size_x"],
                                       pv["size_y"], pv["phioff"], pv["xoff"], pv["yoff"],
                                       int(pv["addstar"]), p.ri.rstar, C.byref(spx),
                                       C.byref(spy), C.byref(unres))
        if rc:
            raise RuntimeError(f"orc_setup_rays_rect: stop code {rc}")
        img = np.zeros((ny, nx, nfr))
        tau = np.zeros((ny, nx, nfr))
        maser, ovs = C.c_int(0), C.c_long(0)
        rc = lib().orc_posvel(C.cast(C.byref(self.m), C.c_void_p),
                              C.cast(C.byref(self.g), C.c_void_p), C.byref(rays), _p(b),
                              nx, ny, spx.value, spy.value, unres.value, iline, vmax, nfr,
                              nthreads, _p(img), _p(tau), C.byref(maser), C.byref(ovs))
        out = dict(image=img, tau=tau, b=b, spx=spx.value, spy=spy.value,
                   starunres=unres.value, maser=maser.value, overshoot=ovs.value,
                   x0=np.ctypeslib.as_array(rays.x0, shape=(n,)).copy(),
                   z0=np.ctypeslib.as_array(rays.z0, shape=(n,)).copy(), theta0=rays.theta0)
        lib().orc_rays_free(C.byref(rays))
        if rc:
            raise RuntimeError(f"orc_posvel: stop code {rc}")
        return out

    def line_spectrum(self, iline, incl_rad=None, nphi=None, nrext=None, dbdr=None,
                      nfr=None, vmax=None, skip_mode=0, nthreads=0, want_image=True,
                      ray_stride=1):
        p = self.p
        incl_rad = p.incl_rad if incl_rad is None else incl_rad
        nphi = p.ri.nphi if nphi is None else nphi
        nrext = p.ri.b_extra if nrext is None else nrext
        dbdr = p.ri.dbdr if dbdr is None else dbdr
        nfr = p.nfr if nfr is None else nfr
        vmax = p.vmax if vmax is None else vmax
        rays = self.rays(incl_rad, nphi, nrext, dbdr)
        nrays = rays.nrays
        flux = np.zeros(nfr)
        image = np.zeros((nrays, nfr)) if want_image else None
        cmask = np.zeros((nrays, nfr), dtype=np.int32) if want_image else None
        mn = np.zeros(nrays, dtype=np.float32)
        mx = np.zeros(nrays, dtype=np.float32)
        dnu = np.zeros(nfr)
        maser = C.c_int(0); ovs = C.c_long(0); nint = C.c_long(0)
        rc = lib().orc_line_spectrum(
            C.byref(self.m), C.byref(self.g), C.byref(rays), C.c_int(iline),
            C.c_double(vmax), C.c_int(nfr), C.c_int(skip_mode), C.c_int(nthreads),
            C.c_int(ray_stride),
            _p(flux), _p(image) if want_image else None,
            _p(cmask, ip) if want_image else None, _p(mn, fp), _p(mx, fp), _p(dnu),
            C.byref(maser), C.byref(ovs), C.byref(nint))
        if rc:
            raise RuntimeError(f"orc_line_spectrum: stop code {rc}")
        return dict(flux=flux, image=image, cmask=cmask, minvel=mn, maxvel=mx, dnu=dnu,
                    maser=maser.value, overshoot=ovs.value, n_integrated=nint.value,
                    nrays=nrays, nrings=rays.nrings,
                    ring_ri=np.ctypeslib.as_array(rays.ring_ri, shape=(rays.nrings + 2,)).copy(),
                    ring_r=np.ctypeslib.as_array(rays.ring_r, shape=(rays.nrings + 1,)).copy(),
                    x0=np.ctypeslib.as_array(rays.x0, shape=(nrays,)).copy(),
                    z0=np.ctypeslib.as_array(rays.z0, shape=(nrays,)).copy(),
                    theta0=rays.theta0)
