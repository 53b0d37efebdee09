"""Python host of the MI355X raytracer: ctypes binding of librl.so.

This mirrors the reference's operator interface for the image/spectrum
raytrace: a model (grid, medium, levels, lines, dust, star) is loaded
once, a circular-image ray set is built per inclination
(setup_rays_circular, telescope.F:715), and calc_line_spectrum
(telescope.F:1465) is run for a batch of (line, inclination) items.
Errors carry the reference's stop codes.

librl.so must be built (``python -m radlite_amd.build``); there is no
CPU fallback -- importing the engine without the library raises.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

from . import l4io

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "librl.so")
_LIB = None

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)

RL_FP64, RL_FP32 = 0, 1
OPT_GRAPH, OPT_AUTOTUNE, OPT_TILE, OPT_PERSISTENT, OPT_CKPT = 1, 2, 3, 4, 5
SKIP_REFERENCE, SKIP_NONE = 0, 1
Q_TABLE = -1.0          # set_physics: the partition sum from set_partition_table
WARN_MASER, WARN_OVERSHOOT = 1, 2


class RLItem(C.Structure):
    _fields_ = [("line", C.c_int), ("incl", C.c_int), ("nfr", C.c_int),
                ("vmax_kms", C.c_double), ("skip_mode", C.c_int)]


class RadliteError(RuntimeError):
    """A reference stop code (or RL_E_HIP=900) raised by the engine."""

    def __init__(self, code, msg):
        super().__init__(f"stop {code}: {msg}")
        self.code = code


def load_library(path=LIB_PATH):
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build it with python -m radlite_amd.build")
    L = C.CDLL(path)
    vp = C.c_void_p
    sig = {
        "rl_create": (C.c_int, [C.POINTER(vp), C.c_int, C.c_int]),
        "rl_destroy": (None, [vp]),
        "rl_last_error": (C.c_char_p, [vp]),
        "rl_set_grid": (C.c_int, [vp, C.c_int, _dp, C.c_int, _dp]),
        "rl_set_medium": (C.c_int, [vp, _dp, _dp, _dp, _dp, C.c_double, C.c_double]),
        "rl_set_levels": (C.c_int, [vp, C.c_int, _dp]),
        "rl_set_lines": (C.c_int, [vp, C.c_int, _dp, _dp, _dp, _dp, _ip, _ip]),
        "rl_set_dust_line": (C.c_int, [vp, _dp, _dp]),
        "rl_set_star": (C.c_int, [vp, C.c_double, C.c_int, _dp, _dp]),
        "rl_set_outer_bc": (C.c_int, [vp, C.c_int]),
        "rl_setup_rays_circular": (C.c_int, [vp, C.c_int, _dp, C.c_int, C.c_int, C.c_int,
                                             _ip, _ip, _dp, _dp]),
        "rl_setup_rays_rectang": (C.c_int, [vp, C.c_int, _dp, C.c_int, C.c_int, C.c_double,
                                            C.c_double, C.c_double, C.c_double, C.c_double,
                                            C.c_int, _ip]),
        "rl_trace_lines_forward": (C.c_int, [vp, C.c_int, C.POINTER(RLItem), _dp, _dp,
                                             C.POINTER(C.c_uint)]),
        "rl_trace_lines_posvel": (C.c_int, [vp, C.c_int, C.POINTER(RLItem), _dp, _dp,
                                            C.POINTER(C.c_uint)]),
        "rl_trace_lines_forward_noalloc": (C.c_int, [vp, C.c_int, C.POINTER(RLItem), vp, vp,
                                                     C.POINTER(C.c_uint)]),
        "rl_resample_image": (C.c_int, [vp, C.c_int, C.c_int, _dp, C.c_int, C.c_int, _ip, _dp,
                                        _dp]),
        "rl_trace_lines_forward_ex": (C.c_int, [vp, C.c_int, C.POINTER(RLItem), _dp, _dp,
                                                C.POINTER(C.c_ubyte), C.POINTER(C.c_uint)]),
        "rl_trace_lines_adjoint": (C.c_int, [vp, C.c_int, C.POINTER(RLItem), _dp, _dp,
                                             _dp, _dp, _dp, _dp, _dp, _dp]),
        "rl_trace_lines_forward_tape": (C.c_int, [vp, C.c_int, C.POINTER(RLItem), _dp]),
        "rl_tape_release": (C.c_int, [vp]),
        "rl_last_adjoint_info": (C.c_int, [vp, _ip, _ip, _ip]),
        "rl_memory_info": (C.c_int, [vp, _dp]),
        "rl_plan": (C.c_int, [vp, C.c_int, C.POINTER(RLItem), _dp]),
        "rl_set_level_data": (C.c_int, [vp, C.c_int, _dp, _dp]),
        "rl_set_dust_species": (C.c_int, [vp, C.c_int, _dp, _dp, _dp]),
        "rl_set_physics": (C.c_int, [vp, _dp, _dp, _dp, _dp, C.c_double, C.c_double]),
        "rl_physics_adjoint": (C.c_int, [vp, _dp, _dp, _dp, _dp]),
        "rl_set_partition_table": (C.c_int, [vp, C.c_int, _dp, _dp]),
        "rl_set_turbulence_alpha": (C.c_int, [vp, C.c_double, C.c_double, C.c_double]),
        "rl_set_velocity_kepler": (C.c_int, [vp, C.c_double]),
        "rl_get_tables": (C.c_int, [vp, _dp, _dp, _dp, _dp, _dp]),
        "rl_sync": (C.c_int, [vp]),
        "rl_set_memory_budget": (C.c_int, [vp, C.c_double]),
        "rl_set_option": (C.c_int, [vp, C.c_int, C.c_int]),
        "rl_last_batch_info": (C.c_int, [vp, _ip, _ip, _ip]),
        "rl_last_stats": (C.c_int, [vp, _dp, C.POINTER(C.c_longlong)]),
        "rl_prepare_line_freqs": (C.c_int, [C.c_int, _dp, _ip, _ip, _dp]),
        "rl_prepare_line_dust": (C.c_int, [C.c_int, C.c_int, C.c_int, _dp, _dp, C.c_int,
                                           _dp, _dp, _dp, C.c_int, _dp, _dp, _dp]),
        "rl_prepare_line_dust_scat": (C.c_int, [C.c_int, C.c_int, C.c_int, _dp, _dp, C.c_int,
                                                _dp, _dp, _dp, _dp, C.c_int, _dp, _dp, _dp]),
        "rl_set_interstellar_field": (C.c_int, [vp, C.c_int, _dp, _dp]),
        "rl_set_dust_scatter": (C.c_int, [vp, C.c_int, _dp, _dp]),
        "rl_prepare_star": (C.c_int, [C.c_int, _dp, C.c_double, _dp]),
        "rl_prepare_aksmax": (C.c_double, [C.c_int, _dp, C.c_double]),
        "rl_passband": (C.c_int, [C.c_double, C.c_double, _ip, _dp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _LIB = L
    return L


def _d(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(_dp)


def _i(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(_ip)


@dataclass
class LineTables:
    """Host pre-pass results (the engine's inputs besides the fields)."""
    nu0: np.ndarray
    A: np.ndarray
    gu: np.ndarray
    gl: np.ndarray
    lev_up: np.ndarray     # 0-based
    lev_low: np.ndarray
    jd: np.ndarray         # [nr, nth, nlines]
    ad: np.ndarray
    istar: np.ndarray
    cont_nu: np.ndarray
    aksmax: float


def prepare_tables(ri: l4io.RunInputs) -> LineTables:
    """R13 pre-pass (line.F:1980-1983, 3560-3607; star.F:508-525;
    line.F:3020-3033) with the reference's arithmetic, in librl."""
    L = load_library()
    nl = len(ri.A)
    lu, lup = _i(ri.lev_up - 1)
    ll, llp = _i(ri.lev_low - 1)
    E, Ep = _d(ri.E_cm)
    nu0 = np.zeros(nl)
    rc = L.rl_prepare_line_freqs(nl, Ep, lup, llp, nu0.ctypes.data_as(_dp))
    if rc:
        raise RadliteError(rc, "rl_prepare_line_freqs")
    nr, nth = ri.rho_gas.shape
    jd = np.zeros((nr, nth, nl))
    ad = np.zeros((nr, nth, nl))
    rd, rdp = _d(ri.rho_dust)
    td, tdp = _d(ri.t_dust)
    nu, nup = _d(ri.cont_nu)
    ka, kap = _d(ri.kappa_abs)
    ks, ksp = _d(ri.kappa_scat)
    sc = getattr(ri, "scat_src", None)
    scp = None
    if sc is not None:
        sc, scp = _d(sc)
    rc = L.rl_prepare_line_dust_scat(nr, nth, ri.rho_dust.shape[0], rdp, tdp, len(nu), nup, kap,
                                     ksp, scp, nl, nu0.ctypes.data_as(_dp),
                                     jd.ctypes.data_as(_dp), ad.ctypes.data_as(_dp))
    if rc:
        raise RadliteError(rc, "rl_prepare_line_dust_scat")
    istar = np.zeros(len(nu))
    fn, fnp = _d(ri.star_fnu)
    rc = L.rl_prepare_star(len(nu), fnp, ri.rstar, istar.ctypes.data_as(_dp))
    if rc:
        raise RadliteError(rc, "rl_prepare_star")
    lw, lwp = _d(ri.lw_kms)
    aks = L.rl_prepare_aksmax(lw.size, lwp, ri.fix_lw)
    return LineTables(nu0=nu0, A=np.asarray(ri.A, float), gu=ri.g[ri.lev_up - 1].astype(float),
                      gl=ri.g[ri.lev_low - 1].astype(float), lev_up=lu, lev_low=ll, jd=jd, ad=ad,
                      istar=istar, cont_nu=nu, aksmax=aks)


def passband(vmax, dv):
    """(nfr, vmax) as main.F:209-210 derive them."""
    L = load_library()
    n = C.c_int(0)
    v = C.c_double(0.0)
    rc = L.rl_passband(vmax, dv, C.byref(n), C.byref(v))
    if rc:
        raise RadliteError(rc, "rl_passband")
    return n.value, v.value


class Raytracer:
    """One engine context on one GPU."""

    def __init__(self, device=0, precision="fp64"):
        self.L = load_library()
        self.ctx = C.c_void_p()
        prec = {"fp64": RL_FP64, "fp32": RL_FP32}[precision]
        self.precision = precision
        rc = self.L.rl_create(C.byref(self.ctx), device, prec)
        if rc:
            raise RadliteError(rc, "rl_create (no HIP device?)")
        self.nrays = []
        self.nrings = []
        self.ring_r = None
        self.ring_edge = None
        self.nr = self.nth = 0
        self.nlev = self.nlines = self.nspec = 0

    def close(self):
        if self.ctx:
            self.L.rl_destroy(self.ctx)
            self.ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc, what):
        if rc:
            msg = self.L.rl_last_error(self.ctx)
            raise RadliteError(rc, f"{what}: {msg.decode() if msg else ''}")

    # ---- model -----------------------------------------------------
    def set_grid(self, r, theta_half):
        r, rp = _d(r)
        t, tp = _d(theta_half)
        self.nr, self.nth = len(r), len(t)
        self._chk(self.L.rl_set_grid(self.ctx, len(r), rp, len(t), tp), "rl_set_grid")

    def set_medium(self, rho, abund, vel, lw_kms, umass_av, aksmax):
        a = [_d(x) for x in (rho, abund, vel, lw_kms)]
        self._chk(self.L.rl_set_medium(self.ctx, a[0][1], a[1][1], a[2][1], a[3][1],
                                       umass_av, aksmax), "rl_set_medium")

    def set_levels(self, pops):
        p, pp = _d(pops)
        self.nlev = p.shape[-1]
        self._chk(self.L.rl_set_levels(self.ctx, p.shape[-1], pp), "rl_set_levels")

    def set_lines(self, nu0, A, gu, gl, lev_up, lev_low):
        a = [_d(x) for x in (nu0, A, gu, gl)]
        u, up = _i(lev_up)
        l, lp = _i(lev_low)
        self.nlines = len(a[0][0])
        self._chk(self.L.rl_set_lines(self.ctx, len(a[0][0]), a[0][1], a[1][1], a[2][1],
                                      a[3][1], up, lp), "rl_set_lines")

    def set_dust_line(self, jd, ad):
        j, jp = _d(jd)
        a, ap = _d(ad)
        self._chk(self.L.rl_set_dust_line(self.ctx, jp, ap), "rl_set_dust_line")

    def set_star(self, rstar, cont_nu, istar):
        n, npt = _d(cont_nu)
        s, sp = _d(istar)
        self._chk(self.L.rl_set_star(self.ctx, rstar, len(n), npt, sp), "rl_set_star")

    def set_interstellar_field(self, cont_nu, isrf):
        nu, nup = _d(cont_nu)
        f, fp = _d(isrf)
        self._chk(self.L.rl_set_interstellar_field(self.ctx, len(nu), nup, fp),
                  "rl_set_interstellar_field")

    def set_dust_scatter(self, cont_nu, scat):
        """The scattering source [nfreq, nr, nth] for the device pre-pass
        (None removes it)."""
        if scat is None:
            self._chk(self.L.rl_set_dust_scatter(self.ctx, 0, None, None), "rl_set_dust_scatter")
            return
        nu, nup = _d(cont_nu)
        sc, scp = _d(scat)
        self._chk(self.L.rl_set_dust_scatter(self.ctx, len(nu), nup, scp), "rl_set_dust_scatter")

    def set_outer_bc(self, t):
        self._chk(self.L.rl_set_outer_bc(self.ctx, int(t)), "rl_set_outer_bc")

    def load_run(self, ri: l4io.RunInputs, tables: LineTables | None = None):
        """Load a RADLite run directory's contents (see l4io.read_run_dir)."""
        tb = prepare_tables(ri) if tables is None else tables
        self.tables = tb
        self.set_grid(ri.r, ri.theta_half)
        self.set_star(ri.rstar, tb.cont_nu, tb.istar)
        self.set_medium(ri.rho_gas, ri.abund, ri.vel, ri.lw_kms, ri.umass_av, tb.aksmax)
        self.set_levels(ri.pops)
        self.set_lines(tb.nu0, tb.A, tb.gu, tb.gl, tb.lev_up, tb.lev_low)
        self.set_dust_line(tb.jd, tb.ad)
        if getattr(ri, "isrf", None) is not None:
            self.set_interstellar_field(ri.cont_nu, ri.isrf)
        self.set_outer_bc(ri.outer_bc)
        return tb

    # ---- rays and trace --------------------------------------------
    def setup_rays(self, incl_rad, nphi, nrext, dbdr):
        inc, ip_ = _d(np.atleast_1d(incl_rad))
        n = len(inc)
        nrays = np.zeros(n, np.int32)
        nrings = np.zeros(n, np.int32)
        rc = self.L.rl_setup_rays_circular(self.ctx, n, ip_, nphi, nrext, dbdr,
                                           nrays.ctypes.data_as(_ip),
                                           nrings.ctypes.data_as(_ip), None, None)
        self._chk(rc, "rl_setup_rays_circular")
        rr = np.zeros(nrings[0] + 1)
        re = np.zeros(nrings[0] + 2)
        rc = self.L.rl_setup_rays_circular(self.ctx, n, ip_, nphi, nrext, dbdr, None, None,
                                           rr.ctypes.data_as(_dp), re.ctypes.data_as(_dp))
        self._chk(rc, "rl_setup_rays_circular")
        self.nrays, self.nrings = [int(x) for x in nrays], [int(x) for x in nrings]
        self.ring_r, self.ring_edge = rr, re
        return self.nrays

    def setup_rays_rectang(self, incl_rad, nx, ny, size_x, size_y, phioff=0.0, xoff=0.0,
                           yoff=0.0, addstar=0):
        """Rectangular-image (position-velocity) ray sets, one per
        inclination: nx x ny pixel rays over size_x x size_y cm, rotated by
        phioff and offset by (xoff, yoff) cm; addstar > 0 adds an
        unresolved star.  Replaces the circular ray sets."""
        inc, ip_ = _d(np.atleast_1d(incl_rad))
        n = len(inc)
        nrays = np.zeros(n, np.int32)
        rc = self.L.rl_setup_rays_rectang(self.ctx, n, ip_, int(nx), int(ny), float(size_x),
                                          float(size_y), float(phioff), float(xoff),
                                          float(yoff), int(addstar), nrays.ctypes.data_as(_ip))
        self._chk(rc, "rl_setup_rays_rectang")
        self.nrays, self.nrings = [int(x) for x in nrays], [0] * n
        self.pv_shape = (int(ny), int(nx))
        return self.nrays

    def trace_posvel(self, items, want_tau=True):
        """Position-velocity cubes of (line, incl, nfr, vmax_kms[, skip])
        items on rectangular ray sets.  Returns (list of intensity cubes
        [ny, nx, nfr], list of optical-depth cubes or None, warn bits)."""
        its = [tuple(it[:4]) + (SKIP_NONE,) for it in items]
        arr = (RLItem * len(its))(*[RLItem(*it) for it in its])
        ny, nx = self.pv_shape
        tot = sum(ny * nx * it[2] for it in its)
        img = np.zeros(tot)
        tau = np.zeros(tot) if want_tau else None
        warn = np.zeros(len(its), np.uint32)
        rc = self.L.rl_trace_lines_posvel(
            self.ctx, len(its), arr, img.ctypes.data_as(_dp),
            tau.ctypes.data_as(_dp) if tau is not None else None,
            warn.ctypes.data_as(C.POINTER(C.c_uint)))
        self._chk(rc, "rl_trace_lines_posvel")
        im, ta, o = [], [], 0
        for it in its:
            n = ny * nx * it[2]
            im.append(img[o:o + n].reshape(ny, nx, it[2]))
            if tau is not None:
                ta.append(tau[o:o + n].reshape(ny, nx, it[2]))
            o += n
        return im, (ta if tau is not None else None), warn

    def resample_image(self, image, vtx, w):
        """image [nrays, nfr] resampled with per-pixel vertices vtx [npix, nv]
        (-1: none) and weights w [npix, nv] on the device.  Returns
        [nfr, npix]."""
        img = np.ascontiguousarray(image, dtype=np.float64)
        vt = np.ascontiguousarray(vtx, dtype=np.int32)
        ww = np.ascontiguousarray(w, dtype=np.float64)
        nrays, nfr = img.shape
        npix, nv = vt.shape
        out = np.zeros((nfr, npix))
        self._chk(self.L.rl_resample_image(self.ctx, nrays, nfr, img.ctypes.data_as(_dp), npix, nv,
                                           vt.ctypes.data_as(_ip), ww.ctypes.data_as(_dp),
                                           out.ctypes.data_as(_dp)), "rl_resample_image")
        return out

    def trace(self, items, want_image=False, want_cmask=False):
        """items: list of (line, incl, nfr, vmax_kms, skip_mode).
        Returns (list of flux arrays, list of images or None, warn bits),
        and the list of integrated-cell masks [nrays, nfr] when want_cmask."""
        arr = (RLItem * len(items))(*[RLItem(*it) for it in items])
        nf = [it[2] for it in items]
        flux = np.zeros(sum(nf))
        nimg = sum(self.nrays[it[1]] * it[2] for it in items)
        img = np.zeros(nimg) if want_image else None
        cm = np.zeros(nimg, np.uint8) if want_cmask else None
        warn = np.zeros(len(items), np.uint32)
        rc = self.L.rl_trace_lines_forward_ex(
            self.ctx, len(items), arr, flux.ctypes.data_as(_dp),
            img.ctypes.data_as(_dp) if img is not None else None,
            cm.ctypes.data_as(C.POINTER(C.c_ubyte)) if cm is not None else None,
            warn.ctypes.data_as(C.POINTER(C.c_uint)))
        self._chk(rc, "rl_trace_lines_forward")
        fl, im, cms, o, io = [], [], [], 0, 0
        for it in items:
            fl.append(flux[o:o + it[2]])
            o += it[2]
            n = self.nrays[it[1]] * it[2]
            if img is not None:
                im.append(img[io:io + n].reshape(self.nrays[it[1]], it[2]))
            if cm is not None:
                cms.append(cm[io:io + n].reshape(self.nrays[it[1]], it[2]))
            io += n
        out = (fl, (im if img is not None else None), warn)
        return out + (cms,) if want_cmask else out

    def trace_noalloc(self, items, flux_dev, image_dev=None):
        """rl_trace_lines_forward_noalloc: the spectra (and images) of items
        written to device memory at flux_dev (and image_dev), addresses as
        ints (e.g. a torch tensor's data_ptr()), no device allocation by the
        library.  Returns the warn bits."""
        arr = (RLItem * len(items))(*[RLItem(*it) for it in items])
        warn = np.zeros(len(items), np.uint32)
        rc = self.L.rl_trace_lines_forward_noalloc(
            self.ctx, len(items), arr, C.c_void_p(int(flux_dev)),
            C.c_void_p(int(image_dev)) if image_dev else None,
            warn.ctypes.data_as(C.POINTER(C.c_uint)))
        self._chk(rc, "rl_trace_lines_forward_noalloc")
        return warn

    def trace_tape(self, items):
        """Forward trace of SKIP_NONE items that keeps the adjoint's tape
        (node records and I checkpoints) on the device; a following
        adjoint() on the same items with unchanged inputs runs only the
        reverse sweep.  Returns the list of flux arrays."""
        arr = (RLItem * len(items))(*[RLItem(*it) for it in items])
        nf = [it[2] for it in items]
        flux = np.zeros(sum(nf))
        self._chk(self.L.rl_trace_lines_forward_tape(self.ctx, len(items), arr,
                                                     flux.ctypes.data_as(_dp)),
                  "rl_trace_lines_forward_tape")
        fl, o = [], 0
        for n in nf:
            fl.append(flux[o:o + n])
            o += n
        return fl

    def tape_release(self):
        self._chk(self.L.rl_tape_release(self.ctx), "rl_tape_release")

    def adjoint_info(self):
        """(checkpoint interval of the last adjoint/tape call, 1 if the last
        adjoint ran from the tape, most ray chunks of an item)."""
        v = [C.c_int(0) for _ in range(3)]
        self._chk(self.L.rl_last_adjoint_info(self.ctx, *[C.byref(x) for x in v]),
                  "rl_last_adjoint_info")
        return tuple(x.value for x in v)

    def adjoint(self, items, dL_dflux, outputs=None):
        """Gradient of L with respect to the inputs the trace reads, given
        dL/dflux per item (list of arrays, same layout as trace's flux).
        Every item must use SKIP_NONE.  Returns (list of flux arrays,
        dict of gradients g_rho, g_abund, g_lw [nr, nth], g_pops
        [nr, nth, nlev], g_jd, g_ad [nr, nth, nlines]); ``outputs`` (names)
        limits the dict to those gradients."""
        arr = (RLItem * len(items))(*[RLItem(*it) for it in items])
        nf = [it[2] for it in items]
        g = np.ascontiguousarray(np.concatenate([np.asarray(x, float) for x in dL_dflux]))
        assert g.size == sum(nf)
        flux = np.zeros(sum(nf))
        nr, nth = self.nr, self.nth
        shapes = dict(g_rho=(nr, nth), g_abund=(nr, nth), g_lw=(nr, nth),
                      g_pops=(nr, nth, self.nlev), g_jd=(nr, nth, self.nlines),
                      g_ad=(nr, nth, self.nlines))
        out = {k: np.zeros(v) for k, v in shapes.items() if outputs is None or k in outputs}
        p = {k: (out[k].ctypes.data_as(_dp) if k in out else None) for k in shapes}
        rc = self.L.rl_trace_lines_adjoint(self.ctx, len(items), arr, g.ctypes.data_as(_dp),
                                           flux.ctypes.data_as(_dp), p["g_rho"], p["g_abund"],
                                           p["g_lw"], p["g_pops"], p["g_jd"], p["g_ad"])
        self._chk(rc, "rl_trace_lines_adjoint")
        fl, o = [], 0
        for n in nf:
            fl.append(flux[o:o + n])
            o += n
        return fl, out

    def memory_info(self):
        """dict: budget, resident trajectory bytes, work bytes of the last
        call's largest unit, free and total device bytes."""
        v = np.zeros(5)
        self._chk(self.L.rl_memory_info(self.ctx, v.ctypes.data_as(_dp)), "rl_memory_info")
        return dict(budget=v[0], traj=v[1], work=v[2], free=v[3], total=v[4])

    # ---- device pre-pass ---------------------------------------------
    def set_level_data(self, E_K, g):
        e, ep = _d(E_K)
        gg, gp = _d(g)
        self._chk(self.L.rl_set_level_data(self.ctx, len(e), ep, gp), "rl_set_level_data")
        self.nlev = len(e)

    def set_dust_species(self, rho_d, kabs_nu0, kscat_nu0):
        """rho_d [nspec, nr, nth]; kabs/kscat at the line centres [nspec, nlines]."""
        r, rp = _d(rho_d)
        a, ap = _d(kabs_nu0)
        b, bp = _d(kscat_nu0)
        self._chk(self.L.rl_set_dust_species(self.ctx, r.shape[0], rp, ap, bp),
                  "rl_set_dust_species")
        self.nspec = r.shape[0]

    def set_partition_table(self, T, Q):
        """Tabulated partition sum (T ascending); set_physics(q_over=Q_TABLE)
        interpolates it linearly, extrapolating past the ends.  T=None
        removes it."""
        if T is None:
            self._chk(self.L.rl_set_partition_table(self.ctx, 0, None, None),
                      "rl_set_partition_table")
            return
        t, tp = _d(T)
        q, qp = _d(Q)
        self._chk(self.L.rl_set_partition_table(self.ctx, len(t), tp, qp),
                  "rl_set_partition_table")

    def set_turbulence_alpha(self, alpha, gamma=1.4, mu=2.3):
        """pyradlite's alpha turbulence model (alpha < 0: explicit v_turb)."""
        self._chk(self.L.rl_set_turbulence_alpha(self.ctx, float(alpha), float(gamma), float(mu)),
                  "rl_set_turbulence_alpha")

    def set_velocity_kepler(self, mstar_g):
        """Keplerian velocity about a star of mass mstar [g], on the device."""
        self._chk(self.L.rl_set_velocity_kepler(self.ctx, float(mstar_g)),
                  "rl_set_velocity_kepler")

    def set_physics(self, T_gas, v_turb_kms, log_abund=None, T_dust=None, m_mol=28.0,
                    q_over=2.77):
        """q_over > 0: Q = T/q_over; 0: populations renormalised; Q_TABLE:
        the partition table.  v_turb_kms may be None under the alpha model."""
        t, tp = _d(T_gas)
        v, vp_ = (None, None) if v_turb_kms is None else _d(v_turb_kms)
        la, lap = (None, None) if log_abund is None else _d(log_abund)
        td, tdp = (None, None) if T_dust is None else _d(T_dust)
        self._chk(self.L.rl_set_physics(self.ctx, tp, tdp, lap, vp_, float(m_mol),
                                        float(q_over)), "rl_set_physics")

    def physics_adjoint(self, want_dust=False):
        """dict g_T, g_log_abund, g_v_turb [nr, nth] (and g_T_dust when the
        dust temperature was given) from the last adjoint call."""
        nr, nth = self.nr, self.nth
        out = dict(g_T=np.zeros((nr, nth)), g_log_abund=np.zeros((nr, nth)),
                   g_v_turb=np.zeros((nr, nth)))
        gd = np.zeros((self.nspec, nr, nth)) if want_dust else None
        self._chk(self.L.rl_physics_adjoint(
            self.ctx, out["g_T"].ctypes.data_as(_dp),
            gd.ctypes.data_as(_dp) if gd is not None else None,
            out["g_log_abund"].ctypes.data_as(_dp), out["g_v_turb"].ctypes.data_as(_dp)),
            "rl_physics_adjoint")
        if gd is not None:
            out["g_T_dust"] = gd
        return out

    def get_tables(self):
        """The device tables: pops [nr, nth, nlev], lw, abund [nr, nth], jd,
        ad [nr, nth, nlines]."""
        nr, nth = self.nr, self.nth
        t = dict(pops=np.zeros((nr, nth, self.nlev)), lw=np.zeros((nr, nth)),
                 abund=np.zeros((nr, nth)), jd=np.zeros((nr, nth, self.nlines)),
                 ad=np.zeros((nr, nth, self.nlines)))
        self._chk(self.L.rl_get_tables(self.ctx, *[t[k].ctypes.data_as(_dp) for k in
                                                   ("pops", "lw", "abund", "jd", "ad")]),
                  "rl_get_tables")
        return t

    def plan(self, items):
        """The memory plan of a call on ``items`` without running it: forward
        batches and the largest batch's bytes, the adjoint's checkpoint
        interval, chunks per item and chunk bytes, the budget, whether
        each item's tape fits, the most nodes of an inclination."""
        arr = (RLItem * len(items))(*[RLItem(*it) for it in items])
        v = np.zeros(8)
        self._chk(self.L.rl_plan(self.ctx, len(items), arr, v.ctypes.data_as(_dp)), "rl_plan")
        return dict(batches=int(v[0]), batch_bytes=v[1], ckpt_interval=int(v[2]),
                    adjoint_chunks=int(v[3]), chunk_bytes=v[4], budget=v[5],
                    tape_fits=bool(v[6]), max_nodes=int(v[7]))

    def set_memory_budget(self, nbytes):
        """Device bytes a batch of items may use (0: 60% of free memory)."""
        self._chk(self.L.rl_set_memory_budget(self.ctx, float(nbytes)), "rl_set_memory_budget")

    def set_option(self, option, value):
        self._chk(self.L.rl_set_option(self.ctx, int(option), int(value)), "rl_set_option")

    def batch_info(self):
        """(batches of the last forward call, tile mapping, persistent blocks)."""
        v = [C.c_int(0) for _ in range(3)]
        self._chk(self.L.rl_last_batch_info(self.ctx, *[C.byref(x) for x in v]),
                  "rl_last_batch_info")
        return tuple(x.value for x in v)

    def stats(self):
        t = np.zeros(5)
        n = np.zeros(4, np.int64)
        self.L.rl_last_stats(self.ctx, t.ctypes.data_as(_dp),
                             n.ctypes.data_as(C.POINTER(C.c_longlong)))
        return dict(t_traj_ms=t[0], t_nodes_ms=t[1], t_formal_ms=t[2], t_flux_ms=t[3],
                    t_total_ms=t[4], rays=int(n[0]), nodes=int(n[1]),
                    overshoot=int(n[2]), nominal=int(n[3]))


def run_directory(path, precision="fp64", skip_mode=SKIP_REFERENCE, device=0,
                  write_output=True, save_image=False):
    """The RADlite executable's job for a run directory: every requested
    line at the linespectrum.inp inclination, written as
    linespectrum_<mol>.dat (main.F:1037-1049, telescope.F:1662-1804).
    With save_image, also the circular image cube of each line,
    lineposvelcirc_<mol>_<line>.dat (the reference's SAVE_IMCIR build,
    telescope.F:1348-1373, 1595-1621), whose integrated-cell mask keeps
    the cells of the earlier lines as the reference's never-cleared
    imcir_cmask does."""
    ri = l4io.read_run_dir(path)
    if ri.extra.get("posvel"):
        return run_posvel_directory(path, ri, precision, device, write_output)
    rt = Raytracer(device=device, precision=precision)
    tb = rt.load_run(ri)
    nfr, vmax = passband(ri.vmax, ri.dv)
    incl = ri.incl_deg * 0.0174532925199                      # telescope.F:188
    rt.setup_rays([incl], ri.nphi, ri.b_extra, ri.dbdr)
    lines = list(range(ri.iline_start - 1, ri.iline_start - 1 + ri.nlines_run))
    items = [(il, 0, nfr, vmax, skip_mode) for il in lines]
    if save_image:
        fl, img, warn, cms = rt.trace(items, want_image=True, want_cmask=True)
        cum = np.zeros_like(cms[0])
        for k, il in enumerate(lines):
            cum |= cms[k]
            iline = il + 1
            if iline >= 100:          # the reference names files for lines 1-99 only
                continue
            nu0s = tb.nu0[il]
            passb = 3.33567e-6 * abs(nu0s) * vmax
            dnu = 2.0 * passb / (nfr - 1.0)
            velo = np.array([((-passb + ch * dnu) / nu0s) * 2.99792458e5 for ch in range(nfr)])
            l4io.write_image_cube(os.path.join(path, f"lineposvelcirc_{ri.molname}_{iline}.dat"),
                                  nu0s, ri.nphi, rt.nrings[0], rt.ring_edge,
                                  rt.ring_r, velo, img[k], cum)
    else:
        fl, _, warn = rt.trace(items)
    blocks = []
    for il, f in zip(lines, fl):
        nu0 = abs(tb.nu0[il])
        passb = 3.33567e-6 * nu0 * vmax
        dnu = 2.0 * passb / (nfr - 1.0)
        nu = nu0 + (-passb + np.arange(nfr) * dnu)
        blocks.append(dict(lev_up=int(ri.lev_up[il]), lev_low=int(ri.lev_low[il]), nu0=nu0,
                           nu=nu, flux=f))
    if write_output:
        l4io.write_line_spectrum_file(os.path.join(path, f"linespectrum_{ri.molname}.dat"),
                                      ri.molname, nfr, incl, blocks, vrad=ri.vrad)
    rt.close()
    return blocks, warn


def run_posvel_directory(path, ri=None, precision="fp64", device=0, write_output=True):
    """The RADlite executable's job for a run directory whose
    linespectrum.inp asks for position-velocity cubes (linespec_command 2,
    main.F:1062-1075): every requested line's cube on the rectangular
    image, written as lineposvel_<mol>_<line>.dat (calc_write_line_posvel,
    telescope.F:1828-2031).  Returns the list of (image, tau) cubes
    [ny, nx, nfr] and the warn bits."""
    ri = l4io.read_run_dir(path) if ri is None else ri
    pv = ri.extra["posvel"]
    rt = Raytracer(device=device, precision=precision)
    tb = rt.load_run(ri)
    nfr, vmax = passband(ri.vmax, ri.dv)
    incl = ri.incl_deg * 0.0174532925199                      # telescope.F:188
    rt.setup_rays_rectang([incl], pv["nx"], pv["ny"], pv["size_x"], pv["size_y"],
                          pv["phioff"], pv["xoff"], pv["yoff"], pv["addstar"])
    lines = list(range(ri.iline_start - 1, ri.iline_start - 1 + ri.nlines_run))
    img, tau, warn = rt.trace_posvel([(il, 0, nfr, vmax) for il in lines])
    rt.close()
    if write_output:
        pix = dict(nx=pv["nx"], ny=pv["ny"], spx=(pv["size_x"] * 0.5) / (pv["nx"] * 0.5),
                   spy=(pv["size_y"] * 0.5) / (pv["ny"] * 0.5), phioff=pv["phioff"],
                   xoff=pv["xoff"], yoff=pv["yoff"])
        for k, il in enumerate(lines):
            nu0s = tb.nu0[il]
            nu0 = abs(nu0s)
            passb = 3.33567e-6 * nu0 * vmax
            dnu = 2.0 * passb / (nfr - 1.0)
            nu = nu0 + (-passb + np.arange(nfr) * dnu)
            l4io.write_posvel_file(os.path.join(path, l4io.posvel_file_name(ri.molname, il + 1)),
                                   ri.molname, ri.lev_up[il], ri.lev_low[il], nu0s, nu, incl,
                                   pix, img[k], tau[k], vrad=ri.vrad)
    return list(zip(img, tau)), warn
