"""Circular image cube -> rectangular image cube (PRO/imcir_conv.pro).

RADLite's circular image (rings x angles per channel, written as
lineposvelcirc_<mol>_<line>.dat) is turned into a rectangular cube by the
IDL routine imcir_conv: the points of the image are triangulated and the
cube is linearly interpolated onto an nx x ny grid.  Here the geometry is
built once per call on the host and the interpolation of every channel
runs on the GPU (rl_resample_image):

* points (imcir_conv.pro:94-106): ring ir at radius r[ir] (the file's
  ring radii), angle phi_ip = 2 pi ip / nphi, rotated by -posang; ring 0
  is placed at the origin, as the routine does (its x, y are never set);
  coordinates are single precision (IDL fltarr);
* the resolved star (add_resolved_star, :116-132): 30 rings of radius
  rstar k / 27 (rstar = 0.99 times the first ring edge, :50), k = 0..29,
  unrotated; within rstar the central intensity times
  1 - 0.3 (1 - cos asin(r / rstar)), beyond it the value of ring 1;
* triangulation and linear interpolation on it (IDL triangulate and
  trigrid, :139-141): a Delaunay triangulation (Qhull through scipy);
  pixel centres x_i = -xs/2 + i xs/(nx-1); outside the triangulation 0;
  coincident points keep the first in IDL's memory order (angle fastest);
* the unresolved star (:145-148): pi rstar^2 / (dx dy) times the central
  intensity added to pixel (nx/2, ny/2);
* the spectral convolution (specres, :68-79): each ring spectrum (not the
  central one) convolved with exp(-(k - (nfr-1)/2)^2 2.77 / fwhm^2),
  fwhm = specres / dv channels, normalised, edges truncated (IDL CONVOL
  /EDGE_TRUNCATE: the end values repeated);
* the flux of the circular image (:81-92): pi rstar^2 I_c + sum over rings
  of (ri[ir+1]^2 - ri[ir]^2) pi / nphi times the ring's intensities, / pc^2;
* down_samp (:152-162): block averages by an integer factor (FREBIN of an
  integral reduction); pvit (:163-169): the image summed over y per
  channel; specast (:170-186): spectro-astrometric offset sum(I x) /
  sum(I) of the y-collapsed image and the flux sum(I) dx dy / pc^2.

Parity against IDL itself is unpinned: the ring points are co-circular in
fours, so the Delaunay triangulation is not unique and IDL's and Qhull's
choice of diagonal may differ (both interpolate linearly; they differ by
the curvature of the image across the cell).  The GPU interpolation is
checked against the same weights evaluated on the CPU and against closed
forms (a field linear in x, y is reproduced exactly).
"""
from __future__ import annotations

import math

import numpy as np

AU = 1.49597870e13            # cm (PRO/natconst.pro)
PC = 3.085678e18              # cm (PRO/natconst.pro)


def geometry(r, ri0, nphi, nr, xs_au, ys_au, nx, ny, posang_deg=0.0, add_resolved_star=False):
    """Per-pixel vertices and weights of the interpolation.

    r: ring radii [nr] (the file's r(1..nr)); ri0: the first ring edge.
    Returns dict vtx [npix, 3] (image rows: 0 the central beam, 1 + ir*nphi
    + ip ring ir, angle ip; -1 none), w [npix, 3], xgrid, ygrid (cm),
    rstar, and the points (x, y, row, coefficient)."""
    from scipy.spatial import Delaunay
    # the reader keeps radii and intensities in single precision (read_imcir.pro)
    r = np.asarray(r, np.float32).astype(np.float64)
    rstar = float(np.float32(ri0)) * 0.99
    posang = math.pi * posang_deg / 180.0
    phi = 2.0 * math.pi * np.arange(nphi, dtype=np.float32).astype(np.float64) / nphi
    nstar = 30 if add_resolved_star else 0
    nrow = nr + nstar
    x = np.zeros((nrow, nphi), np.float32)          # [ir][ip]: ip fastest, IDL's order
    y = np.zeros((nrow, nphi), np.float32)
    row = np.zeros((nrow, nphi), np.int64)
    coef = np.ones((nrow, nphi))
    for ir in range(nr):
        row[ir] = 1 + ir * nphi + np.arange(nphi)
        if ir >= 1:
            x[ir] = (r[ir] * np.cos(phi - posang)).astype(np.float32)
            y[ir] = (r[ir] * np.sin(phi - posang)).astype(np.float32)
    if nstar:
        star_rad = rstar * np.arange(nstar, dtype=np.float32).astype(np.float64) / (nstar - 2 - 1)
        for k in range(nstar):
            ir = nr + k
            x[ir] = (star_rad[k] * np.cos(phi)).astype(np.float32)
            y[ir] = (star_rad[k] * np.sin(phi)).astype(np.float32)
            if star_rad[k] <= rstar:
                P = math.asin(min(star_rad[k] / rstar, 1.0))
                row[ir] = 0
                coef[ir] = 1.0 - 0.3 * (1.0 - math.cos(P))
            else:
                row[ir] = 1 + 1 * nphi + np.arange(nphi)       # ring 1's value
    pts = np.stack([x.ravel(), y.ravel()], axis=1).astype(np.float64)
    # coincident points: the first in memory order
    _, first = np.unique(pts, axis=0, return_index=True)
    keep = np.sort(first)
    tri = Delaunay(pts[keep])
    xs, ys = xs_au * AU, ys_au * AU
    xgrid = -xs / 2.0 + np.arange(nx) * (xs / (nx - 1))
    ygrid = -ys / 2.0 + np.arange(ny) * (ys / (ny - 1))
    gx, gy = np.meshgrid(xgrid, ygrid)                # [iy][ix]: pixel ix + nx*iy
    q = np.stack([gx.ravel(), gy.ravel()], axis=1)
    simp = tri.find_simplex(q)
    inside = simp >= 0
    T = tri.transform[np.where(inside, simp, 0)]
    b = np.einsum("nij,nj->ni", T[:, :2], q - T[:, 2])
    bary = np.concatenate([b, 1.0 - b.sum(axis=1, keepdims=True)], axis=1)
    verts = keep[tri.simplices[np.where(inside, simp, 0)]]     # indices into pts
    vtx = row.ravel()[verts]
    w = bary * coef.ravel()[verts]
    vtx[~inside] = -1
    w[~inside] = 0.0
    return dict(vtx=vtx.astype(np.int32), w=w, xgrid=xgrid, ygrid=ygrid, rstar=rstar,
                x=x, y=y, row=row, coef=coef, inside=inside)


def _convol_truncate(a, kern):
    """IDL CONVOL(a, kern, /EDGE_TRUNCATE) along the last axis, scale 1:
    R[t] = sum_i a[clip(t + i - n//2)] kern[i]."""
    n, m = kern.size, a.shape[-1]
    out = np.zeros_like(a)
    for i in range(n):
        idx = np.clip(np.arange(m) + i - n // 2, 0, m - 1)
        out = out + a[..., idx] * kern[i]
    return out


def imcir_conv(cube, rt=None, xs=10.0, ys=10.0, nx=500, ny=500, posang=0.0,
               add_resolved_star=False, specres=None, down_samp=None, pvit=False,
               specast=False, saveit=None, int_unit="erg/s/cm^2/Hz/sterad", distance=125.0):
    """The rectangular cube of a circular image cube.

    cube: a lineposvelcirc file name, or a dict as l4io.read_image_cube
    returns (nfr, nphi, nr, ri, r, velo, center [nfr], image [nfr, nphi,
    nr], nu0).  rt: an engine.Raytracer whose device does the
    interpolation (one is opened on device 0 if None).  xs, ys in AU.
    Returns dict imrect [nfr, ny, nx] (erg/s/cm^2/Hz/ster), xgrid, ygrid
    (cm), velo, fl_cir [nfr], and spec2d [nfr, nx], sa, fl when pvit /
    specast are set.  saveit: a FITS file name (image HDU with the
    routine's header keywords, then the table of grids and spectra)."""
    from . import engine, l4io
    if isinstance(cube, str):
        cube = l4io.read_image_cube(cube)
    nfr, nphi, nr = int(cube["nfr"]), int(cube["nphi"]), int(cube["nr"])
    ri, r = np.asarray(cube["ri"], float), np.asarray(cube["r"], float)
    center = np.asarray(cube["center"], float)
    img = np.asarray(cube["image"], float)             # [nfr, nphi, nr]
    velo = np.asarray(cube["velo"], float)
    if specres:
        fwhm = specres / (velo[1] - velo[0])
        k = np.arange(nfr, dtype=np.float64)
        kern = np.exp(-(k - (nfr - 1) / 2.0) ** 2 * 2.77 / fwhm ** 2)
        kern = kern / kern.sum()
        img = np.moveaxis(_convol_truncate(np.moveaxis(img, 0, -1), kern), -1, 0)
    rstar = float(np.float32(ri[0])) * 0.99
    # flux of the circular image (imcir_conv.pro:81-92)
    fl_cir = math.pi * rstar ** 2 * center
    for ir in range(nr):
        surf = (ri[ir + 1] ** 2 - ri[ir] ** 2) * math.pi / nphi
        fl_cir = fl_cir + surf * img[:, :, ir].sum(axis=1)
    fl_cir = fl_cir / PC ** 2
    # image rows as the trace orders them: central beam, then ring-major
    rows = np.concatenate([center[None, :], np.transpose(img, (2, 1, 0)).reshape(nr * nphi, nfr)])
    geo = geometry(r, ri[0], nphi, nr, xs, ys, nx, ny, posang, add_resolved_star)
    own = rt is None
    if own:
        rt = engine.Raytracer(0, "fp64")
    try:
        out = rt.resample_image(rows, geo["vtx"], geo["w"])    # [nfr, npix]
    finally:
        if own:
            rt.close()
    imrect = out.reshape(nfr, ny, nx)
    xgrid, ygrid = geo["xgrid"], geo["ygrid"]
    if not add_resolved_star:                          # :145-148
        s = math.pi * rstar ** 2 / ((xgrid[1] - xgrid[0]) * (ygrid[1] - ygrid[0]))
        imrect[:, ny // 2, nx // 2] += center * s
    if down_samp:
        f = int(down_samp)
        if f != down_samp or f < 1:
            raise ValueError("down_samp: an integer reduction factor")
        nx2, ny2 = nx // f, ny // f
        imrect = imrect[:, :ny2 * f, :nx2 * f].reshape(nfr, ny2, f, nx2, f).mean(axis=(2, 4))
        xgrid = xgrid[:nx2 * f].reshape(nx2, f).mean(axis=1)
        ygrid = ygrid[:ny2 * f].reshape(ny2, f).mean(axis=1)
        nx, ny = nx2, ny2
    res = dict(imrect=imrect, xgrid=xgrid, ygrid=ygrid, velo=velo, fl_cir=fl_cir,
               nu0=cube.get("nu0"), rstar=rstar)
    res["spec2d"] = imrect.sum(axis=1) if pvit else np.zeros((nfr, nx))
    sa, fl = np.zeros(nfr), np.zeros(nfr)
    if specast:
        coll = imrect.sum(axis=1)                      # [nfr, nx]
        sa = (coll * xgrid[None, :]).sum(axis=1) / coll.sum(axis=1)
        fl = imrect.sum(axis=(1, 2)) * (xgrid[1] - xgrid[0]) * (ygrid[1] - ygrid[0]) / PC ** 2
    res["sa"], res["fl"] = sa, fl
    if saveit:
        _write_fits(saveit, res, int_unit, distance, add_resolved_star)
    return res


def _write_fits(fname, res, int_unit, distance, add_resolved_star):
    """The image (NAXIS1 = x, NAXIS2 = y, NAXIS3 = channel, big-endian
    float64) with the routine's keywords (imcir_conv.pro:188-198), then a
    table of the grids and spectra."""
    from .spectrum import _blocks, _card, _fits_bintable
    im = res["imrect"]
    nfr, ny, nx = im.shape
    conv = 1.0e17 if int_unit == "MJy/sterad" else 1.0
    xg, yg, v = res["xgrid"], res["ygrid"], res["velo"]
    nu0 = float(res["nu0"] or 0.0)
    cards = [("SIMPLE", True), ("BITPIX", -64), ("NAXIS", 3), ("NAXIS1", nx), ("NAXIS2", ny),
             ("NAXIS3", nfr), ("EXTEND", True), ("BUNIT", int_unit),
             ("CDELT1", float(abs(xg[1] - xg[0]) / AU / distance / 3600.0)),
             ("CDELT2", float(abs(yg[1] - yg[0]) / AU / distance / 3600.0)),
             ("CENTFREQ", nu0), ("NUDELT", float((v[1] - v[0]) * 1e5 / 2.99792458e10 * nu0)),
             ("VELODELT", float(v[1] - v[0])), ("DISTANCE", float(distance))]
    head = "".join(_card(k, val) for k, val in cards) + "END".ljust(80)
    data = (im * conv).astype(">f8").tobytes()
    cols = [("xgrid", list(xg)), ("ygrid", list(yg) + [0.0] * (len(xg) - len(yg)))] \
        if len(xg) >= len(yg) else [("xgrid", list(xg) + [0.0] * (len(yg) - len(xg))),
                                    ("ygrid", list(yg))]
    spec = [("velo", list(v)), ("sa", list(res["sa"])), ("fl", list(res["fl"])),
            ("fl_cir", list(res["fl_cir"]))]
    with open(fname, "wb") as f:
        f.write(_blocks(head.encode("ascii"), b" "))
        f.write(_blocks(data, b"\0"))
        f.write(_fits_bintable(cols))
        f.write(_fits_bintable(spec))
    return fname
