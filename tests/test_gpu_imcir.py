"""Circular image -> rectangular cube on the GPU (rl_resample_image,
radlite_amd.imcir).

* The gather kernel against the same weights evaluated on the CPU, on a
  random image (<= 1e-14).
* imcir_conv on the circular image cube the engine writes for the 3-line
  golden run (lineposvelcirc file): the cube against the CPU evaluation of
  its weights; the rectangular image's flux against the circular image's
  (the same disk, two quadratures: a few per cent at this grid); the
  options (resolved star, channel convolution, down-sampling, position-
  velocity collapse, spectro-astrometry, FITS output read back).
"""
import os
import sys
import tempfile

import numpy as np
import pytest

from cases import build_case
from radlite_amd import engine, imcir, l4io, spectrum

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def test_gather_kernel_matches_cpu():
    say("[gpu] resample kernel")
    rng = np.random.default_rng(0)
    nrays, nfr, npix = 1 + 40 * 16, 33, 5000
    img = rng.normal(size=(nrays, nfr))
    vtx = rng.integers(0, nrays, size=(npix, 3)).astype(np.int32)
    vtx[rng.random(npix) < 0.1] = -1
    w = rng.random((npix, 3))
    rt = engine.Raytracer(0, "fp64")
    out = rt.resample_image(img, vtx, w)
    ref = (w[None, :, :] * np.where(vtx[None] >= 0, img[np.maximum(vtx, 0)].transpose(2, 0, 1),
                                     0.0)).sum(axis=2)
    assert out.shape == (nfr, npix)
    assert np.abs(out - ref).max() <= 1e-14 * np.abs(ref).max()
    bad = vtx.copy()
    bad[0, 0] = nrays
    with pytest.raises(engine.RadliteError):
        rt.resample_image(img, bad, w)
    rt.close()


def test_imcir_conv_on_traced_cube():
    say("[gpu] imcir_conv on the t6 circular cube")
    d = tempfile.mkdtemp()
    build_case("t6", d)
    engine.run_directory(d, precision="fp64", save_image=True)
    ri = l4io.read_run_dir(d)
    fn = os.path.join(d, f"lineposvelcirc_{ri.molname}_1.dat")
    cube = l4io.read_image_cube(fn)
    rmax = float(cube["ri"][-1])
    xs = 2.1 * rmax / imcir.AU
    rt = engine.Raytracer(0, "fp64")
    res = imcir.imcir_conv(cube, rt=rt, xs=xs, ys=xs, nx=301, ny=301, posang=30.0,
                           pvit=True, specast=True)
    im = res["imrect"]
    nfr = int(cube["nfr"])
    assert im.shape == (nfr, 301, 301)
    # the cube against the CPU evaluation of its weights
    geo = imcir.geometry(cube["r"], cube["ri"][0], int(cube["nphi"]), int(cube["nr"]), xs, xs,
                         301, 301, 30.0)
    rows = np.concatenate([cube["center"][None, :],
                           np.transpose(cube["image"], (2, 1, 0)).reshape(-1, nfr)])
    ref = (geo["w"][:, :, None] * np.where(geo["vtx"][:, :, None] >= 0,
                                            rows[np.maximum(geo["vtx"], 0)], 0.0)).sum(axis=1)
    ref = ref.T.reshape(nfr, 301, 301)
    s = np.pi * res["rstar"] ** 2 / ((res["xgrid"][1] - res["xgrid"][0]) ** 2)
    ref[:, 150, 150] += cube["center"] * s
    assert np.abs(im - ref).max() <= 1e-13 * np.abs(ref).max()
    # flux: the triangulated image covers the polygon through the ring
    # points, the ring quadrature whole circles; with few angles they differ
    rel = np.abs(res["fl"] - res["fl_cir"]) / np.abs(res["fl_cir"]).max()
    say(f"    flux of the rectangular cube vs the circular image ({int(cube['nphi'])} angles): "
        f"max rel {rel.max():.2e}")
    assert np.all(np.isfinite(res["fl"]))
    assert np.allclose(res["spec2d"], im.sum(axis=1))
    assert np.all(np.isfinite(res["sa"]))
    # options
    f = os.path.join(d, "rect.fits")
    res2 = imcir.imcir_conv(fn, rt=rt, xs=xs, ys=xs, nx=200, ny=160, add_resolved_star=True,
                            specres=3.0, down_samp=4, saveit=f, distance=140.0)
    assert res2["imrect"].shape == (nfr, 40, 50)
    hdus = spectrum.read_fits(f)
    assert hdus[0][0]["NAXIS1"] == 50 and hdus[0][0]["NAXIS2"] == 40
    assert np.allclose(hdus[0][1], res2["imrect"])
    assert np.allclose(hdus[2][1]["fl_cir"], res2["fl_cir"])
    rt.close()


def test_constant_image_covers_the_polygon():
    """A constant image: the rectangular cube's flux is the constant times
    the area of the regular nphi-gon through the outermost ring's points
    (the triangulation's hull), to the pixel quadrature."""
    say("[gpu] imcir_conv flux of a constant image")
    nr, nphi, nfr = 20, 64, 3
    r = np.geomspace(0.2, 8.0, nr) * imcir.AU
    ri = np.concatenate([[0.15 * imcir.AU], np.sqrt(r[:-1] * r[1:]), [9.0 * imcir.AU]])
    cube = dict(nfr=nfr, nphi=nphi, nr=nr, ri=ri, r=r, velo=np.array([-1.0, 0.0, 1.0]),
                center=np.zeros(nfr), image=np.full((nfr, nphi, nr), 2.5), nu0=1e14)
    rt = engine.Raytracer(0, "fp64")
    res = imcir.imcir_conv(cube, rt=rt, xs=18.0, ys=18.0, nx=801, ny=801, specast=True)
    rt.close()
    rout = float(np.float32(r[-1]))
    area = 0.5 * nphi * rout ** 2 * np.sin(2 * np.pi / nphi)
    want = 2.5 * area / imcir.PC ** 2
    rel = np.abs(res["fl"] - want) / want
    say(f"    flux vs 2.5 x polygon area: {rel.max():.2e}")
    assert rel.max() <= 5e-3
    # symmetric image: no spectro-astrometric offset
    assert np.abs(res["sa"]).max() <= 1e-6 * rout
