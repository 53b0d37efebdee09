"""Circular image -> rectangular cube (radlite_amd.imcir, PRO/imcir_conv.pro)
on the CPU: the geometry (points, triangulation, per-pixel weights) pinned
by closed forms.

* A field linear in x, y, sampled at the image's points, is reproduced
  exactly by the weights at every pixel inside the triangulation (linear
  interpolation on any triangulation is exact for it); pixels outside get
  no vertex.
* The weights of a pixel inside sum to 1 (times the vertex factors).
* The ring-0 points sit at the origin, deduplicated to the first in IDL's
  memory order; the resolved star's points carry the limb-darkening factor
  1 - 0.3 (1 - cos asin(r/rstar)) of the central beam within rstar.
* The channel convolution (IDL CONVOL /EDGE_TRUNCATE) keeps a constant
  spectrum and moves a spike to the kernel's shape.
"""
import math

import numpy as np

from radlite_amd import imcir


def _rings(nr=12, nphi=16):
    r = np.geomspace(0.1, 5.0, nr) * imcir.AU
    ri = np.concatenate([[0.08 * imcir.AU], np.sqrt(r[:-1] * r[1:]), [6.0 * imcir.AU]])
    return r, ri, nphi, nr


def _linear_rows(geo, nr, nphi, nfr, coef):
    """Image rows whose ring points hold a + b x + c y at their positions."""
    rows = np.zeros((1 + nr * nphi, nfr))
    x, y = geo["x"], geo["y"]
    for ir in range(nr):
        for ip in range(nphi):
            for ch in range(nfr):
                a, b, c = coef[ch]
                rows[1 + ir * nphi + ip, ch] = a + b * float(x[ir, ip]) + c * float(y[ir, ip])
    return rows


def test_linear_field_is_reproduced():
    r, ri, nphi, nr = _rings()
    geo = imcir.geometry(r, ri[0], nphi, nr, 12.0, 10.0, 41, 37, posang_deg=23.0)
    coef = [(1.0, 2e-13, -3e-13), (5.0, 0.0, 1e-13), (-2.0, 1e-13, 1e-13)]
    rows = _linear_rows(geo, nr, nphi, 3, coef)
    vtx, w, inside = geo["vtx"], geo["w"], geo["inside"]
    assert inside.any() and (~inside).any()
    assert np.all(vtx[~inside] == -1)
    assert np.allclose(w[inside].sum(axis=1), 1.0, atol=1e-12)
    gx, gy = np.meshgrid(geo["xgrid"], geo["ygrid"])
    for ch, (a, b, c) in enumerate(coef):
        val = (w * rows[np.maximum(vtx, 0), ch]).sum(axis=1)
        exact = a + b * gx.ravel() + c * gy.ravel()
        err = np.abs(val - exact)[inside].max() / np.abs(exact).max()
        assert err <= 1e-9, (ch, err)        # single-precision point positions


def test_origin_points_and_resolved_star():
    r, ri, nphi, nr = _rings()
    geo = imcir.geometry(r, ri[0], nphi, nr, 0.5, 0.5, 21, 21, add_resolved_star=True)
    x, y, row, coef = geo["x"], geo["y"], geo["row"], geo["coef"]
    assert np.all(x[0] == 0) and np.all(y[0] == 0)          # ring 0 at the origin
    assert np.all(row[0] == 1 + np.arange(nphi))
    rstar = geo["rstar"]
    assert math.isclose(rstar, float(np.float32(ri[0])) * 0.99)
    for k in range(30):
        rad = rstar * k / 27.0
        if rad <= rstar:
            assert np.all(row[nr + k] == 0)
            want = 1.0 - 0.3 * (1.0 - math.cos(math.asin(min(rad / rstar, 1.0))))
            assert np.allclose(coef[nr + k], want)
        else:
            assert np.all(row[nr + k] == 1 + nphi + np.arange(nphi))
    # every weight refers to a row of the image
    assert geo["vtx"].max() < 1 + nr * nphi


def test_channel_convolution_edge_truncate():
    n = 21
    k = np.arange(n, dtype=float)
    kern = np.exp(-(k - (n - 1) / 2.0) ** 2 * 2.77 / 3.0 ** 2)
    kern /= kern.sum()
    const = np.full((2, n), 4.0)
    assert np.allclose(imcir._convol_truncate(const, kern), 4.0)
    spike = np.zeros((1, n))
    spike[0, 10] = 1.0
    out = imcir._convol_truncate(spike, kern)[0]
    # R[t] = sum_i a[t + i - n//2] k[i]: the spike at 10 gives k[10 - t + n//2]
    assert np.allclose(out, [kern[10 - t + n // 2] if 0 <= 10 - t + n // 2 < n else 0.0
                             for t in range(n)])
