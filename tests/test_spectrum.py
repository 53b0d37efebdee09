"""Spectrum synthesis after the trace (radlite_amd.spectrum, pyradlite's
RadliteSpectrum) on the CPU.

* The oracle (scipy/numpy restatement, oracle/spectrum_ref.py) pinned by
  closed forms: a flux linear in velocity has no emission and its
  continuum is carried exactly; a Gaussian line keeps its area and its
  width grows in quadrature with the kernel's; the kernel placement of
  scipy.ndimage.convolve for odd and even lengths.
* The torch synthesis against the oracle on random multi-line sets
  (uneven channel counts, overlapping boxes, lines with the most points
  in the middle) and on the reference's own spectrum files through the
  RadliteSpectrum interface (JSON input, duplicate runs, FITS output).
* Gradients of the synthesis with respect to the line fluxes (a linear
  map) against finite differences.
"""
import json
import os
import shutil

import numpy as np
import pytest
import torch
from scipy import ndimage

import spectrum_ref
from radlite_amd import spectrum

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _lines(seed, L=5, base=4.7e-4):
    """L random lines near 4.7 micron: channel counts 20-60 over +-vmax,
    a sloped continuum and a Gaussian emission bump each."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(L):
        nfr = int(rng.integers(20, 61))
        vmax = float(rng.uniform(15.0, 40.0))
        v = np.linspace(-vmax, vmax, nfr)
        lam = base * (1.0 + 0.004 * rng.uniform(-1, 1))          # cm
        freq = spectrum.C0 / lam
        cont = 1e-20 * (1.0 + 0.01 * rng.normal()) + 1e-23 * v * rng.normal()
        em = 5e-21 * rng.uniform(0.2, 1.0) * np.exp(-0.5 * (v / rng.uniform(2.0, 8.0)) ** 2)
        out.append(dict(vel=v, flux=cont + em, freq=freq, maxnumpoints=int(rng.integers(60, 70))))
    return out


def _torch(lines, **kw):
    out = spectrum.synthesize([d["vel"] for d in lines], [d["flux"] for d in lines],
                              [d["freq"] for d in lines],
                              maxnumpoints=[d["maxnumpoints"] for d in lines], **kw)
    return {k: v.detach().cpu().numpy() for k, v in out.items()}


def _rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


# ---- the oracle pinned ------------------------------------------------
def test_scipy_kernel_placement():
    """ndimage.convolve centres an n-point kernel on point n//2, which the
    torch path's left offset n - 1 - n//2 reproduces for the symmetric
    Gaussian."""
    for n in (3, 4, 5, 6, 9):
        w = np.arange(1.0, n + 1)
        x = np.zeros(21)
        x[10] = 1.0
        y = ndimage.convolve(x, w, mode="reflect")
        assert np.array_equal(np.nonzero(y)[0], np.arange(10 - n // 2, 10 - n // 2 + n))
        # reflect = half-sample symmetric: a constant stays constant
        assert np.allclose(ndimage.convolve(np.ones(7), np.ones(n) / n, mode="reflect"), 1.0)


def test_oracle_linear_flux_has_no_emission():
    v = np.linspace(-20.0, 20.0, 41)
    lines = [dict(vel=v, flux=2e-20 + 3e-23 * v, freq=spectrum.C0 / 4.7e-4, maxnumpoints=41)]
    o = spectrum_ref.process_spectrum(lines, obsres=3.0, vsampling=1.0, dist=140.0)
    assert np.abs(o["emission"]).max() <= 1e-12 * np.abs(o["continuum"]).max()
    # the continuum is the line-centre value (a single line: constant)
    c = 2e-20 * 1e23 / 140.0 ** 2
    assert np.allclose(o["continuum"], c, rtol=1e-14)
    assert np.allclose(o["spectrum"], c, rtol=1e-13)


def test_oracle_gaussian_area_and_width():
    """A Gaussian line (sigma s) on a flat continuum: the emission keeps
    the line's area, and its second moment is s^2 + sigma_k^2 with the
    kernel's sigma_k = obsres sqrt(ln 2) / 2 (the reference's kernel
    exp(-2 x^2 / (F^2 ln 2)), F = obsres in grid points)."""
    s, A, obsres = 4.0, 3e-21, 6.0
    v = np.linspace(-40.0, 40.0, 401)
    lam0 = 4.7e-4
    lines = [dict(vel=v, flux=1e-20 + A * np.exp(-0.5 * (v / s) ** 2), freq=spectrum.C0 / lam0,
                  maxnumpoints=401)]
    dist = 1.0
    for ref in (spectrum_ref.process_spectrum(lines, obsres=obsres, vsampling=0.2, dist=dist),
                _torch(lines, obsres=obsres, vsampling=0.2, dist=dist)):
        mu = ref["wavelength"]
        vel = (mu / (lam0 * 1e4) - 1.0) * spectrum.CINKM0
        e = ref["emission"]
        dv = np.gradient(vel)
        area = float(np.sum(e * dv))
        assert abs(area / (A * 1e23 * np.sqrt(2 * np.pi) * s) - 1.0) < 2e-3
        m1 = np.sum(e * vel * dv) / area
        m2 = np.sum(e * (vel - m1) ** 2 * dv) / area
        sk2 = (obsres * np.sqrt(np.log(2.0)) / 2.0) ** 2
        assert abs(m2 / (s * s + sk2) - 1.0) < 1e-2


# ---- torch against the oracle -----------------------------------------
def _oracle_first_fit(lines, choices, **kw):
    """The oracle at the first user sampling whose grid ends inside the
    model grid (the reference raises otherwise)."""
    for vs in choices:
        try:
            return vs, spectrum_ref.process_spectrum(lines, vsampling=vs, **kw)
        except ValueError:
            continue
    raise AssertionError("no sampling fits")


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_synthesis_matches_oracle(seed):
    lines = _lines(seed)
    kw = dict(obsres=[2.0, 5.0, 12.5][seed], dist=[140.0, 1.0, 55.0][seed])
    vs, o = _oracle_first_fit(lines, [0.5, 1.3, 0.25, 0.77, 2.1, 0.33, 0.9, 1.7, 0.61, 0.43,
                                      0.29, 1.1, 0.15, 2.6], **kw)
    kw["vsampling"] = vs
    t = _torch(lines, **kw)
    assert np.array_equal(t["wavelength"], o["wavelength"])
    for k in ("spectrum", "emission", "continuum"):
        assert _rel(t[k], o[k]) <= 1e-14, (k, _rel(t[k], o[k]))
    assert np.allclose(t["frequency"], o["frequency"], rtol=1e-15)


def test_beyond_the_grid_raises():
    """A user grid that outruns the model grid: interp1d's bounds error."""
    lines = _lines(0, L=1)
    kw = dict(obsres=2.0, dist=1.0)
    for vs in (0.37, 0.71, 1.9, 2.3, 3.1):
        try:
            spectrum_ref.process_spectrum(lines, vsampling=vs, **kw)
            ok = True
        except ValueError:
            ok = False
        if ok:
            _torch(lines, vsampling=vs, **kw)
        else:
            with pytest.raises(ValueError):
                _torch(lines, vsampling=vs, **kw)


def _run_dirs(tmp_path, names):
    paths = []
    for k, name in enumerate(names):
        d = tmp_path / f"run{k}"
        d.mkdir()
        shutil.copy(os.path.join(GOLD, f"spec_{name}.dat"), d / "linespectrum_moldata.dat")
        paths.append(str(d))
    return paths


def test_radlite_spectrum_interface(tmp_path):
    """RadliteSpectrum on the reference's own 3-line spectrum file, the run
    given twice (duplicates kept once): the same spectrum as the oracle on
    the parsed lines; FITS output read back."""
    paths = _run_dirs(tmp_path, ["t6", "t6"])
    for vs in (0.8, 0.65, 1.2, 0.45, 2.2, 0.95, 1.55):
        inp = dict(run_paths=dict(value=paths), dist=dict(value=120.0), obsres=dict(value=4.0),
                   vsampling=dict(value=vs), interpolation=dict(value="linear"),
                   verbose=dict(value=False), numcores=dict(value=1))
        jf = tmp_path / "spec.json"
        jf.write_text(json.dumps(inp))
        rs = spectrum.RadliteSpectrum(str(jf))
        try:
            rs.gen_spec()
            break
        except ValueError:        # user grid past the model grid, as in the reference
            continue
    lines = rs.get_attr("_radlitelineoutput")
    assert len(lines) == 3
    o = spectrum_ref.process_spectrum(lines, obsres=4.0, vsampling=vs, dist=120.0)
    for k in ("spectrum", "emission", "continuum", "wavelength"):
        assert _rel(rs.get_attr(k), o[k]) <= 1e-14, k
    assert rs.get_unit("spectrum") == "Jy"
    with pytest.raises(AttributeError):
        rs.get_attr("nonexistent")
    f = tmp_path / "out.fits"
    rs.write_fits(str(f))
    with pytest.raises(OSError):
        rs.write_fits(str(f))
    rs.write_fits(str(f), overwrite=True)
    assert os.path.getsize(f) % 2880 == 0
    hdus = spectrum.read_fits(str(f))
    assert len(hdus) == 3
    h0 = hdus[0][0]
    assert h0["SIMPLE"] is True and h0["FLUXUNIT"] == "Jy" and h0["INCL_DEG"] == 35.0
    assert h0["DIST_PC"] == 120.0
    tab = hdus[1][1]
    for k in ("wavelength", "frequency", "emission", "spectrum", "continuum"):
        assert np.array_equal(tab[k], rs.get_attr(k))
    assert hdus[2][1]["molname"] == ["moldata"] * 3


def test_gradients_against_finite_differences():
    lines = _lines(4, L=3)
    vs, _ = _oracle_first_fit(lines, [0.6, 0.45, 0.9, 1.1], obsres=3.0, dist=10.0)
    kw = dict(obsres=3.0, vsampling=vs, dist=10.0)
    fl = [torch.tensor(d["flux"], requires_grad=True) for d in lines]
    out = spectrum.synthesize([d["vel"] for d in lines], fl, [d["freq"] for d in lines], **kw)
    w = torch.as_tensor(np.random.default_rng(9).normal(size=out["spectrum"].shape))
    L = (w * out["spectrum"]).sum() + 0.3 * (w * out["emission"]).sum()
    L.backward()
    rng = np.random.default_rng(10)
    for _ in range(6):
        i = int(rng.integers(0, 3))
        k = int(rng.integers(0, len(lines[i]["flux"])))
        h = 1e-22
        vals = []
        for sgn in (1, -1):
            f2 = [torch.tensor(d["flux"]) for d in lines]
            f2[i][k] += sgn * h
            o = spectrum.synthesize([d["vel"] for d in lines], f2, [d["freq"] for d in lines],
                                    **kw)
            vals.append(float((w * o["spectrum"]).sum() + 0.3 * (w * o["emission"]).sum()))
        fd = (vals[0] - vals[1]) / (2 * h)
        g = float(fl[i].grad[k])
        assert abs(g - fd) <= 1e-6 * max(abs(fd), 1e-3 * float(torch.abs(fl[i].grad).max())), \
            (i, k, g, fd)


def test_line_velocities_are_the_files():
    """The channel velocities of a traced line, ascending, against the
    reference's spectrum file (E13.6) for the same run."""
    from cases import load_golden
    from radlite_amd import l4io
    g = load_golden("t6")
    blocks = l4io.read_line_spectrum_file(os.path.join(GOLD, "spec_t6.dat"))
    for il, b in enumerate(blocks):
        nfr = len(b["v"])
        vmax = -float(g["dnu"][il][0]) / (3.33567e-6 * abs(float(g["nu0"][il])))
        v = spectrum.line_velocities(float(g["nu0"][il]), nfr, vmax)
        assert np.allclose(v, b["v"], rtol=5.1e-6, atol=1e-12 * np.abs(v).max())
        # the flux in channel order, flipped, is the file's column
        assert np.allclose(g["flux"][il][::-1], b["flux"], rtol=5.1e-6)
