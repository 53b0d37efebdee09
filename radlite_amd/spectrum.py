"""Spectrum synthesis after the trace, in torch: pyradlite's RadliteSpectrum.

pyradlite turns the per-line spectra RADLite writes into one observed
spectrum (RadliteSpectrum, pyradlite/radlite.py:2229-3270): each line's
continuum is removed, the line emission is placed on a common wavelength
grid, the line-centre continua are interpolated across the grid, the sum
is convolved to the instrument's resolution and resampled.  Here the same
steps run on torch tensors, on the GPU when one is given, and are
differentiable in the line fluxes, so a fit to an observed spectrum can
back-propagate through the synthesis into the trace adjoint
(diff.SpectrumModel).

Steps (RadliteSpectrum._process_spectrum, radlite.py:3001-3184):

* box width = max(3 max_i |v_i[0]|, 3 obsres); model resolution vres =
  v[1] - v[0] of the line with the most points (:3031-3039);
* grids mu_min (1 + vres/c)^k and mu_min (1 + vsampling/c)^k [micron]
  over [mu_min, mu_max], the extreme line centres widened by the box
  (:3044-3067);
* continuum of a line: linear between its first and last channel; its
  emission is the excess; the continuum at v = 0 is kept (:3074-3084);
* the emission, extended by its edge values to +-box, linearly
  interpolated onto the grid points within the box and summed over lines
  (:3091-3112);
* the line-centre continua, sorted by wavelength and held constant to the
  grid ends, linearly interpolated over the grid (:3117-3130);
* Jy at ``dist`` pc: x 1e23 / dist^2 (:3134-3137);
* Gaussian kernel of n = ceil(3 obsres / vres) points,
  exp(-2 (k - (n-1)/2)^2 / ((obsres/vres)^2 ln 2)), convolved with
  half-sample-symmetric ("reflect") edges and normalised by its sum
  (:3143-3154; scipy.ndimage.convolve centres an n-point kernel on
  point n//2);
* linear resampling onto the user grid; frequency c / mu (:3161-3170).

Linear interpolation follows numpy.interp, which scipy's interp1d uses for
one-dimensional float64 data: the value at a knot is the knot's value,
elsewhere y_j + (y_{j+1} - y_j) / (x_{j+1} - x_j) (x - x_j); a point
beyond the grid raises ValueError as interp1d's bounds check does (the
final resampling meets this whenever the user grid's last point falls
past the model grid's, usual for vsampling finer than vres; the
reference stops there too).  Only
the 'linear' kind is provided (the reference's two-point continuum has no
other meaningful kind).

The line centre of a spectrum file block is its nu0 record (RADLite's
nu0 from the level energies, line.F:1980-1983); pyradlite takes c times
the wavenumber of its own molecule files instead (radlite.py:2830), which
is the same line.
"""
from __future__ import annotations

import glob
import json
import math
import os
import struct

import numpy as np
import torch

C0 = 2.99792458e10            # cm/s (astropy const.c.cgs.value)
CINMU0 = C0 * 1.0e4           # micron/s
CINKM0 = C0 / 1.0e5           # km/s


def _interp(x, y, xq):
    """numpy.interp of y(x) at xq inside [x[0], x[-1]] (x ascending, not
    differentiated; y differentiable).  Leading dimensions of x, y and xq
    are batch dimensions; x may be padded with +inf past ``n`` knots."""
    n = x.shape[-1]
    j = torch.searchsorted(x, xq, right=True) - 1
    j = j.clamp(0, n - 1)
    jn = (j + 1).clamp(max=n - 1)
    xj, xj1 = torch.gather(x, -1, j), torch.gather(x, -1, jn)
    yj, yj1 = torch.gather(y, -1, j), torch.gather(y, -1, jn)
    at = (xj == xq) | (jn == j) | torch.isinf(xj1)
    den = torch.where(at, torch.ones_like(xj), xj1 - xj)
    val = (yj1 - yj) / den * (xq - xj) + yj
    return torch.where(at, yj, val)


def _bounds(x_lo, x_hi, xq_lo, xq_hi):
    if xq_lo < x_lo:
        raise ValueError("A value in x_new is below the interpolation range.")
    if xq_hi > x_hi:
        raise ValueError("A value in x_new is above the interpolation range.")


def _geom(a, q, n):
    """a q^k, k = 0..n-1, with the scalar libm pow the reference's
    element-by-element loop calls (a vectorised power may differ in the
    last bit)."""
    return np.fromiter((a * math.pow(q, k) for k in range(n)), dtype=np.float64, count=n)


def _reflect_index(k, n):
    """Half-sample-symmetric extension (d c b a | a b c d | d c b a) of an
    n-point signal at integer positions k."""
    m = np.mod(k, 2 * n)
    return np.where(m >= n, 2 * n - 1 - m, m)


def synthesize(vel, flux, freq, dist, obsres, vsampling, maxnumpoints=None,
               interpolation="linear", device=None):
    """The observed spectrum from per-line spectra.

    vel: list of per-line velocities [km/s] (ascending, as the spectrum
    file lists them); flux: list of per-line fluxes [erg/s/cm^2/Hz at
    1 pc], numpy arrays or torch tensors (may require grad); freq: line
    centre frequencies [Hz]; dist [pc]; obsres and vsampling [km/s];
    maxnumpoints: per line, the channel maximum of its file (the line
    with the largest, the first among equals, sets vres; default: the
    first line).  Returns a dict of float64 tensors on ``device``:
    wavelength [micron], frequency [Hz], spectrum, emission, continuum
    [Jy]."""
    if interpolation not in ("linear",):
        raise NotImplementedError(f"interpolation kind {interpolation!r}: only 'linear'")
    dev = torch.device(device) if device is not None else None
    L = len(vel)
    if L == 0 or len(flux) != L or len(freq) != L:
        raise ValueError("vel, flux and freq need one entry per line")
    vel = [np.asarray(v, dtype=np.float64) for v in vel]
    freq = np.asarray(freq, dtype=np.float64)
    if dev is None:
        dev = flux[0].device if torch.is_tensor(flux[0]) else torch.device("cpu")
    fl = [f.to(dev, torch.float64) if torch.is_tensor(f) else
          torch.as_tensor(np.asarray(f, dtype=np.float64), device=dev) for f in flux]
    # scalars and grids (host, float64, the reference's arithmetic)
    maxvspan = max(abs(float(v[0])) for v in vel)
    boxwidth = max(3 * maxvspan, 3 * obsres)
    mnp = [1] * L if maxnumpoints is None else list(maxnumpoints)
    indmax = int(np.argmax(mnp))
    vres = float(vel[indmax][1] - vel[indmax][0])
    fnz = freq[freq != 0]
    muminraw = float(np.min(CINMU0 / 1.0 / fnz))
    mumin = muminraw - (boxwidth * 1.0E9 * muminraw / CINMU0)
    mumaxraw = float(np.max(CINMU0 / 1.0 / fnz))
    mumax = mumaxraw + (boxwidth * 1.0E9 * mumaxraw / CINMU0)
    growthfrac = 1 + (vres / 1.0 / CINKM0)
    maxlen = int(math.floor(math.log(mumax / 1.0 / mumin) / math.log(growthfrac)) + 1)
    fullmu_h = _geom(mumin, growthfrac, maxlen)
    userfrac = 1 + (vsampling / 1.0 / CINKM0)
    userlen = int(math.floor(math.log(mumax / 1.0 / mumin) / math.log(userfrac)) + 1)
    outmu_h = _geom(mumin, userfrac, userlen)
    fullmu = torch.as_tensor(fullmu_h, device=dev)
    outmu = torch.as_tensor(outmu_h, device=dev)

    # continuum per line and emission excess, padded [L, K] (K = nmax + 2
    # knots: the box edges and the channels)
    n = np.array([len(v) for v in vel])
    K = int(n.max()) + 2
    knots = np.full((L, K), np.inf)
    for i, v in enumerate(vel):
        kv = np.concatenate([[-1 * boxwidth], v, [boxwidth]])
        knots[i, :len(kv)] = (kv * 1.0E9 / freq[i]) + (CINMU0 / 1.0 / freq[i])
    v0 = torch.as_tensor(np.array([v[0] for v in vel]), device=dev)
    vN = torch.as_tensor(np.array([v[-1] for v in vel]), device=dev)
    f0 = torch.stack([f[0] for f in fl])
    fN = torch.stack([f[-1] for f in fl])
    slope = (fN - f0) / (vN - v0)
    contcen = slope * (0 - v0) + f0                      # value at v = 0
    em_rows = []
    for i in range(L):
        vv = torch.as_tensor(vel[i], device=dev)
        c = slope[i] * (vv - v0[i]) + f0[i]
        c = torch.where(vv == v0[i], f0[i], torch.where(vv == vN[i], fN[i], c))
        e = fl[i] - c
        pad = K - (len(vel[i]) + 2)
        em_rows.append(torch.cat([e[:1], e, e[-1:], e.new_zeros(pad)]))
    emolds = torch.stack(em_rows)
    kn = torch.as_tensor(knots, device=dev)

    # stitch: each line's box of grid points, batched over lines
    first = knots[:, 0]
    last = knots[np.arange(L), n + 1]
    lo = np.searchsorted(fullmu_h, first, side="left")
    hi = np.searchsorted(fullmu_h, last, side="right")
    cnt = hi - lo
    M = max(int(cnt.max()), 1)
    ar = np.arange(M)[None, :]
    idx_h = np.minimum(lo[:, None] + ar, maxlen - 1)
    mask_h = ar < cnt[:, None]
    idx = torch.as_tensor(idx_h, device=dev)
    xq = fullmu[idx]
    em_new = _interp(kn, emolds, xq)
    mask = torch.as_tensor(mask_h, device=dev)
    fullem = torch.zeros(maxlen, dtype=torch.float64, device=dev)
    fullem = fullem.index_add(0, idx[mask], em_new[mask])

    # continuum over the grid from the line-centre values
    mu0 = CINMU0 / 1.0 / freq
    order = np.argsort(mu0, kind="stable")
    cx = torch.as_tensor(np.concatenate([[mumin], mu0[order], [mumax]]), device=dev)
    cs = contcen[torch.as_tensor(order, device=dev)]
    cy = torch.cat([cs[:1], cs, cs[-1:]])
    fullcont = _interp(cx, cy, fullmu)

    scale = 1.0E23 / (dist ** 2)
    fullem = fullem * scale
    fullcont = fullcont * scale
    fully = fullcont + fullem

    # Gaussian convolution, reflect edges, kernel centred on point n//2
    numgauss = int(math.ceil(3.0 * obsres / vres))
    kk = np.arange(0, numgauss, dtype=np.float64)
    toppart = -1 * 2 * ((kk - ((numgauss - 1) / 2.0)) ** 2)
    botpart = ((obsres / 1.0 / vres) ** 2) * np.log(2.0)
    kern = np.exp(toppart / botpart)
    ksum = float(np.sum(kern))
    left = numgauss - 1 - numgauss // 2
    pidx = torch.as_tensor(_reflect_index(np.arange(maxlen + numgauss - 1) - left, maxlen),
                           device=dev)

    def convolve(x):
        xp = x[pidx]
        out = xp[0:maxlen] * float(kern[0])
        for m in range(1, numgauss):
            out = out + xp[m:m + maxlen] * float(kern[m])
        return out / 1.0 / ksum

    resem = convolve(fullem)
    resy = convolve(fully)
    _bounds(fullmu_h[0], fullmu_h[-1], outmu_h[0], outmu_h[-1])
    return dict(wavelength=outmu,
                frequency=C0 / (outmu * 1.0E-4),
                spectrum=_interp(fullmu, resy, outmu),
                emission=_interp(fullmu, resem, outmu),
                continuum=_interp(fullmu, fullcont, outmu))


# ----------------------------------------------------------------------
# the RadliteSpectrum interface
# ----------------------------------------------------------------------
class RadliteSpectrum:
    """pyradlite's RadliteSpectrum (radlite.py:2229-3270): initialised from
    a JSON file of {"name": {"value": ...}} entries (run_paths, dist,
    obsres, vsampling, interpolation, verbose, numcores); gen_spec() reads
    every linespectrum_*.dat of the run directories and synthesises the
    spectrum; get_attr / get_unit as the reference's; write_fits.  An
    optional ``device`` entry (or constructor argument) places the
    synthesis on a GPU."""

    def __init__(self, infilename, device=None):
        with open(infilename) as f:
            inputdict = json.load(f)
        self._attrdict = {k: v["value"] for k, v in inputdict.items()}
        self._attrdict["units"] = {}
        self._device = device if device is not None else self._attrdict.get("device")

    def get_attr(self, attrname):
        try:
            return self._attrdict[attrname]
        except KeyError:
            pass
        raise AttributeError(f"'{attrname}' doesn't seem to be a valid attribute.  Valid "
                             f"attributes (sorted) are:\n{sorted(self._attrdict)}.\n"
                             "Run the method gen_spec() (if you haven't yet) to "
                             "automatically populate more attributes.\n")

    def get_unit(self, attrname):
        try:
            return self._attrdict["units"][attrname]
        except KeyError:
            raise KeyError("No unit returned because this attribute is either invalid or "
                           "has not been populated yet.") from None

    def _set_attr(self, attrname, attrval, attrunit=""):
        self._attrdict[attrname] = attrval
        self._attrdict["units"][attrname] = attrunit

    def gen_spec(self):
        self._read_radliteoutput()
        self._process_spectrum()

    def _read_radliteoutput(self):
        """Every block of every linespectrum_*.dat in the run directories
        (radlite.py:2862-2993): duplicates (same molecule, levels and line
        centre) kept once, in the sorted order of their keys as
        numpy.unique gives them; all runs must share one inclination."""
        from . import l4io
        lines, mols = [], []
        for path in self.get_attr("run_paths"):
            for fn in sorted(glob.glob(os.path.join(path, "linespectrum_*.dat"))):
                with open(fn) as f:
                    rec = f.read().splitlines()
                molname = rec[2].strip()
                maxnumpoints = int(rec[5])
                incl = float(rec[6].split()[2])
                for b in l4io.read_line_spectrum_file(fn):
                    lines.append(dict(vel=b["v"], flux=b["flux"], freq=b["nu0"],
                                      maxnumpoints=maxnumpoints))
                    mols.append(dict(molname=molname, incl=incl, lev_up=b["lev_up"],
                                     lev_low=b["lev_low"], freq=b["nu0"],
                                     wavelength=C0 / 1.0 / b["nu0"]))
        if not lines:
            raise ValueError("no linespectrum_*.dat files in the run paths")
        keys = [m["molname"] + str(m["lev_up"]) + str(m["lev_low"]) + repr(m["freq"])
                for m in mols]
        uniq = np.unique(keys, return_index=True)[1]
        lines = [lines[u] for u in uniq]
        mols = [mols[u] for u in uniq]
        if len({m["incl"] for m in mols}) > 1:
            raise ValueError("The RADLite runs you specified don't all have the same "
                             "source inclination!")
        self._set_attr("_radlitelineoutput", lines)
        self._set_attr("molinfo", mols)

    def _process_spectrum(self):
        lines = self.get_attr("_radlitelineoutput")
        out = synthesize([d["vel"] for d in lines], [d["flux"] for d in lines],
                         [d["freq"] for d in lines], self.get_attr("dist"),
                         self.get_attr("obsres"), self.get_attr("vsampling"),
                         maxnumpoints=[d["maxnumpoints"] for d in lines],
                         interpolation=self._attrdict.get("interpolation", "linear"),
                         device=self._device)
        for name, unit in (("spectrum", "Jy"), ("emission", "Jy"), ("continuum", "Jy"),
                           ("wavelength", "$\\mu$m"), ("frequency", "Hz")):
            self._set_attr(name, out[name].detach().cpu().numpy(), unit)

    def write_fits(self, fitsname, overwrite=False):
        """The spectrum as a FITS file (radlite.py:3190-3268): a primary
        header with WAVEUNIT, FREQUNIT, FLUXUNIT, INCL_DEG, DIST_PC; a
        binary table of wavelength, frequency, emission, spectrum,
        continuum (float64); a binary table of the lines (molecule name,
        levels, frequency, wavelength, inclination)."""
        try:
            cols = [(k, self.get_attr(k)) for k in
                    ("wavelength", "frequency", "emission", "spectrum", "continuum")]
            mols = self.get_attr("molinfo")
        except AttributeError:
            raise AttributeError("Oh no! Looks like you haven't processed any RADLite "
                                 "output data yet. You can do so by running the gen_spec() "
                                 "method for this class.") from None
        if os.path.exists(fitsname) and not overwrite:
            raise OSError(f"File {fitsname!r} already exists.")
        hdr = [("SIMPLE", True), ("BITPIX", 8), ("NAXIS", 0), ("EXTEND", True),
               ("WAVEUNIT", "micron"), ("FREQUNIT", "Hz"), ("FLUXUNIT", "Jy"),
               ("INCL_DEG", float(mols[0]["incl"])), ("DIST_PC", float(self.get_attr("dist")))]
        mcols = [("molname", [m["molname"] for m in mols]),
                 ("lev_up", [float(m["lev_up"]) for m in mols]),
                 ("lev_low", [float(m["lev_low"]) for m in mols]),
                 ("freq", [m["freq"] for m in mols]),
                 ("wavelength", [m["wavelength"] for m in mols]),
                 ("incl", [m["incl"] for m in mols])]
        with open(fitsname, "wb") as f:
            f.write(_fits_header(hdr))
            f.write(_fits_bintable(cols))
            f.write(_fits_bintable(mcols))


# ----------------------------------------------------------------------
# FITS (the standard's 2880-byte blocks of 80-character cards)
# ----------------------------------------------------------------------
def _card(key, value):
    if isinstance(value, bool):
        v = ("T" if value else "F").rjust(20)
    elif isinstance(value, (int, np.integer)):
        v = str(int(value)).rjust(20)
    elif isinstance(value, float):
        v = repr(value).upper().rjust(20)
    else:
        v = ("'" + str(value).replace("'", "''").ljust(8) + "'").ljust(20)
    return (f"{key:<8s}= " + v).ljust(80)[:80]


def _blocks(b, fill):
    return b + fill * ((-len(b)) % 2880)


def _fits_header(cards):
    s = "".join(_card(k, v) for k, v in cards) + "END".ljust(80)
    return _blocks(s.encode("ascii"), b" ")


def _fits_bintable(cols):
    nrow = len(cols[0][1])
    forms, packs = [], []
    for name, data in cols:
        if isinstance(data[0], str):
            w = max(1, max(len(s) for s in data))
            forms.append(f"{w}A")
            packs.append((lambda s, w=w: s.encode("ascii").ljust(w, b" ")[:w]))
        else:
            forms.append("D")
            packs.append(lambda x: struct.pack(">d", float(x)))
    rowb = sum(int(fm[:-1]) if fm.endswith("A") else 8 for fm in forms)
    cards = [("XTENSION", "BINTABLE"), ("BITPIX", 8), ("NAXIS", 2), ("NAXIS1", rowb),
             ("NAXIS2", nrow), ("PCOUNT", 0), ("GCOUNT", 1), ("TFIELDS", len(cols))]
    for k, ((name, _), fm) in enumerate(zip(cols, forms), 1):
        cards += [(f"TTYPE{k}", name), (f"TFORM{k}", fm)]
    body = b"".join(b"".join(p(data[r]) for p, (_, data) in zip(packs, cols))
                    for r in range(nrow))
    return _fits_header(cards) + _blocks(body, b"\0")


def read_fits(fname):
    """The header cards and data of a FITS file this module (or
    radlite_amd.imcir) wrote: list of (header dict, {column: values} for a
    binary table, the array for an image, or None)."""
    with open(fname, "rb") as f:
        raw = f.read()
    out, pos = [], 0
    while pos < len(raw):
        hdr = {}
        while True:
            card = raw[pos:pos + 80].decode("ascii")
            pos += 80
            if card.startswith("END"):
                break
            if card[8:10] == "= ":
                v = card[10:].strip()
                if v.startswith("'"):
                    val = v[1:v.rfind("'")].rstrip()
                elif v in ("T", "F"):
                    val = v == "T"
                else:
                    val = float(v) if any(c in v for c in ".E") else int(v)
                hdr[card[:8].strip()] = val
        pos += (-pos) % 2880
        table = None
        naxis = int(hdr.get("NAXIS", 0))
        if hdr.get("XTENSION") != "BINTABLE" and naxis > 0:
            # an image: big-endian, NAXIS1 fastest
            shape = [int(hdr[f"NAXIS{k}"]) for k in range(naxis, 0, -1)]
            dt = {-64: ">f8", -32: ">f4", 8: "u1", 16: ">i2", 32: ">i4", 64: ">i8"}[int(hdr["BITPIX"])]
            nb = int(np.prod(shape)) * np.dtype(dt).itemsize
            table = np.frombuffer(raw[pos:pos + nb], dtype=dt).reshape(shape).astype(np.float64)
            pos += nb
            pos += (-pos) % 2880
        elif hdr.get("XTENSION") == "BINTABLE":
            nrow, rowb = hdr["NAXIS2"], hdr["NAXIS1"]
            table, off = {}, 0
            for k in range(1, hdr["TFIELDS"] + 1):
                fm, name = hdr[f"TFORM{k}"], hdr[f"TTYPE{k}"]
                if fm.endswith("A"):
                    w = int(fm[:-1])
                    table[name] = [raw[pos + r * rowb + off:pos + r * rowb + off + w]
                                   .decode("ascii").rstrip() for r in range(nrow)]
                else:
                    w = 8
                    table[name] = np.array([struct.unpack(">d", raw[pos + r * rowb + off:
                                                                    pos + r * rowb + off + 8])[0]
                                            for r in range(nrow)])
                off += w
            pos += nrow * rowb
            pos += (-pos) % 2880
        out.append((hdr, table))
    return out


def line_velocities(nu0, nfr, vmax_kms, vrad=0.0):
    """Velocities [km/s] of a traced line's channels as the spectrum file
    lists them (write_line_spectrum, telescope.F:1773-1778: channels nfr..1,
    v = -c (nu - nu0) / nu0 + v_rad with nu = |nu0| + dnu_k, line.F:462-469),
    ascending; the trace's flux in channel order is flipped to match."""
    nu0a = abs(nu0)
    passb = 3.33567e-6 * nu0a * vmax_kms
    dnu = 2.0 * passb / (nfr - 1.0)
    nu = nu0a + (-passb + np.arange(nfr) * dnu)
    v = -2.99792458e5 * (nu - nu0) / nu0 + vrad
    return v[::-1].copy()


def synthesize_items(flux, items, nu0, dist, obsres, vsampling, vrad=0.0, device=None):
    """synthesize() of a trace's output: ``flux`` the concatenated spectra
    of (line, incl, nfr, vmax_kms, ...) ``items`` (one inclination), as
    engine.Raytracer.trace or diff.SpectrumModel return them; nu0 the
    lines' rest frequencies [Hz] (indexed by the items' line)."""
    vel, fl, fr, o = [], [], [], 0
    for it in items:
        il, nfr, vmax = it[0], it[2], it[3]
        vel.append(line_velocities(nu0[il], nfr, vmax, vrad))
        f = flux[o:o + nfr]
        fl.append(torch.flip(f, [0]) if torch.is_tensor(f) else np.asarray(f)[::-1].copy())
        fr.append(abs(nu0[il]))
        o += nfr
    return synthesize(vel, fl, fr, dist, obsres, vsampling, device=device)
