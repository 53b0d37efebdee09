"""CPU restatement of pyradlite's spectrum synthesis (numpy + scipy).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of
radlite_amd.spectrum; never by the product path.

Follows RadliteSpectrum._process_spectrum (pyradlite/radlite.py:3001-3184)
step by step with the same library calls the reference makes
(scipy.interpolate.interp1d, scipy.ndimage.convolve), the same loop order
and the same scalar arithmetic:

* box width 3 max(|v_first|) or 3 obsres, finest resolution from the line
  with the most points (:3031-3039);
* wavelength grids mu_min (1 + v/c)^k at the model and at the user
  resolution (:3044-3067);
* continuum per line linear between the first and last channel, the
  emission its excess, the continuum at v = 0 kept (:3074-3084);
* each line's emission, extended by its edge values to +-box, linearly
  interpolated onto the grid points inside its box and added (:3091-3112);
* the line-centre continua, sorted by wavelength and extended to the grid
  ends, interpolated over the grid (:3117-3130);
* Jy at dist (:3134-3137); a Gaussian kernel of ceil(3 obsres / vres)
  points, exp(-2 (k - (n-1)/2)^2 / ((obsres/vres)^2 ln 2)), convolved with
  ndimage "reflect" edges and normalised by its sum (:3143-3154);
* linear resampling onto the user grid, frequency c / mu (:3161-3170).

Only the 'linear' interpolation kind is restated (the two-point continuum
of the reference has no other meaningful kind).
"""
from __future__ import annotations

import numpy as np
from scipy import ndimage
from scipy.interpolate import interp1d

C0 = 2.99792458e10          # astropy const.c.cgs.value
CINMU0 = C0 * 1.0e4
CINKM0 = C0 / 1.0e5


def process_spectrum(lines, obsres, vsampling, dist, interpkind="linear"):
    """lines: list of dicts vel (km/s, ascending), flux (erg/s/cm^2/Hz at
    1 pc), freq (Hz), maxnumpoints.  Returns dict of numpy arrays."""
    maxvspan = np.max([np.abs(d["vel"][0]) for d in lines])
    boxwidth = np.max([(3 * maxvspan), (3 * obsres)])
    indmax = np.argmax([d["maxnumpoints"] for d in lines])
    vres = lines[indmax]["vel"][1] - lines[indmax]["vel"][0]

    freq_list = np.array([d["freq"] for d in lines])
    freq_list = freq_list[freq_list != 0]
    muminraw = np.min(CINMU0 / 1.0 / freq_list)
    mumin = muminraw - (boxwidth * 1.0E9 * muminraw / CINMU0)
    mumaxraw = np.max(CINMU0 / 1.0 / freq_list)
    mumax = mumaxraw + (boxwidth * 1.0E9 * mumaxraw / CINMU0)

    growthfrac = (1 + (vres / 1.0 / CINKM0))
    maxlen = int(np.floor(np.log(mumax / 1.0 / mumin) / np.log(growthfrac)) + 1)
    fullmu = np.array([(mumin * (growthfrac ** ai)) for ai in range(0, maxlen)])
    fullem = np.zeros(maxlen)

    userfrac = (1 + (vsampling / 1.0 / CINKM0))
    userlen = int(np.floor(np.log(mumax / 1.0 / mumin) / np.log(userfrac)) + 1)
    outmu = np.array([(mumin * (userfrac ** ai)) for ai in range(0, userlen)])

    emonly, contcen = [None] * len(lines), [None] * len(lines)
    for ai, d in enumerate(lines):
        f = interp1d(x=[d["vel"][0], d["vel"][-1]], y=[d["flux"][0], d["flux"][-1]],
                     kind=interpkind)
        emonly[ai] = d["flux"] - f(d["vel"])
        contcen[ai] = f(0)

    for ai, d in enumerate(lines):
        muolds = ((np.concatenate([[-1 * boxwidth], d["vel"], [boxwidth]]) * 1.0E9 / d["freq"])
                  + (CINMU0 / 1.0 / d["freq"]))
        emolds = np.concatenate([[emonly[ai][0]], emonly[ai], [emonly[ai][-1]]])
        newinds = np.where(((fullmu <= muolds[-1]) & (fullmu >= muolds[0])))[0]
        f = interp1d(x=muolds, y=emolds, kind=interpkind)
        fullem[newinds] = fullem[newinds] + f(fullmu[newinds])

    order = np.argsort([(CINMU0 / 1.0 / d["freq"]) for d in lines])
    musort = np.array([(CINMU0 / 1.0 / d["freq"]) for d in lines])[order]
    csort = np.array(contcen)[order]
    musort = np.concatenate([[mumin], musort, [mumax]])
    csort = np.concatenate([[csort[0]], csort, [csort[-1]]])
    fullcont = interp1d(x=musort, y=csort, kind=interpkind)(fullmu)

    fullem = fullem * 1.0E23 / (dist ** 2)
    fullcont = fullcont * 1.0E23 / (dist ** 2)
    fully = fullcont + fullem

    numgauss = np.ceil(3.0 * obsres / vres)
    toppart = -1 * 2 * ((np.arange(0, numgauss) - ((numgauss - 1) / 2.0)) ** 2)
    botpart = ((obsres / 1.0 / vres) ** 2) * np.log(2.0)
    kern = np.exp(toppart / botpart)
    resem = ndimage.convolve(fullem, kern, mode="reflect") / 1.0 / np.sum(kern)
    resy = ndimage.convolve(fully, kern, mode="reflect") / 1.0 / np.sum(kern)

    outem = interp1d(x=fullmu, y=resem)(outmu)
    outy = interp1d(x=fullmu, y=resy)(outmu)
    outcont = interp1d(x=fullmu, y=fullcont)(outmu)
    return dict(spectrum=outy, emission=outem, continuum=outcont, wavelength=outmu,
                frequency=C0 / (outmu * 1.0E-4), fullmu=fullmu, kernel=kern)
