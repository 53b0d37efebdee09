"""The HIP engine on case r1, the reference's own tutorial RADMC inputs
(120 x 100 grid, 130-point continuum, 3 lines x 40 channels, 4,193 rays):
FP64 against the reference's golden spectra (<= 1e-10 relative per
channel) and per-ray intensities (<= 1e-8, the per-ray bound of the other
golden tests), with the reference's NONREDUNDANT skipping; FP32 line
integrals within 1e-4 of the reference."""
import sys

import numpy as np
import pytest

import radmc_case
from cases import load_golden
from radlite_amd import engine

pytestmark = pytest.mark.gpu


def _engine(precision):
    g = load_golden("r1")
    ri = radmc_case.radmc_inputs(g)
    rt = engine.Raytracer(0, precision)
    rt.load_run(ri)
    nfr, vmax = engine.passband(ri.vmax, ri.dv)
    rt.setup_rays([ri.incl_deg * 0.0174532925199], ri.nphi, ri.b_extra, ri.dbdr)
    return g, rt, nfr, vmax


def test_fp64_matches_reference_on_tutorial_inputs():
    print("[gpu] r1 fp64", file=sys.stderr, flush=True)
    g, rt, nfr, vmax = _engine("fp64")
    nl = len(g["flux"])
    fl, im, warn = rt.trace([(il, 0, nfr, vmax, engine.SKIP_REFERENCE) for il in range(nl)],
                            want_image=True)
    for il in range(nl):
        rel = np.abs(fl[il] - g["flux"][il]) / np.abs(g["flux"][il])
        assert rel.max() <= 1e-10, (il, rel.max())
        img = im[il][g["image_rows"]]
        scale = np.abs(g["image"][il]).max()
        rel = np.abs(img - g["image"][il]) / np.maximum(np.abs(g["image"][il]), 1e-12 * scale)
        assert rel.max() <= 1e-8, (il, rel.max())
        assert bool(warn[il] & engine.WARN_MASER) == bool(g["maser"][il])
    rt.close()


def test_fp32_line_integral_on_tutorial_inputs():
    print("[gpu] r1 fp32", file=sys.stderr, flush=True)
    g, rt, nfr, vmax = _engine("fp32")
    nl = len(g["flux"])
    fl, _, _ = rt.trace([(il, 0, nfr, vmax, engine.SKIP_REFERENCE) for il in range(nl)])
    for il in range(nl):
        ref = g["flux"][il]
        cont = 0.5 * (ref[0] + ref[-1])
        li_ref = float(np.sum(ref - cont))
        f = fl[il]
        li = float(np.sum(f - 0.5 * (f[0] + f[-1])))
        err = abs(li - li_ref) / abs(li_ref)
        print(f"    line {il}: FP32 line-integral error {err:.2e}", file=sys.stderr, flush=True)
        assert err <= 1e-4, (il, err)
    rt.close()
