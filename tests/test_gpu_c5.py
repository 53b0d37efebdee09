"""The large workload on the GPU: the 1024x512 disk (529,409 rays per
inclination, 300 channels over +-30 km/s, non-LTE populations of the
366-level ro-vibrational ladder), first line of the 1000-line set at 40
degrees.  At this size the forward builds its trajectories in batches
under the memory budget and the adjoint cuts the item into ray chunks
with per-chunk stencil transposes.

* FP32 against FP64: line-integrated flux within 1e-4 (the bound for the
  FP32 mode), every channel within 1e-4 of the peak.
* FP64 against the oracle on a sample of rays (every 2003rd): per-ray
  intensities within 1e-8, the per-ray bound of the golden tests.
* The chunked adjoint with respect to the populations: FP32 against FP64
  (relative norm over the cells with a gradient above 1e-3 of the
  largest), and the FP64 gradient against central differences of the
  FP64 forward (step 1e-4 relative) on the three most influential cells of
  the upper level.
"""
import sys

import numpy as np
import pytest

import pyorc
from radlite_amd import engine, synth

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(900)]

INCL = 40.0


def say(*a):
    print(*a, file=sys.stderr, flush=True)


_CACHE = {}


def _inputs():
    if "ri" not in _CACHE:
        ri, items, incl = synth.c5_inputs(0, keep_lines=2)
        n, v = items[0][2], items[0][3]
        _CACHE["ri"] = (ri, n, v)
    return _CACHE["ri"]


def _engine(precision):
    ri, n, v = _inputs()
    rt = engine.Raytracer(0, precision)
    rt.load_run(ri)
    rt.setup_rays([INCL * 0.0174532925199], ri.nphi, ri.b_extra, ri.dbdr)
    return rt, n, v


def _line_integral(f):
    return float(np.sum(f - 0.5 * (f[0] + f[-1])))


def test_c5_forward_fp32_fp64_and_oracle_rays():
    ri, n, v = _inputs()
    say(f"[gpu] c5 forward fp64, {n} channels at {INCL} deg")
    rt, n, v = _engine("fp64")
    assert rt.nrays == [529409]
    item = [(0, 0, n, v, engine.SKIP_NONE)]
    f64, im, _ = rt.trace(item, want_image=True)
    st = rt.stats()
    say(f"    fp64: {st['nodes']} nodes, {rt.batch_info()[0]} batch(es)")
    rt.close()
    say("[gpu] c5 forward fp32")
    rt, _, _ = _engine("fp32")
    f32, _, _ = rt.trace(item)
    rt.close()
    li64, li32 = _line_integral(f64[0]), _line_integral(f32[0])
    e_line = abs(li32 - li64) / abs(li64)
    e_ch = float(np.max(np.abs(f32[0] - f64[0])) / np.abs(f64[0]).max())
    say(f"    fp32 vs fp64: line integral {e_line:.2e}, channels (of the peak) {e_ch:.2e}")
    assert e_line <= 1e-4 and e_ch <= 1e-4
    orc = pyorc.Oracle(pyorc.Prepared(ri))
    o = orc.line_spectrum(0, incl_rad=INCL * 0.0174532925199, nfr=n, vmax=v, skip_mode=1,
                          ray_stride=2003, nthreads=16)
    rows = np.arange(0, o["nrays"], 2003)
    ref = o["image"][rows]
    got = im[0][rows]
    scale = np.abs(ref).max()
    rel = np.abs(got - ref) / np.maximum(np.abs(ref), 1e-12 * scale)
    say(f"    fp64 vs oracle on {len(rows)} rays: max rel {rel.max():.2e}")
    assert rel.max() <= 1e-8


def test_c5_chunked_adjoint_populations():
    ri, n, v = _inputs()
    rng = np.random.default_rng(5)
    w = rng.normal(size=n)
    item = [(0, 0, n, v, engine.SKIP_NONE)]
    say("[gpu] c5 adjoint fp64 (populations)")
    rt, _, _ = _engine("fp64")
    f0, g = rt.adjoint(item, [w], outputs=("g_pops",))
    K, tape, chunks = rt.adjoint_info()
    say(f"    fp64 adjoint: checkpoint interval {K}, {chunks} ray chunk(s)")
    g64 = g["g_pops"]
    lu = int(ri.lev_up[0]) - 1
    ll = int(ri.lev_low[0]) - 1
    # central differences of the FP64 forward on the most influential cells
    # of the upper level (|x dL/dx| largest)
    pops = ri.pops
    mag = np.abs(g64[..., lu] * pops[..., lu])
    cells = np.argsort(mag.ravel())[::-1][:3]
    L0 = float(np.dot(w, f0[0]))
    scale = float(np.sum(np.abs(w * f0[0])))
    worst = 0.0
    for flat in cells:
        c = np.unravel_index(flat, mag.shape)
        x0 = pops[c + (lu,)]
        dx = 1e-4 * x0
        Ls = []
        for sgn in (1, -1):
            p = pops.copy()
            p[c + (lu,)] = x0 + sgn * dx
            rt.set_levels(p)
            fl, _, _ = rt.trace(item)
            Ls.append(float(np.dot(w, fl[0])))
        fd = (Ls[0] - Ls[1]) / (2 * dx)
        ga = float(g64[c + (lu,)])
        floor = 1e-13 * scale / dx
        err = abs(ga - fd) / abs(fd)
        worst = max(worst, err)
        say(f"    cell {c}: adjoint {ga:.6e}, central difference {fd:.6e}, rel {err:.2e}")
        assert abs(ga - fd) <= 1e-5 * abs(fd) + floor, (c, ga, fd)
    rt.set_levels(pops)
    rt.close()
    say("[gpu] c5 adjoint fp32")
    rt, _, _ = _engine("fp32")
    _, g = rt.adjoint(item, [w], outputs=("g_pops",))
    K32, _, chunks32 = rt.adjoint_info()
    rt.close()
    g32 = g["g_pops"]
    for lev in (lu, ll):
        a, b = g32[..., lev], g64[..., lev]
        big = np.abs(b) > 1e-3 * np.abs(b).max()
        rel = float(np.linalg.norm((a - b)[big]) / np.linalg.norm(b[big]))
        say(f"    fp32 vs fp64 dL/dpop level {lev + 1}: norm rel over {int(big.sum())} cells "
            f"{rel:.2e} (fp32: K {K32}, {chunks32} chunk(s))")
        assert rel <= 1e-3, (lev, rel)
    assert abs(L0) > 0
