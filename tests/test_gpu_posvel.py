"""Position-velocity cubes on the rectangular imager (linespec_command 2).

* FP64 cubes (intensity and line-of-sight optical depth of every pixel at
  every channel) against the reference's own cubes for the three golden
  runs: corner pixels outside the grid, rays within R*, the added
  unresolved star, the CMB outside the disk, rotated and offset images,
  several lines (<= 1e-10 relative).
* The run-directory driver writes the reference's lineposvel files byte
  for byte.
* FP32 against FP64 on the golden runs and on a 48x40 cube of the 200x100
  disk; FP64 against the oracle on that cube.
"""
import sys
import tempfile

import numpy as np
import pytest

import pyorc
from cases import POSVEL_CASES, build_case, input_hash, load_golden
from radlite_amd import engine, l4io, synth

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def _case(name):
    g = load_golden(name)
    d = tempfile.mkdtemp()
    build_case(name, d)
    assert str(g["input_sha256"]) == input_hash(d)
    return g, d, l4io.read_run_dir(d)


def _cubes(ri, precision, lines=None):
    pv = ri.extra["posvel"]
    nfr, vmax = engine.passband(ri.vmax, ri.dv)
    rt = engine.Raytracer(0, precision)
    rt.load_run(ri)
    n = rt.setup_rays_rectang([ri.incl_deg * 0.0174532925199], pv["nx"], pv["ny"],
                              pv["size_x"], pv["size_y"], pv["phioff"], pv["xoff"], pv["yoff"],
                              pv["addstar"])
    assert n == [pv["nx"] * pv["ny"] + (1 if pv["addstar"] > 0 else 0)]
    lines = range(ri.nlines_run) if lines is None else lines
    img, tau, warn = rt.trace_posvel([(il, 0, nfr, vmax) for il in lines])
    st = rt.stats()
    rt.close()
    return img, tau, warn, st


def _rel(a, r):
    return float((np.abs(a - r) / np.maximum(np.abs(r), 1e-12 * np.abs(r).max())).max())


@pytest.mark.parametrize("name", sorted(POSVEL_CASES))
def test_fp64_cube_matches_reference(name):
    say(f"[gpu] posvel {name} fp64 against the reference")
    g, d, ri = _case(name)
    img, tau, warn, st = _cubes(ri, "fp64")
    for il in range(len(g["nu0"])):
        ri_ = _rel(img[il], g["image"][il])
        t = g["tau"][il]
        assert np.array_equal(t == 0, tau[il] == 0)
        rt_ = float((np.abs(tau[il] - t) / np.maximum(np.abs(t), 1e-300)).max())
        say(f"    line {il}: intensity {ri_:.2e}, tau {rt_:.2e}, warn {warn[il]}")
        assert ri_ <= 1e-10 and rt_ <= 1e-10, (il, ri_, rt_)
        assert bool(warn[il] & 2) == (int(g["overshoot"][il]) > 0)
        assert bool(warn[il] & 1) == bool(g["maser"][il])


@pytest.mark.parametrize("name", sorted(POSVEL_CASES))
def test_run_directory_writes_reference_files(name):
    say(f"[gpu] posvel {name}: run directory")
    g, d, ri = _case(name)
    cubes, warn = engine.run_directory(d, precision="fp64")
    assert len(cubes) == len(g["nu0"])
    import os
    gold = os.path.join(os.path.dirname(__file__), "golden")
    for il in range(len(g["nu0"])):
        iline = ri.iline_start + il
        fn = l4io.posvel_file_name(ri.molname, iline)
        got = open(os.path.join(d, fn)).read()
        ref = open(os.path.join(gold, f"posvel_{name}_{ri.molname}_{iline}.dat")).read()
        assert got == ref, (name, iline)


@pytest.mark.parametrize("name", sorted(POSVEL_CASES))
def test_fp32_cube_against_fp64(name):
    say(f"[gpu] posvel {name} fp32 against fp64")
    g, d, ri = _case(name)
    a, ta, _, _ = _cubes(ri, "fp32")
    b, tb, _, _ = _cubes(ri, "fp64")
    for il in range(len(a)):
        # per pixel relative to the cube's brightest pixel, and the
        # channel totals (the cube's spectrum)
        e_pix = float(np.abs(a[il] - b[il]).max() / np.abs(b[il]).max())
        sa, sb = a[il].sum(axis=(0, 1)), b[il].sum(axis=(0, 1))
        e_tot = float((np.abs(sa - sb) / np.abs(sb)).max())
        e_tau = float(np.abs(ta[il] - tb[il]).max() / max(np.abs(tb[il]).max(), 1e-300))
        say(f"    line {il}: pixel {e_pix:.2e}, channel totals {e_tot:.2e}, tau {e_tau:.2e}")
        assert e_pix <= 1e-4 and e_tot <= 1e-4 and e_tau <= 1e-4


def test_large_cube_fp32_fp64_and_oracle():
    """48x40 pixels over the 200x100 disk, 32 channels, rotated, with the
    star added: FP32 against FP64 and FP64 against the oracle."""
    say("[gpu] posvel 48x40 cube on the 200x100 disk")
    ri, _ = synth.workload_inputs("c4", seed=0)
    pv = dict(nx=48, ny=40, size_x=2.4 * ri.r[-1], size_y=1.6 * ri.r[-1], phioff=0.4,
              xoff=0.0, yoff=0.0, addstar=1)
    ri.extra["posvel"] = pv
    nfr, vmax = 32, 12.0
    incl = 50.0 * 0.0174532925199
    out = {}
    for prec in ("fp64", "fp32"):
        rt = engine.Raytracer(0, prec)
        rt.load_run(ri)
        rt.setup_rays_rectang([incl], pv["nx"], pv["ny"], pv["size_x"], pv["size_y"],
                              pv["phioff"], pv["xoff"], pv["yoff"], pv["addstar"])
        img, tau, warn = rt.trace_posvel([(0, 0, nfr, vmax)])
        st = rt.stats()
        say(f"    {prec}: {st['t_total_ms']:.2f} ms, {st['rays']} rays, {st['nodes']} nodes")
        out[prec] = (img[0], tau[0])
        rt.close()
    a, b = out["fp32"][0], out["fp64"][0]
    e_pix = float(np.abs(a - b).max() / np.abs(b).max())
    say(f"    fp32 vs fp64: pixel {e_pix:.2e}")
    assert e_pix <= 1e-4
    orc = pyorc.Oracle(pyorc.Prepared(ri))
    o = orc.posvel(0, pv, incl_rad=incl, nfr=nfr, vmax=vmax)
    r = _rel(out["fp64"][0], o["image"])
    t = o["tau"]
    rt_ = float((np.abs(out["fp64"][1] - t) / np.maximum(np.abs(t), 1e-300)).max())
    say(f"    fp64 vs oracle: intensity {r:.2e}, tau {rt_:.2e}")
    assert r <= 1e-10 and rt_ <= 1e-10
    # the outside pixels are black and have no optical depth
    assert np.all(out["fp64"][0][t == 0] == 0.0)
