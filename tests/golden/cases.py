"""Golden parity cases: seeded models rendered by the compiled reference.

Each case is a run directory produced by radlite_amd.synth from a seed;
make_golden.py runs the reference build (oracle/_ref/RADlite_probe) on
it and stores the reference's full-precision results in <name>.npz.
The tests regenerate the identical inputs (checked by a hash of the
files) and compare the oracle and the HIP engine against them.
"""
from __future__ import annotations

import hashlib
import os

import numpy as np

from radlite_amd import synth

# name -> parameters.  Small grids exercise every branch the reference
# takes (refinement rings, dbdr insertions, the incl < 0.1 rad clamp,
# near edge-on views, CMB outer boundary, several lines in one run).
CASES = {
    "t1": dict(nr=8, nth=6, nphi=8, b_extra=-2, dbdr=1, nfr=12, vmax=30.0, incl=45.0, seed=1),
    "t2": dict(nr=10, nth=8, nphi=6, b_extra=-1, dbdr=2, nfr=15, vmax=25.0, incl=20.0, seed=2),
    "t3": dict(nr=12, nth=10, nphi=12, b_extra=-3, dbdr=1, nfr=21, vmax=40.0, incl=80.0, seed=3),
    "t4": dict(nr=8, nth=6, nphi=8, b_extra=-1, dbdr=1, nfr=9, vmax=20.0, incl=3.0, seed=4),
    "t5": dict(nr=9, nth=7, nphi=10, b_extra=-2, dbdr=3, nfr=16, vmax=35.0, incl=60.0, seed=5,
               outer_bc=2),
    "t6": dict(nr=10, nth=8, nphi=8, b_extra=-2, dbdr=1, nfr=24, vmax=30.0, incl=35.0, seed=6,
               nlines=3),
    "t7": dict(nr=16, nth=12, nphi=16, b_extra=-4, dbdr=2, nfr=30, vmax=30.0, incl=45.0, seed=7,
               t_scale=0.2),
    # the dust scattering/PAH source: scatsource.dat and quantsource.dat
    # (summed), interpolated to each of three line centres
    "t8": dict(nr=10, nth=8, nphi=8, b_extra=-2, dbdr=1, nfr=24, vmax=30.0, incl=35.0, seed=8,
               nlines=3, scat=True),
    # outer boundary type 3 (interstellar field per channel): with 20
    # channels, the lines whose centres fall in continuum interval 20 get a
    # zero field (the reference compares the interval with freq_nr, the
    # passband's channel count), the others the interpolated field
    "t9": dict(nr=10, nth=8, nphi=8, b_extra=-2, dbdr=1, nfr=20, vmax=30.0, incl=50.0, seed=7,
               nlines=4, outer_bc=3, lam_range=(4.4, 4.9)),
    # the CPU-runnable workload: nr=100 x nth=60, 110 rings x 64 phi, 50 ch, 45 deg
    "c1": dict(nr=100, nth=60, nphi=64, b_extra=-1, dbdr=1, nfr=50, vmax=30.0, incl=45.0, seed=0,
               sigma_t=0.01, image_stride=7),
}

# position-velocity cubes (rectangular images, linespec_command 2,
# telescope.F:196-224): an unresolved star added to the 4 central pixels;
# pixels inside the inner hole, some of whose rays pass within R*,
# rotated and offset; an image larger than the grid (corner rays not
# traced) with the CMB outer boundary
POSVEL_CASES = {
    "p1": dict(nr=10, nth=8, nphi=8, b_extra=-2, dbdr=1, nfr=24, vmax=30.0, incl=35.0, seed=6,
               nlines=3, posvel=dict(nx=16, ny=16, size_x=40.0 * synth.AU,
                                     size_y=40.0 * synth.AU, phioff=0.0, xoff=0.0, yoff=0.0,
                                     addstar=1)),
    "p2": dict(nr=8, nth=6, nphi=8, b_extra=-2, dbdr=1, nfr=12, vmax=30.0, incl=45.0, seed=1,
               posvel=dict(nx=12, ny=10, size_x=0.12 * synth.AU, size_y=0.1 * synth.AU,
                           phioff=0.3, xoff=0.002 * synth.AU, yoff=-0.001 * synth.AU,
                           addstar=0)),
    "p3": dict(nr=9, nth=7, nphi=10, b_extra=-2, dbdr=3, nfr=16, vmax=35.0, incl=60.0, seed=5,
               outer_bc=2, posvel=dict(nx=14, ny=14, size_x=230.0 * synth.AU,
                                       size_y=200.0 * synth.AU, phioff=-0.7,
                                       xoff=3.0 * synth.AU, yoff=0.0, addstar=0)),
    # outer boundary type 3 on the pos-vel path: calc_write_line_posvel never
    # fills the per-channel field, so rays start from (and pixels off the
    # grid hold) zero
    "p4": dict(nr=9, nth=7, nphi=10, b_extra=-2, dbdr=3, nfr=16, vmax=35.0, incl=60.0, seed=5,
               outer_bc=3, posvel=dict(nx=14, ny=14, size_x=230.0 * synth.AU,
                                       size_y=200.0 * synth.AU, phioff=-0.7,
                                       xoff=3.0 * synth.AU, yoff=0.0, addstar=0)),
}


def isrf_field(nu):
    """A synthetic interstellar field: 0.3 B_nu(1200 K)."""
    return 0.3 * synth.bnu(nu, 1200.0)


def build_case(name, path):
    c = dict(CASES[name] if name in CASES else POSVEL_CASES[name])
    disk = synth.make_disk(c["nr"], c["nth"], seed=c["seed"], sigma_t=c.get("sigma_t", 0.05),
                           t_scale=c.get("t_scale", 1.0))
    if c.get("nlines", 1) == 1:
        lines = synth.single_line()
    elif "lam_range" in c:
        lines = synth.random_lines(c["nlines"], seed=c["seed"] + 100, Jmax=10,
                                   lam_range=c["lam_range"])
    else:
        lines = synth.random_lines(c["nlines"], seed=c["seed"] + 100, Jmax=10)
    pops = synth.lte_pops(disk, lines,
                          q_mode="t_over_2.77" if c.get("nlines", 1) == 1 else "norm")
    vmax, dv = synth.passband(c["vmax"], c["nfr"])
    synth.write_run_dir(path, disk, lines, pops, nphi=c["nphi"], dbdr=c["dbdr"],
                        b_extra=c["b_extra"], vmax=vmax, dv=dv, incl_deg=c["incl"],
                        outer_bc=c.get("outer_bc", 0), posvel=c.get("posvel"),
                        scat={"scatsource.dat": lambda nu: synth.scat_source(disk, nu, c["seed"]),
                              "quantsource.dat": lambda nu: synth.scat_source(
                                  disk, nu, c["seed"] + 1, level=0.2)}
                        if c.get("scat") else None,
                        isrf=isrf_field if c.get("outer_bc", 0) == 3 else None)
    return c


INPUT_FILES = ("radlite.inp", "linespectrum.inp", "line.inp", "moldata.dat",
               "levelpop_moldata.dat", "radius.inp", "theta.inp", "density.inp",
               "abundance.inp", "velocity.inp", "turbulence.inp", "dustdens.inp",
               "dusttemp_final.dat", "frequency.inp", "dustopac_1.inp",
               "starinfo.inp", "starspectrum.inp")


OPTIONAL_INPUT_FILES = ("scatsource.dat", "quantsource.dat", "interstellfield.inp")


def input_hash(path):
    h = hashlib.sha256()
    for f in INPUT_FILES + tuple(f for f in OPTIONAL_INPUT_FILES
                                 if os.path.exists(os.path.join(path, f))):
        with open(os.path.join(path, f), "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()


def golden_path(name):
    return os.path.join(os.path.dirname(os.path.abspath(__file__)), f"{name}.npz")


def load_golden(name):
    z = np.load(golden_path(name), allow_pickle=False)
    return {k: z[k] for k in z.files}


def inputs_digest(ri) -> str:
    """sha256 of the arrays of an in-memory RunInputs (the workload
    generators' output), so that a fixture made from them detects drift."""
    h = hashlib.sha256()
    for k in ("r", "theta_half", "rho_gas", "abund", "vel", "lw_kms", "pops", "rho_dust",
              "t_dust", "cont_nu", "kappa_abs", "kappa_scat", "star_fnu", "E_cm", "g", "lev_up",
              "lev_low", "A"):
        h.update(np.ascontiguousarray(getattr(ri, k), dtype=np.float64).tobytes())
    h.update(repr((ri.rstar, ri.umass_av, ri.fix_lw, ri.nphi, ri.dbdr, ri.b_extra, ri.outer_bc,
                   ri.vmax, ri.dv)).encode())
    return h.hexdigest()
