"""A golden case on the reference's own RADMC inputs.

The reference ships the outputs of a RADMC continuum run for its tutorial
(DOCS/DOCS_VERSION_1-3/files_for_tutorials/radmc_outputs: a 120-radius x
100-theta grid with dust density and temperatures, a 130-point frequency
grid, dust opacities, star and line.inp).  Case r1 is a run directory made
of those files as they are, plus the gas-side files pyradlite writes
before a run (radlite.py:1481-1677, 1739-2221): gas density = 100 x dust
density, constant abundance, Keplerian velocity about the tutorial star's
mass (spherical radius), thermal + turbulent line width, LTE populations
of three CO-like lines at the dust temperature, radlite.inp and
linespectrum.inp.

The reference's files exist only where the reference is mounted (the
build container): build_radmc_case needs them.  make_golden.py stores the
inputs as read (l4io.read_run_dir) in r1.npz beside the reference's
results, so that radmc_inputs() rebuilds the same RunInputs anywhere, and
a CPU test checks that reading the tutorial files gives exactly those
arrays.
"""
from __future__ import annotations

import os
import shutil

import numpy as np

from radlite_amd import l4io, synth

TUTORIAL = "/root/reference/DOCS/DOCS_VERSION_1-3/files_for_tutorials/radmc_outputs"
COPIED = ("radius.inp", "theta.inp", "dustdens.inp", "dusttemp.info", "dusttemp_final.dat",
          "frequency.inp", "dustopac.inp", "dustopac_1.inp", "starinfo.inp",
          "starspectrum.inp", "line.inp")

RADMC_CASES = {
    "r1": dict(nphi=32, b_extra=-2, dbdr=1, nfr=40, vmax=30.0, incl=45.0, nlines=3, seed=31,
               image_stride=4,
               v_turb=0.1, m_mol=28.0, abund=1e-4, gas_to_dust=100.0),
}


def available():
    return os.path.isdir(TUTORIAL)


def build_radmc_case(name, path):
    """Write the run directory of case ``name`` (needs the reference's
    tutorial files)."""
    c = RADMC_CASES[name]
    os.makedirs(path, exist_ok=True)
    for f in COPIED:
        shutil.copy(os.path.join(TUTORIAL, f), os.path.join(path, f))
    t = l4io._tokens(os.path.join(path, "radius.inp"))
    nr = int(t[0])
    r = np.array(t[1:1 + nr], float)
    t = l4io._tokens(os.path.join(path, "theta.inp"))
    nth = int(t[0])
    th = np.array(t[2:2 + nth], float)
    t = l4io._tokens(os.path.join(path, "dustdens.inp"))
    rho_d = np.array(t[4:4 + nr * nth], float).reshape(nr, nth)
    t = l4io._tokens(os.path.join(path, "dusttemp_final.dat"))
    Td = np.array(t[5:5 + nr * nth], float).reshape(nr, nth)
    t = l4io._tokens(os.path.join(path, "starinfo.inp"))
    mstar = float(t[2])
    # pyradlite's gas-side fields: _calc_gasdensity (radlite.py:1549),
    # _calc_velocity (:1653-1664, spherical r), _calc_turbulence
    # (:1618-1622), constant abundance (:1513)
    rho_g = rho_d * c["gas_to_dust"]
    vel = np.zeros((nr, nth, 3))
    vel[..., 2] = np.sqrt(synth.G_CGS * mstar / r)[:, None]
    lw = np.sqrt(c["v_turb"] ** 2 + 2 * synth.KB * Td / (c["m_mol"] * synth.MH) / 1e10)
    abund = np.full((nr, nth), c["abund"])
    with open(os.path.join(path, "density.inp"), "w") as f:
        f.write(f"{nr} {nth} 1\n" + "\n".join(l4io._fmt(x) for x in rho_g.ravel()) + "\n")
    with open(os.path.join(path, "abundance.inp"), "w") as f:
        f.write(f"{nr} {nth}\n" + "\n".join(l4io._fmt(a) + " 0.0" for a in abund.ravel()) + "\n")
    with open(os.path.join(path, "velocity.inp"), "w") as f:
        f.write(f"{nr} {nth}\n")
        f.write("\n".join(" ".join(l4io._fmt(v) for v in row)
                          for row in vel.reshape(nr * nth, 3)) + "\n")
    with open(os.path.join(path, "turbulence.inp"), "w") as f:
        f.write(f"1\n{nr} {nth}\n" + "\n".join(l4io._fmt(x) for x in lw.ravel()) + "\n")
    lines = synth.random_lines(c["nlines"], seed=c["seed"], Jmax=20)
    disk = synth.Disk(r=r, theta_half=th, rho_gas=rho_g, t_gas=Td, abund=abund, vel=vel,
                      lw_kms=lw, rho_dust=rho_d, t_dust=Td)
    pops = synth.lte_pops(disk, lines, q_mode="renorm")
    l4io.write_moldata(path, "synth", lines.weight, lines.E_cm, lines.g, lines.lev_up,
                       lines.lev_low, lines.A)
    l4io.write_levelpop(path, pops)
    vmax, dv = synth.passband(c["vmax"], c["nfr"])
    l4io.write_radlite_inp(path, c["nphi"], c["dbdr"], c["b_extra"])
    l4io.write_linespectrum_inp(path, vmax, dv, c["incl"], c["nlines"])
    return c


# RunInputs fields stored in the golden file (everything the trace reads)
ARRAYS = ("r", "theta_half", "rho_gas", "abund", "vel", "lw_kms", "pops", "rho_dust", "t_dust",
          "cont_nu", "kappa_abs", "kappa_scat", "star_fnu", "E_cm", "g", "lev_up", "lev_low", "A")
SCALARS = ("rstar", "umass_av", "fix_lw", "nphi", "dbdr", "b_extra", "outer_bc", "vmax", "dv",
           "incl_deg", "vrad", "nlines_run", "iline_start")


def inputs_to_arrays(ri: l4io.RunInputs) -> dict:
    out = {"in_" + k: np.asarray(getattr(ri, k)) for k in ARRAYS}
    out.update({"in_" + k: np.array(getattr(ri, k)) for k in SCALARS})
    out["in_molname"] = np.array(ri.molname)
    return out


def radmc_inputs(g: dict) -> l4io.RunInputs:
    """The RunInputs of a stored case (no reference files needed)."""
    kw = {k: g["in_" + k] for k in ARRAYS}
    for k in SCALARS:
        v = g["in_" + k]
        kw[k] = v.item()
    kw["molname"] = str(g["in_molname"])
    return l4io.RunInputs(**kw)
