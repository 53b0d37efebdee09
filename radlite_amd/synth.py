"""Seeded synthetic disk models for the benchmark and parity workloads.

One named generator (numpy PCG64 ``default_rng(seed)``) per model.  The
physics is the one the workload definitions use:

* spherical grid, ``nr`` radii log-spaced over [0.1, 100] AU and
  ``nth_half`` theta cell centres (j - 1/2)(pi/2)/nth_half, mirrored
  about the midplane (theta.inp flag 1, grid.F:1104-1163);
* Sigma = 1000 (R/AU)^-1 g/cm^2, H/R = 0.05 (R/AU)^0.25, Gaussian
  vertical profile, with R = r sin(theta), z = r cos(theta);
* T = 1000 K (R/AU)^-0.5 times lognormal(0, sigma_T) per cell; dust
  temperature = T; gas/dust = 100; grey kappa_abs = 100 cm^2/g;
* Keplerian v_phi = sqrt(G M_sun / r) with the spherical r (as
  pyradlite's _calc_velocity, radlite.py:1653-1664); v_r = v_theta = 0;
* line width lw = sqrt(v_turb^2 + 2kT/(m_mol m_H)) in km/s
  (radlite.py:1618-1622) with v_turb = 0.1 km/s;
* abundance 1e-4, umass_av = 2.3;
* LTE populations n_i = g_i exp(-E_i/kT)/Q(T), Q = T/2.77
  (radlite.py:1111-1119 with the workload's closed-form Q).

The model for one line at 4.7 um with E_up = 3000 K has a lower level at
E_up - h nu0/k = -61.3 K: the reference derives nu0 from the level
energies (line.F:1980-1983), so the lower level energy is written as-is.
"""
from __future__ import annotations

import os
from dataclasses import dataclass

import numpy as np

from . import l4io

AU = 1.495978707e13
G_CGS = 6.674e-8
MSUN = 1.989e33
RSUN = 6.957e10
KB = 1.380649e-16
MH = 1.6726e-24
HPLANCK = 6.62607015e-27
CLIGHT = 2.99792458e10
PC = 3.08572e18
K_PER_CM = HPLANCK * CLIGHT / KB          # 1.4387769 K per cm^-1


@dataclass
class Disk:
    r: np.ndarray
    theta_half: np.ndarray
    rho_gas: np.ndarray
    t_gas: np.ndarray
    abund: np.ndarray
    vel: np.ndarray
    lw_kms: np.ndarray
    rho_dust: np.ndarray
    t_dust: np.ndarray


def make_disk(nr, nth_half, seed=0, sigma_t=0.01, r_in_au=0.1, r_out_au=100.0,
              m_mol=28.0, v_turb_kms=0.1, abund=1e-4, t_scale=1.0):
    rng = np.random.default_rng(seed)
    r = np.logspace(np.log10(r_in_au), np.log10(r_out_au), nr) * AU
    th = (np.arange(nth_half) + 0.5) * (np.pi / 2) / nth_half
    R = r[:, None] * np.sin(th)[None, :]
    z = r[:, None] * np.cos(th)[None, :]
    Rau = R / AU
    sigma = 1000.0 / Rau
    H = 0.05 * Rau ** 0.25 * R
    rho = sigma / (np.sqrt(2 * np.pi) * H) * np.exp(-z * z / (2 * H * H))
    T = 1000.0 * Rau ** -0.5 * np.exp(rng.normal(0.0, sigma_t, size=R.shape)) * t_scale
    vphi = np.sqrt(G_CGS * MSUN / r)[:, None] * np.ones_like(R)
    vel = np.zeros(R.shape + (3,))
    vel[..., 2] = vphi
    lw = np.sqrt(v_turb_kms ** 2 + 2 * KB * T / (m_mol * MH) / 1e10)
    return Disk(r=r, theta_half=th, rho_gas=rho, t_gas=T,
                abund=np.full(R.shape, abund), vel=vel, lw_kms=lw,
                rho_dust=rho / 100.0, t_dust=T.copy())


@dataclass
class LineSet:
    E_cm: np.ndarray         # level energies [cm^-1] (quantised as read)
    g: np.ndarray
    lev_up: np.ndarray       # 1-based
    lev_low: np.ndarray
    A: np.ndarray            # quantised as read
    weight: float = 28.0


def single_line(lam_um=4.7, A=30.0, E_up_K=3000.0, g_up=21.0, g_low=19.0):
    """Two-level molecule with one line (level 1 = lower, 2 = upper)."""
    dE_cm = 1e4 / lam_um
    Eu = E_up_K / K_PER_CM
    El = Eu - dE_cm
    Eq, gq, Aq = l4io.quantize_moldata([El, Eu], [g_low, g_up], [A])
    return LineSet(E_cm=Eq, g=gq, lev_up=np.array([2]), lev_low=np.array([1]), A=Aq)


def random_lines(nlines, seed, lam_range=(4.5, 5.0), logA=(0.0, 2.0),
                 Eup=(100.0, 6000.0), Jmax=40):
    """Independent two-level pairs per line, stacked into one level list
    (line k uses levels 2k+1 (lower) and 2k+2 (upper))."""
    rng = np.random.default_rng(seed)
    E, g, up, lo, A = [], [], [], [], []
    for k in range(nlines):
        lam = rng.uniform(*lam_range)
        a = 10 ** rng.uniform(*logA)
        eu_k = rng.uniform(*Eup)
        J = int(rng.integers(1, Jmax + 1))
        eu = eu_k / K_PER_CM
        el = eu - 1e4 / lam
        E += [el, eu]
        g += [2 * J - 1, 2 * J + 1]
        lo.append(2 * k + 1); up.append(2 * k + 2); A.append(a)
    Eq, gq, Aq = l4io.quantize_moldata(E, g, A)
    return LineSet(E_cm=Eq, g=gq, lev_up=np.array(up), lev_low=np.array(lo), A=Aq)


def rovib_lines(nlines, seed, nv=6, jmax=60, e_vib_k=1400.0, b_rot_k=2.0, Eup=(100.0, 8000.0),
                logA=(0.0, 2.0)):
    """A ro-vibrational-like ladder: levels (v, J) for v < nv, J <= jmax with
    E = v e_vib + b J(J+1) [K] and g = 2J+1; lines (v', J') -> (v'', J'')
    with v'' < v', |J' - J''| = 1 and E_up in ``Eup``, ``nlines`` of them
    drawn without replacement, A log-uniform in 10**logA."""
    rng = np.random.default_rng(seed)
    nl = jmax + 1
    EK = np.array([v * e_vib_k + b_rot_k * J * (J + 1) for v in range(nv) for J in range(nl)])
    g = np.array([2.0 * J + 1.0 for v in range(nv) for J in range(nl)])
    cand = []
    for vu in range(1, nv):
        for vl in range(vu):
            for Ju in range(nl):
                for Jl in (Ju - 1, Ju + 1):
                    if 0 <= Jl < nl:
                        iu, il = vu * nl + Ju, vl * nl + Jl
                        if Eup[0] <= EK[iu] <= Eup[1]:
                            cand.append((iu, il))
    if len(cand) < nlines:
        raise ValueError(f"{len(cand)} candidate lines < {nlines}")
    pick = rng.choice(len(cand), size=nlines, replace=False)
    up = np.array([cand[k][0] + 1 for k in pick])
    lo = np.array([cand[k][1] + 1 for k in pick])
    A = 10 ** rng.uniform(*logA, size=nlines)
    Eq, gq, Aq = l4io.quantize_moldata(list(EK / K_PER_CM), list(g), list(A))
    return LineSet(E_cm=Eq, g=gq, lev_up=up, lev_low=lo, A=Aq)


def nonlte_pops(disk: Disk, lines: LineSet, sigma=0.2, seed=0):
    """Boltzmann populations times lognormal(sigma) per cell and level,
    renormalised to sum 1 per cell; [nr, nth, nlev]."""
    rng = np.random.default_rng(seed)
    T = disk.t_gas[..., None]
    EK = lines.E_cm[None, None, :] * K_PER_CM
    n = lines.g[None, None, :] * np.exp(-EK / T)
    n *= np.exp(rng.normal(0.0, sigma, size=n.shape))
    n /= n.sum(axis=-1, keepdims=True)
    n[n < 1e-99] = 0.0
    return n


def lte_pops(disk: Disk, lines: LineSet, q_mode="t_over_2.77"):
    """n_i = g_i exp(-E_i/kT)/Q(T); shape [nr, nth, nlev]."""
    T = disk.t_gas[..., None]
    EK = lines.E_cm[None, None, :] * K_PER_CM
    n = lines.g[None, None, :] * np.exp(-EK / T)
    if q_mode == "t_over_2.77":
        n = n / (T / 2.77)
    else:
        n = n / n.sum(axis=-1, keepdims=True)
    n[n < 1e-99] = 0.0
    return n


def star_and_dust(nfreq=40, rstar=2 * RSUN, tstar=4000.0, kappa=100.0):
    """Frequency grid bracketing 1-20 um (ascending nu), grey dust, and a
    blackbody star written as F_nu at 1 pc."""
    lam = np.logspace(np.log10(20.0), np.log10(1.0), nfreq) * 1e-4
    nu = CLIGHT / lam
    x = HPLANCK * nu / (KB * tstar)
    bnu = 2 * HPLANCK * nu ** 3 / CLIGHT ** 2 / np.expm1(x)
    fnu = np.pi * bnu * (rstar / PC) ** 2
    return nu, np.full(nfreq, kappa), np.zeros(nfreq), rstar, fnu


def passband(vmax_target, nfr):
    """(vmax, dv) such that the reference's int(2 vmax/dv + 1) == nfr
    (main.F:209): the quotient is placed just above nfr - 1."""
    dv = 2.0 * vmax_target / (nfr - 1) * (1.0 - 1e-9)
    return vmax_target, dv


def bnu(nu, T):
    """Planck B_nu(T) [erg/s/cm^2/Hz/sr]."""
    return 2 * HPLANCK * nu ** 3 / CLIGHT ** 2 / np.expm1(HPLANCK * nu / (KB * T))


def scat_source(disk: Disk, nu, seed, kappa=100.0, level=0.5):
    """A synthetic dust scattering/PAH emissivity [nfreq, nr, nth] of the
    size of the thermal dust emissivity: level x rho_d kappa B_nu(T_d)
    times a per-cell lognormal(0.5) factor and a slope in frequency."""
    rng = np.random.default_rng(seed)
    T = disk.t_dust
    th = disk.rho_dust[None] * kappa * bnu(nu[:, None, None], T[None])
    f = np.exp(rng.normal(0.0, 0.5, size=T.shape))[None]
    slope = (nu / nu[len(nu) // 2])[:, None, None] ** 0.5
    return level * th * f * slope


def write_run_dir(path, disk: Disk, lines: LineSet, pops, *, nphi, dbdr,
                  b_extra, vmax, dv, incl_deg, nlines_run=None, outer_bc=0,
                  fix_lw=0.0, umass_av=2.3, nfreq=40, posvel=None, scat=None, isrf=None):
    """A RADLite run directory.  scat: optional {file name: [nfreq, nr, nth]}
    written as scatsource.dat / quantsource.dat; isrf: optional
    function of the continuum frequencies giving interstellfield.inp."""
    os.makedirs(path, exist_ok=True)
    nl = len(lines.A)
    l4io.write_radlite_inp(path, nphi, dbdr, b_extra, outer_bc=outer_bc)
    l4io.write_linespectrum_inp(path, vmax, dv, incl_deg,
                                nl if nlines_run is None else nlines_run, posvel=posvel)
    l4io.write_line_inp(path, nl, umass_av=umass_av, fix_lw=fix_lw)
    l4io.write_moldata(path, "synth", lines.weight, lines.E_cm, lines.g,
                       lines.lev_up, lines.lev_low, lines.A)
    l4io.write_levelpop(path, pops)
    l4io.write_fields(path, disk.r, disk.theta_half, disk.rho_gas, disk.abund,
                      disk.vel, disk.lw_kms, disk.rho_dust, disk.t_dust)
    nu, ka, ks, rstar, fnu = star_and_dust(nfreq=nfreq)
    l4io.write_dust_and_star(path, nu, ka, ks, rstar, fnu)
    for name, arr in (scat or {}).items():
        l4io.write_scatsource(path, arr(nu) if callable(arr) else arr, name)
    if isrf is not None:
        l4io.write_interstellfield(path, isrf(nu))


# ----------------------------------------------------------------------
# the workloads
# ----------------------------------------------------------------------
WORKLOADS = {
    # name: grid, rays, channels, inclinations, noise
    "c1": dict(nr=100, nth=60, nphi=64, b_extra=-1, dbdr=1, nfr=50, vmax=30.0,
               incl=(45.0,), sigma_t=0.01),
    "c2": dict(nr=400, nth=200, nphi=256, b_extra=-113, dbdr=1, nfr=200, vmax=40.0,
               incl=(30.0, 60.0), sigma_t=0.1),
    "c4": dict(nr=200, nth=100, nphi=128, b_extra=-57, dbdr=1, nfr=100, vmax=30.0,
               incl=(45.0,), sigma_t=0.0),
}


def write_workload(path, name, seed=0, **over):
    w = dict(WORKLOADS[name]); w.update(over)
    disk = make_disk(w["nr"], w["nth"], seed=seed, sigma_t=w["sigma_t"])
    lines = single_line()
    pops = lte_pops(disk, lines)
    vmax, dv = passband(w["vmax"], w["nfr"])
    write_run_dir(path, disk, lines, pops, nphi=w["nphi"], dbdr=w["dbdr"],
                  b_extra=w["b_extra"], vmax=vmax, dv=dv, incl_deg=w["incl"][0])
    return w


def run_inputs(disk: Disk, lines: LineSet, pops, *, nphi, dbdr, b_extra, vmax, dv,
               incl_deg, outer_bc=0, fix_lw=0.0, umass_av=2.3, nfreq=40):
    """The same RunInputs that l4io.read_run_dir would return for the
    directory write_run_dir writes (every value round-trips through the
    %.17e text), built in memory."""
    nu, ka, ks, rstar, fnu = star_and_dust(nfreq=nfreq)
    return l4io.RunInputs(
        r=disk.r, theta_half=disk.theta_half, rho_gas=disk.rho_gas, abund=disk.abund,
        vel=disk.vel, lw_kms=disk.lw_kms, pops=pops, rho_dust=disk.rho_dust[None],
        t_dust=disk.t_dust[None], cont_nu=nu, kappa_abs=ka[None], kappa_scat=ks[None],
        rstar=rstar, star_fnu=fnu, umass_av=umass_av, fix_lw=fix_lw, E_cm=lines.E_cm,
        g=lines.g, lev_up=lines.lev_up, lev_low=lines.lev_low, A=lines.A, molname="moldata",
        nphi=nphi, dbdr=dbdr, b_extra=b_extra, outer_bc=outer_bc, vmax=vmax, dv=dv,
        incl_deg=incl_deg, vrad=0.0, nlines_run=len(lines.A), iline_start=1)


def workload_inputs(name, seed=0, **over):
    """RunInputs of a named workload (first inclination in incl_deg)."""
    w = dict(WORKLOADS[name]); w.update(over)
    disk = make_disk(w["nr"], w["nth"], seed=seed, sigma_t=w["sigma_t"])
    lines = single_line()
    pops = lte_pops(disk, lines)
    vmax, dv = passband(w["vmax"], w["nfr"])
    ri = run_inputs(disk, lines, pops, nphi=w["nphi"], dbdr=w["dbdr"], b_extra=w["b_extra"],
                    vmax=vmax, dv=dv, incl_deg=w["incl"][0])
    return ri, w


def c3_inputs(seed=0, nlines=64, incl=(10.0, 20.0, 30.0, 40.0, 50.0, 60.0, 70.0, 80.0),
              nr=256, nth=128, nphi=64, vmax=30.0, nfr_range=(20, 400)):
    """The batched workload: nr x nth disk, ``nlines`` random lines (lambda0
    in 4.5-5.0 um, log10 A in [0, 2], E_up in 100-6000 K, g = 2J+1, 2J-1
    for J in 1..40), populations Boltzmann-then-renormalised, one channel
    count per line drawn from ``nfr_range`` (inclusive) over +-vmax.
    Returns (RunInputs, items) with items (line, incl index, nfr, vmax)
    for every line at every inclination."""
    rng = np.random.default_rng(seed)
    disk = make_disk(nr, nth, seed=seed, sigma_t=0.1)
    lines = random_lines(nlines, seed=seed + 1)
    pops = lte_pops(disk, lines, q_mode="renorm")
    nfrs = rng.integers(nfr_range[0], nfr_range[1] + 1, size=nlines)
    v0, dv0 = passband(vmax, int(nfrs[0]))
    ri = run_inputs(disk, lines, pops, nphi=nphi, dbdr=1, b_extra=-1, vmax=v0, dv=dv0,
                    incl_deg=incl[0])
    items = []
    for ii in range(len(incl)):
        for il in range(nlines):
            v, dv = passband(vmax, int(nfrs[il]))
            n = int(2.0 * v / dv + 1)
            items.append((il, ii, n, 0.5 * (n - 1) * dv))
    return ri, items, list(incl)


def c5_inputs(seed=0, nlines=1000, incl=(20.0, 40.0, 60.0, 80.0), nr=1024, nth=512, nphi=512,
              nfr=300, vmax=30.0, keep_lines=None):
    """The large workload: nr x nth disk, ``nlines`` ro-vibrational-like
    lines (J = 0-60 ladders, A = 1-100 s^-1, E_up = 100-8000 K), non-LTE
    populations (Boltzmann x lognormal(0.2), renormalised), ``nfr``
    channels per line over +-vmax, nphi azimuths, the given inclinations.
    Returns (RunInputs, items) with items (line, incl index, nfr, vmax) for
    every line at every inclination, inclination-major.  keep_lines = k
    keeps only the first k of the drawn lines (same levels, populations and
    disk): a slice of the same job whose per-line tables are cheap."""
    disk = make_disk(nr, nth, seed=seed, sigma_t=0.1)
    lines = rovib_lines(nlines, seed=seed + 1)
    pops = nonlte_pops(disk, lines, seed=seed + 2)
    if keep_lines is not None:
        k = int(keep_lines)
        lines = LineSet(E_cm=lines.E_cm, g=lines.g, lev_up=lines.lev_up[:k],
                        lev_low=lines.lev_low[:k], A=lines.A[:k])
        nlines = k
    v, dv = passband(vmax, nfr)
    ri = run_inputs(disk, lines, pops, nphi=nphi, dbdr=1, b_extra=-1, vmax=v, dv=dv,
                    incl_deg=incl[0])
    n = int(2.0 * v / dv + 1)
    items = [(il, ii, n, 0.5 * (n - 1) * dv) for ii in range(len(incl)) for il in range(nlines)]
    return ri, items, list(incl)
