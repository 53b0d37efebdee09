"""Differentiable line spectra: pyradlite's pre-pass in torch and the
raytrace as a torch.autograd.Function whose backward is the HIP adjoint.

Chain (all per cell on the half theta grid [nr, nth]):

* LTE populations n_i = g_i exp(-E_i/kT) / Q(T), Q = T/2.77 (the
  closed-form partition function of the workloads) or Q interpolated
  linearly in a table and extrapolated past its ends (pyradlite
  ``_prep_mol_forcore``, radlite.py:1111-1119, with the ParSum.dat
  interpolation of radlite.py:1150-1153);
* line width lw = sqrt(v_turb^2 + 2kT/(m_mol m_p)) in km/s, or the
  alpha model sqrt((alpha c_s)^2 + 2kT/(m_mol m_p))
  (``_calc_turbulence``, radlite.py:1613-1622);
* Keplerian velocity v_phi = sqrt(G M / r) (``_calc_velocity``,
  radlite.py:1653-1664);
* abundance = exp(log_abund);
* dust emissivity at each line centre j_d = rho_d kappa_abs(nu0)
  (B_nu(T) + 1e-290) and extinction a_d = rho_d (kappa_abs + kappa_scat)
  (line_dust_compute_src_alp, line.F:3560-3607; bplanck, setup.F:937-952);
* the trace: flux = F(rho, abund, lw, pops, jd, ad) through librl, its
  gradient through rl_trace_lines_adjoint.

PyTorch is the autograd plumbing here; the trace and its adjoint run in
the HIP library.  Tensors are float64 on the CPU (the C ABI takes host
arrays).  ``flux_from_physics`` runs the pre-pass in torch and hands the
tables to the trace; ``flux_from_physics_device`` runs the same pre-pass
on the device (rl_set_physics) and its adjoint there (rl_physics_adjoint),
so that only the three per-cell fields and their gradients cross PCIe.
"""
from __future__ import annotations

import numpy as np
import torch

from . import engine, synth


# pyradlite's constants: astropy's CODATA 2018 values (radlite.py:43-51)
KB_PYR = 1.380649e-16            # erg/K
MP_PYR = 1.67262192369e-24       # proton mass [g]
G_PYR = 6.6743e-8                # cm^3 g^-1 s^-2


def partition_sum(T, T_tab, Q_tab):
    """Q(T) from a table as interp1d(kind="linear", bounds_error=False,
    fill_value="extrapolate") evaluates it (radlite.py:1150-1153): the
    interval is the first tabulated T >= T, clipped to the first and last
    intervals, Q = slope (T - T_lo) + Q_lo."""
    x = torch.as_tensor(np.asarray(T_tab, float), dtype=T.dtype)
    y = torch.as_tensor(np.asarray(Q_tab, float), dtype=T.dtype)
    i = torch.searchsorted(x, T.detach().contiguous()).clamp(1, len(x) - 1)
    lo = i - 1
    slope = (y[i] - y[lo]) / (x[i] - x[lo])
    return slope * (T - x[lo]) + y[lo]


def lte_populations(T, E_cm, g, q_over=2.77, q_table=None):
    """n_i [nr, nth, nlev] = g_i exp(-E_i/kT) / Q(T), Q = T/q_over or, with
    q_table = (T_tab, Q_tab), the interpolated table; below 1e-99 -> 0."""
    EK = torch.as_tensor(np.asarray(E_cm) * synth.K_PER_CM, dtype=T.dtype)
    gg = torch.as_tensor(np.asarray(g, float), dtype=T.dtype)
    Tc = T[..., None]
    Q = (Tc / q_over) if q_table is None else partition_sum(T, *q_table)[..., None]
    n = gg * torch.exp(-EK / Tc) / Q
    return torch.where(n < 1e-99, torch.zeros_like(n), n)


def line_width_kms(T, v_turb_kms, m_mol=28.0, m_p=MP_PYR):
    return torch.sqrt(v_turb_kms ** 2 + 2 * KB_PYR * T / (m_mol * m_p) / 1e10)


def line_width_alpha_kms(T_dust, alpha, gamma=1.4, mu=2.3, m_mol=28.0):
    """pyradlite's turbulence (radlite.py:1613-1622) in km/s: the alpha
    fraction of the sound speed and the thermal broadening in quadrature,
    both at the dust temperature."""
    cs = torch.sqrt(gamma * KB_PYR * T_dust / (mu * MP_PYR))
    th = torch.sqrt(2.0 * KB_PYR * T_dust / (m_mol * MP_PYR))
    return torch.sqrt((alpha * cs) ** 2 + th ** 2) / 1e5


def kepler_velocity(r, nth, mstar):
    """(v_r, v_theta, v_phi) [nr, nth, 3] in cm/s about a star of mass
    mstar [g] (radlite.py:1653-1664), r the cell radii [nr]."""
    v = np.zeros((len(r), nth, 3))
    v[..., 2] = np.sqrt(G_PYR * mstar / np.asarray(r, float))[:, None]
    return v


def kappa_at_lines(cont_nu, kappa, nu0):
    """kappa on the continuum grid interpolated at each |nu0| the way
    line_dust_compute_src_alp does (hunt, then linear; 0 outside)."""
    x = np.asarray(cont_nu, float)
    out = np.zeros(len(nu0))
    asc = x[-1] > x[0]
    for i, f in enumerate(np.abs(nu0)):
        n = len(x)
        lo, hi = 0, n + 1
        while hi - lo != 1:
            m = (hi + lo) >> 1
            if (f > x[m - 1]) == asc:
                lo = m
            else:
                hi = m
        if lo == 0 or lo == n:
            continue
        w = (f - x[lo - 1]) / (x[lo] - x[lo - 1])
        out[i] = w * kappa[lo] + (1.0 - w) * kappa[lo - 1]
    return out


def dust_line(T_dust, rho_dust, kabs_nu0, kscat_nu0, nu0):
    """(j_d, a_d) [nr, nth, nlines] at the line centres for one dust species."""
    nu = torch.as_tensor(np.abs(np.asarray(nu0, float)), dtype=T_dust.dtype)
    ka = torch.as_tensor(kabs_nu0, dtype=T_dust.dtype)
    ks = torch.as_tensor(kscat_nu0, dtype=T_dust.dtype)
    Tc = T_dust[..., None]
    bnu = 1.47455e-47 * nu ** 3 / (torch.exp(4.7989e-11 * nu / Tc) - 1.0) + 1.0e-290
    rd = rho_dust[..., None]
    return rd * ka * bnu, rd * (ka + ks)


class _Trace(torch.autograd.Function):
    @staticmethod
    def forward(ctx, rho, abund, lw, pops, jd, ad, model):
        model._upload(rho, abund, lw, pops, jd, ad)
        fl = None
        if any(ctx.needs_input_grad[:6]):
            # keep the adjoint's tape (I checkpoints) on the device, unless it
            # does not fit the memory budget (the adjoint then recomputes)
            try:
                fl = model.rt.trace_tape(model.items)
            except engine.RadliteError as e:
                if e.code != 13:
                    raise
        if fl is None:
            fl, _, _ = model.rt.trace(model.items)
        model._calls += 1
        ctx.model = model
        ctx.call = model._calls
        ctx.save_for_backward(rho, abund, lw, pops, jd, ad)
        return torch.from_numpy(np.concatenate(fl))

    @staticmethod
    def backward(ctx, g):
        m = ctx.model
        if m._calls != ctx.call:
            # another forward ran since: restore the inputs this forward saw
            # (the adjoint then recomputes its forward instead of the tape)
            m._upload(*ctx.saved_tensors)
            m._calls += 1
        g = g.detach().cpu().numpy().astype(np.float64)
        parts, o = [], 0
        for it in m.items:
            parts.append(g[o:o + it[2]])
            o += it[2]
        _, gr = m.rt.adjoint(m.items, parts)
        t = lambda a: torch.from_numpy(a)
        return (t(gr["g_rho"]), t(gr["g_abund"]), t(gr["g_lw"]), t(gr["g_pops"]),
                t(gr["g_jd"]), t(gr["g_ad"]), None)


class _PhysTrace(torch.autograd.Function):
    """flux(T, log_abund, v_turb) with the pre-pass and the trace on the
    device (rl_set_physics, rl_trace_lines_forward_tape); the backward is
    rl_trace_lines_adjoint followed by rl_physics_adjoint, the cell
    gradients staying on the device."""

    @staticmethod
    def forward(ctx, T, log_abund, v_turb, model):
        model._set_physics(T, log_abund, v_turb)
        fl = None
        if any(ctx.needs_input_grad[:3]):
            try:
                fl = model.rt.trace_tape(model.items)
            except engine.RadliteError as e:
                if e.code != 13:
                    raise
        if fl is None:
            fl, _, _ = model.rt.trace(model.items)
        model._calls += 1
        ctx.model = model
        ctx.call = model._calls
        ctx.save_for_backward(T, log_abund, v_turb)
        return torch.from_numpy(np.concatenate(fl))

    @staticmethod
    def backward(ctx, g):
        m = ctx.model
        if m._calls != ctx.call:
            m._set_physics(*ctx.saved_tensors)
            m._calls += 1
        g = g.detach().cpu().numpy().astype(np.float64)
        parts, o = [], 0
        for it in m.items:
            parts.append(g[o:o + it[2]])
            o += it[2]
        m.rt.adjoint(m.items, parts, outputs=())
        gr = m.rt.physics_adjoint()
        t = lambda a: torch.from_numpy(a)
        return t(gr["g_T"]), t(gr["g_log_abund"]), t(gr["g_v_turb"]), None


class SpectrumModel:
    """A loaded run (grid, velocity, lines, star, rays) whose spectrum is
    differentiable with respect to the per-cell fields.

    ``flux(rho, abund, lw, pops, jd, ad)`` returns the concatenated
    spectra of ``items`` (every item SKIP_NONE) as a torch tensor."""

    def __init__(self, ri, precision="fp64", incl_rad=None, device=0, items=None):
        self.ri = ri
        self._calls = 0
        self._device_physics = False
        self._m_mol = 28.0
        self.rt = engine.Raytracer(device, precision)
        self.tables = self.rt.load_run(ri)
        inc = [ri.incl_deg * 0.0174532925199] if incl_rad is None else list(incl_rad)
        self.rt.setup_rays(inc, ri.nphi, ri.b_extra, ri.dbdr)
        nfr, vmax = engine.passband(ri.vmax, ri.dv)
        self.nfr, self.vmax = nfr, vmax
        self.items = items or [(il, ii, nfr, vmax, engine.SKIP_NONE)
                               for ii in range(len(inc)) for il in range(len(ri.A))]
        tb = self.tables
        self.kabs_nu0 = kappa_at_lines(ri.cont_nu, ri.kappa_abs[0], tb.nu0)
        self.kscat_nu0 = kappa_at_lines(ri.cont_nu, ri.kappa_scat[0], tb.nu0)

    def _upload(self, rho, abund, lw, pops, jd, ad):
        n = lambda x: np.ascontiguousarray(x.detach().cpu().numpy(), dtype=np.float64)
        lwn = n(lw)
        aks = float(lwn.max())
        self.rt.set_medium(n(rho), n(abund), self.ri.vel, lwn, self.ri.umass_av, aks)
        self.rt.set_levels(n(pops))
        self.rt.set_dust_line(n(jd), n(ad))

    def flux(self, rho, abund, lw, pops, jd, ad):
        return _Trace.apply(rho, abund, lw, pops, jd, ad, self)

    def _set_physics(self, T, log_abund, v_turb):
        n = lambda x: np.ascontiguousarray(x.detach().cpu().numpy(), dtype=np.float64)
        self.rt.set_physics(n(T), n(v_turb), log_abund=n(log_abund), m_mol=self._m_mol,
                            q_over=2.77)

    def flux_from_physics_device(self, T, log_abund, v_turb_kms, m_mol=28.0):
        """flux_from_physics with the pre-pass on the device (and its
        adjoint): per call only T, log-abundance and v_turb cross to the
        device and their gradients back; the dust temperature is T, for
        every dust species of the run."""
        if not self._device_physics:
            ri = self.ri
            self.rt.set_level_data(np.asarray(ri.E_cm, float) * synth.K_PER_CM,
                                   np.asarray(ri.g, float))
            ns = ri.rho_dust.shape[0]
            ka = np.array([kappa_at_lines(ri.cont_nu, ri.kappa_abs[s], self.tables.nu0)
                           for s in range(ns)])
            ks = np.array([kappa_at_lines(ri.cont_nu, ri.kappa_scat[s], self.tables.nu0)
                           for s in range(ns)])
            self.rt.set_dust_species(ri.rho_dust, ka, ks)
            self._device_physics = True
        self._m_mol = m_mol
        return _PhysTrace.apply(T, log_abund, v_turb_kms, self)

    def flux_from_physics(self, T, log_abund, v_turb_kms, m_mol=28.0):
        """Spectrum as a function of per-cell temperature (gas = dust),
        log-abundance and turbulent width, through the pre-pass."""
        ri = self.ri
        rho = torch.as_tensor(ri.rho_gas, dtype=torch.float64)
        pops = lte_populations(T, ri.E_cm, ri.g)
        lw = line_width_kms(T, v_turb_kms, m_mol)
        rho_d = torch.as_tensor(ri.rho_dust[0], dtype=torch.float64)
        jd, ad = dust_line(T, rho_d, self.kabs_nu0, self.kscat_nu0, self.tables.nu0)
        return self.flux(rho, torch.exp(log_abund), lw, pops, jd, ad)

    def close(self):
        self.rt.close()
