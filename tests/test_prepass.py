"""The differentiable pre-pass (radlite_amd.diff) against the generator's
numpy formulas and the library's host pre-pass (rl_prepare_line_dust,
the reference's line_dust_compute_src_alp arithmetic), and its
derivatives against torch's numerical gradient check.  CPU only."""
import numpy as np
import torch

from radlite_amd import diff, engine, synth


def _model():
    disk = synth.make_disk(12, 8, seed=5, sigma_t=0.1)
    lines = synth.random_lines(3, seed=2)
    pops = synth.lte_pops(disk, lines)
    vmax, dv = synth.passband(30.0, 9)
    ri = synth.run_inputs(disk, lines, pops, nphi=6, dbdr=1, b_extra=-1, vmax=vmax, dv=dv,
                          incl_deg=40.0)
    return disk, lines, pops, ri


def test_populations_and_width_match_generator():
    disk, lines, pops, ri = _model()
    T = torch.tensor(disk.t_gas)
    n = diff.lte_populations(T, lines.E_cm, lines.g).numpy()
    assert np.allclose(n, pops, rtol=1e-14, atol=0)
    lw = diff.line_width_kms(T, 0.1).numpy()
    assert np.allclose(lw, disk.lw_kms, rtol=1e-15, atol=0)


def test_dust_line_matches_library_prepass():
    disk, lines, pops, ri = _model()
    tb = engine.prepare_tables(ri)
    ka = diff.kappa_at_lines(ri.cont_nu, ri.kappa_abs[0], tb.nu0)
    ks = diff.kappa_at_lines(ri.cont_nu, ri.kappa_scat[0], tb.nu0)
    jd, ad = diff.dust_line(torch.tensor(ri.t_dust[0]), torch.tensor(ri.rho_dust[0]), ka, ks,
                            tb.nu0)
    assert np.allclose(jd.numpy(), tb.jd, rtol=1e-13, atol=0)
    assert np.allclose(ad.numpy(), tb.ad, rtol=1e-13, atol=0)


def test_prepass_derivatives():
    disk, lines, pops, ri = _model()
    T = torch.tensor(disk.t_gas[:3, :2], requires_grad=True)
    assert torch.autograd.gradcheck(lambda t: diff.lte_populations(t, lines.E_cm, lines.g),
                                    (T,), eps=1e-4, rtol=1e-5, atol=0)
    vt = torch.full_like(T, 0.1).requires_grad_(True)
    assert torch.autograd.gradcheck(lambda t, v: diff.line_width_kms(t, v), (T, vt),
                                    eps=1e-6, rtol=1e-6, atol=0)
    rd = torch.tensor(disk.rho_dust[:3, :2])
    nu0 = np.array([6.3e13, 6.5e13])
    assert torch.autograd.gradcheck(
        lambda t: diff.dust_line(t, rd, np.array([100.0, 90.0]), np.zeros(2), nu0)[0],
        (T,), eps=1e-4, rtol=1e-5, atol=0)


def test_large_workload_inputs_small_instance():
    """The 1000-line workload's generator on a small grid: ro-vibrational
    ladder lines (J' - J'' = +-1, v' > v'', E_up in 100-8000 K, A in
    1-100 s^-1, g = 2J+1), non-LTE populations that sum to 1 per cell and
    deviate from Boltzmann by the lognormal factor, items inclination-major."""
    ri, items, incl = synth.c5_inputs(0, nlines=200, nr=32, nth=16, nphi=8, nfr=30)
    assert len(items) == 200 * 4 and items[0][:2] == (0, 0) and items[200][:2] == (0, 1)
    assert all(it[2] == 30 for it in items)
    nl = 61
    up, lo = ri.lev_up - 1, ri.lev_low - 1
    vu, ju, vl, jl = up // nl, up % nl, lo // nl, lo % nl
    assert np.all(vu > vl) and np.all(np.abs(ju - jl) == 1)
    EK = ri.E_cm * synth.K_PER_CM
    assert np.all((EK[up] >= 100.0 - 1.0) & (EK[up] <= 8000.0 + 1.0))
    assert np.all((ri.A >= 1.0 * (1 - 1e-3)) & (ri.A <= 100.0 * (1 + 1e-3)))
    assert np.allclose(ri.g, 2.0 * (np.arange(len(ri.g)) % nl) + 1.0)
    assert len(set(zip(up, lo))) == 200                    # drawn without replacement
    p = ri.pops
    assert p.shape == (32, 16, 6 * nl) and np.allclose(p.sum(-1), 1.0, rtol=0, atol=1e-12)
    # the lognormal factor: log(p / Boltzmann) minus its per-cell offset has
    # the stated spread on the levels Boltzmann does not push to zero
    T = synth.make_disk(32, 16, seed=0, sigma_t=0.1).t_gas[..., None]
    b = ri.g[None, None, :] * np.exp(-EK[None, None, :] / T)
    ok = (b > 1e-30 * b.max(-1, keepdims=True)) & (p > 0)
    r = np.where(ok, np.log(np.where(ok, p, 1.0) / np.where(ok, b, 1.0)), 0.0)
    r = r - (r.sum(-1, keepdims=True) / ok.sum(-1, keepdims=True))
    sd = np.sqrt((r[ok] ** 2).mean())
    assert 0.18 < sd < 0.22, sd
