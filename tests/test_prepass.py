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
    lw = diff.line_width_kms(T, 0.1, m_p=synth.MH).numpy()   # the generator's m_H
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
                                    eps=1e-5, rtol=1e-5, atol=0)
    assert torch.autograd.gradcheck(lambda t: diff.line_width_alpha_kms(t, 0.9), (T,),
                                    eps=1e-5, rtol=1e-5, atol=0)
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


def test_partition_sum_matches_scipy_interp1d():
    """Q(T) from a table: the interval choice and extrapolation of
    scipy's interp1d(kind="linear", bounds_error=False,
    fill_value="extrapolate"), the call radlite.py:1152 makes, inside the
    table, on its nodes and past both ends."""
    from scipy.interpolate import interp1d
    Tt = np.array([10.0, 25.0, 50.0, 100.0, 200.0, 400.0, 1000.0, 3000.0])
    Qt = 0.3617 * Tt + 2.1e-5 * Tt ** 2 + 0.5
    T = np.concatenate([np.linspace(1.0, 5000.0, 397), Tt, [Tt[0] - 3.0, Tt[-1] + 7.0]])
    ref = interp1d(Tt, Qt, kind="linear", bounds_error=False, fill_value="extrapolate")(T)
    got = diff.partition_sum(torch.tensor(T), Tt, Qt).numpy()
    assert np.array_equal(got, ref) or np.abs(got / ref - 1).max() <= 2e-16
    # its derivative is the slope of the interval that forms Q
    Tg = torch.tensor([30.0, 150.0, 4000.0], dtype=torch.float64, requires_grad=True)
    assert torch.autograd.gradcheck(lambda t: diff.partition_sum(t, Tt, Qt), (Tg,), eps=1e-3,
                                    rtol=1e-8, atol=0)


def test_prepass_hand_values():
    """Special values fixed by the physics, with pyradlite's constants
    (k = 1.380649e-16, m_p = 1.67262192369e-24, G = 6.6743e-8)."""
    T = torch.tensor([100.0], dtype=torch.float64)
    # the width is the turbulence alone at T = 0 (sqrt of a square is exact)
    vt = torch.tensor([0.1, 0.37], dtype=torch.float64)
    assert torch.equal(diff.line_width_kms(torch.zeros(2, dtype=torch.float64), vt), vt)
    # thermal width of CO (m = 28 m_p) at 100 K: sqrt(2kT/m) = 0.24282 km/s;
    # lw^2 - v_t^2 is linear in T with slope 2k/(m m_p) = 5.896e-4 (km/s)^2/K
    assert abs(float(diff.line_width_kms(T, 0.0)) - 0.2428167997942543) <= 1e-15
    for Tk in (10.0, 300.0, 2500.0):
        lw = float(diff.line_width_kms(torch.tensor(Tk, dtype=torch.float64), 0.2))
        assert abs((lw * lw - 0.04) / Tk / 0.0005895999826232297 - 1) <= 1e-12
    # the alpha model (alpha 0.9, gamma 1.4, mu 2.3) at 100 K: (0.9 c_s)^2 +
    # the thermal width^2, c_s = 0.70883 km/s
    assert abs(float(diff.line_width_alpha_kms(T, 0.9)) - 0.6825963631337855) <= 1e-15
    # Keplerian: v sqrt(r) = sqrt(G M); the Sun's G M = 1.3271244e26 cm^3/s^2
    # gives 29.785 km/s at 1 au
    M = 1.3271244e26 / 6.6743e-8
    r = np.array([1.495978707e13, 4 * 1.495978707e13])
    v = diff.kepler_velocity(r, 3, M)
    assert np.all(v[..., :2] == 0.0)
    assert abs(v[0, 0, 2] / 1e5 - 29.78469182967693) <= 1e-12
    assert np.allclose(v[..., 2] * np.sqrt(r)[:, None], 11520088541326.408, rtol=1e-15, atol=0)
    # an upper level at E = 100 K, g = 3, at T = 100 K with Q(100) = 36 from
    # the table {50: 18.5, 150: 53.5}: n = 3 e^-1 / 36
    E_cm = np.array([0.0, 100.0 / synth.K_PER_CM])
    n = diff.lte_populations(T, E_cm, [1.0, 3.0], q_table=([50.0, 150.0], [18.5, 53.5]))
    assert abs(float(n[0, 1]) / 0.030656620097620196 - 1) <= 4e-16
    assert abs(float(n[0, 0]) / (1.0 / 36.0) - 1) <= 2e-16
