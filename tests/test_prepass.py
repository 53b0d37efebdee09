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
