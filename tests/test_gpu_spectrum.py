"""Spectrum synthesis on the GPU and the fit chain through it.

* synthesize() on cuda:0 against the same call on the CPU (the oracle
  comparison itself runs on the CPU, tests/test_spectrum.py).
* The whole chain of a fit to an observed spectrum: the trace of the
  3-line golden run (FP64) -> the synthesis to a convolved, resampled
  spectrum -> chi^2 -> the trace adjoint; the gradient with respect to
  the per-cell temperature (device pre-pass) against a central finite
  difference of the same chain along a random direction.
"""
import sys
import tempfile

import numpy as np
import pytest
import torch

from cases import build_case
from radlite_amd import diff, l4io, spectrum

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def _lines(seed, L=40):
    rng = np.random.default_rng(seed)
    vel, flux, freq = [], [], []
    for _ in range(L):
        nfr = int(rng.integers(20, 200))
        v = np.linspace(-30.0, 30.0, nfr)
        vel.append(v)
        flux.append(1e-20 + 4e-21 * np.exp(-0.5 * (v / rng.uniform(2, 8)) ** 2))
        freq.append(spectrum.C0 / (4.6e-4 + 2e-5 * rng.uniform()))
    return vel, flux, freq


def test_synthesis_on_gpu_matches_cpu():
    say("[gpu] spectrum synthesis, 40 lines, cuda vs cpu")
    vel, flux, freq = _lines(0)
    for vs in (0.8, 0.55, 1.3, 0.35, 2.1):
        try:
            c = spectrum.synthesize(vel, flux, freq, 140.0, 3.0, vs, device="cpu")
            break
        except ValueError:
            continue
    g = spectrum.synthesize(vel, flux, freq, 140.0, 3.0, vs, device="cuda:0")
    assert g["spectrum"].device.type == "cuda"
    for k in ("spectrum", "emission", "continuum"):
        a, b = g[k].cpu().numpy(), c[k].numpy()
        e = float(np.abs(a - b).max() / np.abs(b).max())
        say(f"    {k}: {e:.1e}")
        assert e <= 1e-13
    assert np.array_equal(g["wavelength"].cpu().numpy(), c["wavelength"].numpy())


def test_fit_chain_gradient_through_synthesis_and_adjoint():
    say("[gpu] trace -> synthesis -> chi^2 -> adjoint (t6, fp64)")
    d = tempfile.mkdtemp()
    build_case("t6", d)
    ri = l4io.read_run_dir(d)
    m = diff.SpectrumModel(ri, "fp64")
    nu0 = list(m.tables.nu0)
    T0 = torch.tensor(ri.t_dust[0], dtype=torch.float64)
    la0 = torch.log(torch.tensor(ri.abund, dtype=torch.float64))
    vt0 = torch.full_like(T0, 0.1)
    kw = None

    def chain(T):
        nonlocal kw
        F = m.flux_from_physics_device(T, la0, vt0)
        if kw is None:
            errs = []
            for vs in (3.1, 4.4, 5.3, 6.7, 2.9, 8.2, 3.7, 4.9, 1.0, 0.7):
                try:
                    spectrum.synthesize_items(F.detach(), m.items, nu0, 140.0, 4.0, vs,
                                              device="cuda:0")
                    kw = dict(dist=140.0, obsres=4.0, vsampling=vs, device="cuda:0")
                    break
                except ValueError as e:          # user grid past the model grid
                    errs.append(f"{vs}: {e}")
            assert kw is not None, errs
        return spectrum.synthesize_items(F, m.items, nu0, **kw)

    with torch.no_grad():
        s0 = chain(T0 * 1.05)
    obs = s0["spectrum"].detach()
    sig = 1e-3 * float(obs.abs().max())

    def loss(T):
        s = chain(T)
        return (((s["spectrum"] - obs) / sig) ** 2).sum()

    T = T0.clone().requires_grad_(True)
    L = loss(T)
    L.backward()
    g = T.grad.numpy()
    assert np.isfinite(g).all() and np.abs(g).max() > 0
    rng = np.random.default_rng(2)
    dirn = torch.from_numpy(rng.normal(size=T0.shape)) * T0 * 1e-3
    h = 1e-2
    with torch.no_grad():
        lp = float(loss(T0 + h * dirn))
        lm = float(loss(T0 - h * dirn))
    fd = (lp - lm) / (2 * h)
    an = float((T.grad * dirn).sum())
    say(f"    directional derivative: adjoint {an:.6e}, finite difference {fd:.6e}, "
        f"rel {abs(an - fd) / abs(fd):.1e}")
    assert abs(an - fd) <= 1e-5 * abs(fd)
    m.close()
