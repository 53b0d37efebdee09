"""Adjoint of the raytrace (rl_trace_lines_adjoint) on the GPU.

* The adjoint's own forward pass reproduces the forward spectrum.
* FP64 gradients of a linear functional of the spectrum with respect to
  every input the trace reads (populations, abundance, density, width,
  dust j and alpha) against central differences of the CPU oracle on
  small golden cases.  The tolerance is relative to the gradient plus
  the finite-difference rounding floor of the case.
* The chi^2 fit of the 200x100 disk (100 channels, 45 deg, 34,049 rays)
  with gradients with respect to per-cell temperature, log-abundance and
  turbulence through the torch pre-pass: FP64 adjoint against central
  differences (step 1e-4 relative) of the FP64 forward on 50 random
  cells, and the FP32 adjoint against the FP64 adjoint.
"""
import sys
import tempfile

import numpy as np
import pytest
import torch

import pyorc
from cases import build_case, input_hash, load_golden
from radlite_amd import diff, engine, l4io, synth

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, file=sys.stderr, flush=True)


_CACHE = {}


def _case(name):
    if name not in _CACHE:
        g = load_golden(name)
        with tempfile.TemporaryDirectory() as d:
            build_case(name, d)
            assert str(g["input_sha256"]) == input_hash(d)
            ri = l4io.read_run_dir(d)
        _CACHE[name] = ri
    return _CACHE[name]


def _engine(ri, precision):
    rt = engine.Raytracer(0, precision)
    rt.load_run(ri)
    nfr, vmax = engine.passband(ri.vmax, ri.dv)
    rt.setup_rays([ri.incl_deg * 0.0174532925199], ri.nphi, ri.b_extra, ri.dbdr)
    return rt, nfr, vmax


@pytest.mark.parametrize("name", ["t1", "t5", "t6"])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_adjoint_forward_pass_matches_forward(name, precision):
    ri = _case(name)
    say(f"[gpu] adjoint forward pass {name} {precision}")
    rt, nfr, vmax = _engine(ri, precision)
    items = [(il, 0, nfr, vmax, engine.SKIP_NONE) for il in range(len(ri.A))]
    fl, _, _ = rt.trace(items)
    fa, gr = rt.adjoint(items, [np.zeros(nfr)] * len(items))
    for a, b in zip(fa, fl):
        rel = np.abs(a - b) / np.abs(b)
        # FP64: same quantity through the scaled-record formulation; FP32: same code
        assert rel.max() <= (1e-12 if precision == "fp64" else 1e-6), rel.max()
    for k, v in gr.items():
        assert np.all(v == 0.0), k
    rt.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_tape_and_checkpoint_intervals(precision):
    """rl_trace_lines_forward_tape + rl_trace_lines_adjoint (reverse sweep
    only) against the adjoint that runs its own forward; the tape goes
    stale when an input is set; every checkpoint interval (I recomputed
    from checkpoints every 2..16 segments) against every segment kept."""
    ri = _case("t6")
    say(f"[gpu] tape and checkpoint intervals t6 {precision}")
    rt, nfr, vmax = _engine(ri, precision)
    items = [(il, 0, nfr, vmax, engine.SKIP_NONE) for il in range(len(ri.A))]
    rng = np.random.default_rng(7)
    gf = [rng.normal(size=nfr) * 1e20 for _ in items]
    fa, ga = rt.adjoint(items, gf)
    assert rt.adjoint_info() == (1, 0, 1)         # small case: every segment kept, one chunk
    ft = rt.trace_tape(items)
    fb, gb = rt.adjoint(items, gf)
    assert rt.adjoint_info() == (1, 1, 1)
    for a, b, c in zip(fa, fb, ft):
        assert np.array_equal(a, b) and np.array_equal(a, c)
    for k in ga:
        assert np.array_equal(ga[k], gb[k]), k
    rt.set_levels(ri.pops)                        # same values, but an input was set
    fc, gc = rt.adjoint(items, gf)
    assert rt.adjoint_info()[1] == 0
    for k in ga:
        assert np.array_equal(ga[k], gc[k]), k
    rt.tape_release()
    for K in (2, 4, 8, 16):
        rt.set_option(engine.OPT_CKPT, K)
        fk, gk = rt.adjoint(items, gf)
        assert rt.adjoint_info() == (K, 0, 1)
        for a, b in zip(fa, fk):
            assert np.array_equal(a, b)
        worst = 0.0
        for k in ga:
            d = float(np.abs(gk[k] - ga[k]).max())
            s = float(np.abs(ga[k]).max())
            worst = max(worst, d / s if s > 0 else d)
        say(f"    K={K}: max |g_K - g_1| / max |g_1| = {worst:.2e}")
        # the recomputed I follows the same arithmetic as the stored one
        assert worst <= (1e-14 if precision == "fp64" else 1e-6), (K, worst)
        ft = rt.trace_tape(items)
        fk2, gk2 = rt.adjoint(items, gf)
        assert rt.adjoint_info() == (K, 1, 1)
        for k in ga:
            assert np.array_equal(gk[k], gk2[k]), (K, k)
        rt.tape_release()
    with pytest.raises(engine.RadliteError):
        rt.set_option(engine.OPT_CKPT, 3)
    rt.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_adjoint_repeats_bit_for_bit_many_channels(precision):
    """300 channels (five groups of 64 per ray in the reverse sweep): two
    adjoint calls and a tape + reverse call give bit-identical gradients
    (the channel groups of a ray add their node sums in a fixed order)."""
    ri = _case("t6")
    say(f"[gpu] adjoint bit-repeat t6 nfr=300 {precision}")
    rt, _, vmax = _engine(ri, precision)
    nfr = 300
    items = [(il, 0, nfr, vmax, engine.SKIP_NONE) for il in range(len(ri.A))]
    rng = np.random.default_rng(9)
    gf = [rng.normal(size=nfr) * 1e20 for _ in items]
    fa, ga = rt.adjoint(items, gf)
    fb, gb = rt.adjoint(items, gf)
    rt.trace_tape(items)
    fc, gc = rt.adjoint(items, gf)
    assert rt.adjoint_info()[1] == 1
    for a, b, c in zip(fa, fb, fc):
        assert np.array_equal(a, b) and np.array_equal(a, c)
    nonzero = 0
    for k in ga:
        assert np.array_equal(ga[k], gb[k]), k
        assert np.array_equal(ga[k], gc[k]), k
        nonzero += int(np.count_nonzero(ga[k]))
    assert nonzero > 0
    rt.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_adjoint_ray_chunks(precision):
    """An item whose buffers do not fit the memory budget is traced and
    reversed a range of rays at a time (the largest checkpoint interval,
    2..64 chunks, each with its own stencil transpose): the spectrum is
    bit-identical and the cell gradients agree with the whole item at the
    same interval to the rounding of the per-chunk cell sums; the tape
    refuses an item that does not fit."""
    ri = _case("t6")
    say(f"[gpu] adjoint ray chunks t6 {precision}")
    rt, nfr, vmax = _engine(ri, precision)
    items = [(il, 0, nfr, vmax, engine.SKIP_NONE) for il in range(len(ri.A))]
    rng = np.random.default_rng(8)
    gf = [rng.normal(size=nfr) * 1e20 for _ in items]
    rt.set_option(engine.OPT_CKPT, 16)
    fa, ga = rt.adjoint(items, gf)
    assert rt.adjoint_info() == (16, 0, 1)
    mi = rt.memory_info()                         # the whole item's buffers
    rt.set_option(engine.OPT_CKPT, 0)
    # no room at all (the most chunks), and room for about a third of the item
    for budget, lo, hi in ((1, 32, 64), (mi["traj"] + 0.3 * mi["work"], 2, 63)):
        rt.set_memory_budget(budget)
        fc, gc = rt.adjoint(items, gf)
        K, tape, nch = rt.adjoint_info()
        say(f"    budget {budget:.3g}: K={K}, {nch} chunks")
        assert K == 16 and tape == 0 and lo <= nch <= hi, (K, tape, nch)
        for a, b in zip(fa, fc):
            assert np.array_equal(a, b)
        for k in ga:
            scale = np.abs(ga[k]).max()
            d = np.abs(gc[k] - ga[k]).max()
            assert d <= 1e-12 * scale, (k, d, scale)
        with pytest.raises(engine.RadliteError) as e:
            rt.trace_tape(items)
        assert e.value.code == 13
    rt.set_memory_budget(0)
    rt.close()


def test_forward_evicts_trajectories_to_fit():
    """Two inclinations under a budget that holds one inclination's
    trajectories and a batch: the second batch evicts the first
    inclination's nodes (rebuilt when needed again); the spectra are
    bit-identical to the unconstrained run."""
    ri = _case("t6")
    say("[gpu] trajectory eviction t6")
    rt = engine.Raytracer(0, "fp32")
    rt.load_run(ri)
    nfr, vmax = engine.passband(ri.vmax, ri.dv)
    rt.setup_rays([0.6, 1.1], ri.nphi, ri.b_extra, ri.dbdr)
    items = [(0, 0, nfr, vmax, engine.SKIP_NONE), (0, 1, nfr, vmax, engine.SKIP_NONE),
             (0, 0, nfr, vmax, engine.SKIP_REFERENCE)]
    ref, _, _ = rt.trace(items)
    both = rt.memory_info()["traj"]
    work = []
    for it in items[:2]:                          # one item with its inclination's nodes
        rt.trace([it])
        work.append(rt.memory_info()["work"])
    rt.set_memory_budget(1.02 * max(work))
    f, _, _ = rt.trace(items)
    mi = rt.memory_info()
    say(f"    both {both:.3g} B, budget {mi['budget']:.3g} B, resident {mi['traj']:.3g} B, "
        f"batches {rt.batch_info()[0]}")
    assert rt.batch_info()[0] == 3
    assert mi["traj"] < both
    for a, b in zip(ref, f):
        assert np.array_equal(a, b)
    rt.close()


def test_adjoint_rejects_channel_skipping():
    ri = _case("t1")
    rt, nfr, vmax = _engine(ri, "fp64")
    with pytest.raises(engine.RadliteError) as e:
        rt.adjoint([(0, 0, nfr, vmax, engine.SKIP_REFERENCE)], [np.ones(nfr)])
    assert e.value.code == 13
    rt.close()


def _oracle_loss(orc, prep, weights, nlines):
    """L = sum_lines w . F and sum |w F| (the scale of L's rounding)."""
    tot, scale = 0.0, 0.0
    for il in range(nlines):
        o = orc.line_spectrum(il, skip_mode=1, want_image=False)
        tot += float(np.dot(weights[il], o["flux"]))
        scale += float(np.abs(weights[il] * o["flux"]).sum())
    return tot, scale


@pytest.mark.parametrize("name", ["t3", "t5", "t6"])
def test_fp64_adjoint_matches_oracle_finite_differences(name):
    """L = sum_lines w . F; dL/dx by the adjoint vs (L(x+h) - L(x-h))/2h of
    the CPU oracle, for the cells with the largest gradients and random
    cells among those with non-negligible ones, every input kind."""
    ri = _case(name)
    say(f"[gpu] adjoint vs oracle FD {name}")
    rt, nfr, vmax = _engine(ri, "fp64")
    nl = len(ri.A)
    rng = np.random.default_rng(11)
    w = [rng.normal(size=nfr) for _ in range(nl)]
    items = [(il, 0, nfr, vmax, engine.SKIP_NONE) for il in range(nl)]
    fa, gr = rt.adjoint(items, w)
    rt.close()
    prep = pyorc.Prepared(ri)
    orc = pyorc.Oracle(prep)
    L0, S0 = _oracle_loss(orc, prep, w, nl)
    lev_u, lev_l = int(prep.lev_up[0]), int(prep.lev_low[0])
    # (name, gradient array, oracle array, index function)
    fields = [
        ("pops_up", gr["g_pops"][..., lev_u], prep.pops, lambda c: c + (lev_u,)),
        ("pops_low", gr["g_pops"][..., lev_l], prep.pops, lambda c: c + (lev_l,)),
        ("abund", gr["g_abund"], prep.abund, lambda c: c),
        ("rho", gr["g_rho"], prep.rho, lambda c: c),
        ("lw", gr["g_lw"], prep.lw, lambda c: c),
        ("jd", gr["g_jd"][..., 0], prep.jd, lambda c: c + (0,)),
        ("ad", gr["g_ad"][..., 0], prep.ad, lambda c: c + (0,)),
    ]
    # central differences at h = 1e-4 relative (truncation ~h^2, rounding
    # ~ eps sum|w F| / (h x)); the forward is piecewise smooth (the theomax
    # bracket and the dtau thresholds of qdr_src_2 are kinks), so a cell that
    # misses at 1e-4 is checked again at 1e-5: a kink inside the wider
    # stencil drops out, a wrong gradient does not
    worst = 0.0
    checked = 0
    for fname, g, arr, idx in fields:
        # cells ranked by the sensitivity to a relative change, |x dL/dx|
        xv = np.array([arr[idx(tuple(int(v) for v in c))] for c in np.ndindex(g.shape)])
        mag = np.abs(g.ravel() * xv).reshape(g.shape)
        gmax = mag.max()
        assert gmax > 0, fname
        order = np.argsort(mag.ravel())[::-1]
        top = list(order[:4])
        sig = np.flatnonzero(mag.ravel() > 1e-3 * gmax)
        rnd = list(rng.choice(sig, size=min(4, len(sig)), replace=False))
        for flat in dict.fromkeys(top + rnd):
            c = np.unravel_index(flat, g.shape)
            ix = idx(tuple(int(v) for v in c))
            x0 = arr[ix]
            if x0 == 0.0:
                continue
            for h in (1e-4, 1e-5):
                dx = h * abs(x0)
                arr[ix] = x0 + dx
                Lp = _oracle_loss(orc, prep, w, nl)[0]
                arr[ix] = x0 - dx
                Lm = _oracle_loss(orc, prep, w, nl)[0]
                arr[ix] = x0
                fd = (Lp - Lm) / (2 * dx)
                # the FP64 flux carries ~1e-14 relative rounding (recurrences
                # along the ray), measured on these cases as FD noise
                floor = 1e-13 * S0 / dx
                # |g - fd| <= 1e-5 |fd| + floor
                ok = abs(g[c] - fd) <= 1e-5 * abs(fd) + floor
                if ok:
                    break
            worst = max(worst, abs(g[c] - fd) / abs(fd))
            checked += 1
            assert ok, (name, fname, c, g[c], fd, floor)
    say(f"    {name}: {checked} FD checks, worst rel err {worst:.2e}")


def _c4_problem(precision):
    ri, w = synth.workload_inputs("c4", seed=0)
    model = diff.SpectrumModel(ri, precision)
    T0 = torch.tensor(ri.t_dust[0], dtype=torch.float64)
    la0 = torch.log(torch.tensor(ri.abund, dtype=torch.float64))
    vt0 = torch.full_like(T0, 0.1)
    return ri, model, T0, la0, vt0


def _chi2(model, T, la, vt, Ft, sigma, with_scale=False):
    F = model.flux_from_physics(T, la, vt)
    r = (F - Ft) / sigma
    L = (r ** 2).sum()
    if with_scale:     # rounding scale of chi^2: sum 2|r| |F|/sigma eps_F
        return L, float((2 * r.abs() * F.abs()).sum() / sigma)
    return L


def test_c4_chi2_gradients_fd_and_fp32():
    """The 200x100 disk: target from T x lognormal(0.2); gradients of
    chi^2 with respect to T, log-abundance and turbulence."""
    say("[gpu] c4 chi^2 adjoint")
    ri, m64, T0, la0, vt0 = _c4_problem("fp64")
    rng = np.random.default_rng(4)
    Tt = T0 * torch.from_numpy(np.exp(rng.normal(0.0, 0.2, size=T0.shape)))
    with torch.no_grad():
        Ft = m64.flux_from_physics(Tt, la0, vt0)
    sigma = 1e-3 * float(Ft.abs().max())
    T = T0.clone().requires_grad_(True)
    la = la0.clone().requires_grad_(True)
    vt = vt0.clone().requires_grad_(True)
    L, S0 = _chi2(m64, T, la, vt, Ft, sigma, with_scale=True)
    L.backward()
    assert m64.rt.adjoint_info()[1] == 1          # the backward ran from the forward's tape
    g64 = {"T": T.grad.numpy().copy(), "la": la.grad.numpy().copy(), "vt": vt.grad.numpy().copy()}
    L0 = float(L)
    say(f"    chi2 = {L0:.6e}")
    # central differences on 50 random cells, step 1e-4 relative (additive
    # for log-abundance); FD rounding floor 1e-13 of the chi^2 scale
    cells = rng.choice(T0.numel(), size=50, replace=False)
    nth = T0.shape[1]
    fails, worst_sig = [], 0.0
    nchk = retried = 0
    with torch.no_grad():
        for flat in cells:
            c = (int(flat) // nth, int(flat) % nth)
            for key, base in (("T", T0), ("la", la0), ("vt", vt0)):
                x0 = float(base[c])
                ga = float(g64[key][c])
                # a step that straddles a kink of the piecewise-smooth forward
                # (sub-grid point counts, the theomax bracket) is re-checked
                # with a 10x smaller one
                for rel in (1e-4, 1e-5):
                    dx = rel * abs(x0) if key != "la" else rel
                    args = {"T": T0.clone(), "la": la0.clone(), "vt": vt0.clone()}
                    args[key][c] = x0 + dx
                    Lp = float(_chi2(m64, args["T"], args["la"], args["vt"], Ft, sigma))
                    args[key][c] = x0 - dx
                    Lm = float(_chi2(m64, args["T"], args["la"], args["vt"], Ft, sigma))
                    fd = (Lp - Lm) / (2 * dx)
                    floor = 1e-13 * S0 / dx      # FD rounding floor of chi^2 over the step
                    # |g - fd| <= 1e-4 |fd| + floor (step 1e-4: truncation ~1e-6 relative)
                    ok = abs(ga - fd) <= 1e-4 * abs(fd) + floor
                    if ok:
                        break
                    retried += rel == 1e-4
                nchk += 1
                if abs(fd) > 100 * floor:
                    worst_sig = max(worst_sig, abs(ga - fd) / abs(fd))
                if not ok:
                    fails.append((key, c, ga, fd, floor))
    say(f"    {nchk} FD checks ({retried} re-checked at 1e-5), {len(fails)} beyond tolerance, "
        f"worst rel err (well above the floor) {worst_sig:.2e}")
    for f in fails:
        say("      ", f)
    assert not fails
    m64.close()
    # FP32 adjoint against the FP64 adjoint
    _, m32, _, _, _ = _c4_problem("fp32")
    T = T0.clone().requires_grad_(True)
    la = la0.clone().requires_grad_(True)
    vt = vt0.clone().requires_grad_(True)
    L32 = _chi2(m32, T, la, vt, Ft, sigma)
    L32.backward()
    m32.close()
    say(f"    fp32 chi2 = {float(L32):.6e}")
    for key, g in (("T", T.grad), ("la", la.grad), ("vt", vt.grad)):
        a, b = g.numpy(), g64[key]
        rel = np.linalg.norm(a - b) / np.linalg.norm(b)
        big = np.abs(b) > 1e-2 * np.abs(b).max()
        cell = (np.abs(a - b)[big] / np.abs(b)[big]).max()
        say(f"    fp32 vs fp64 adjoint d/d{key}: norm rel {rel:.2e}, max over large cells {cell:.2e}")
        # measured: d/dT 2.5e-8 / 4.5e-7, d/dla 4.0e-6 / 7.3e-5, d/dvt 3.6e-6 / 2.9e-4
        assert rel <= 1e-3 and cell <= 1e-3, key
