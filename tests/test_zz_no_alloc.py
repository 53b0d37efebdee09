"""rl_trace_lines_forward_noalloc: the forward trace with every work buffer
taken from the library's fixed device workspace, spectra and images
written to the caller's device memory.  Checked against
rl_trace_lines_forward on the same items (the same kernels: bit-identical
spectra, images and warning bits), on two calls in a row, the second with
a smaller input (fewer items, fewer channels), so that what the first
call left in the workspace cannot leak into the second; and its refusals
(trajectories not resident, a request larger than the workspace)."""
import sys
import tempfile

import numpy as np
import pytest
import torch

from cases import build_case
from radlite_amd import engine, l4io

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def _run(name):
    with tempfile.TemporaryDirectory() as d:
        build_case(name, d)
        return l4io.read_run_dir(d)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_noalloc_matches_forward_two_calls(precision):
    ri = _run("c1")
    say(f"[gpu] noalloc forward c1 {precision}")
    rt = engine.Raytracer(0, precision)
    rt.load_run(ri)
    nfr, vmax = engine.passband(ri.vmax, ri.dv)
    rt.setup_rays([ri.incl_deg * 0.0174532925199], ri.nphi, ri.b_extra, ri.dbdr)
    calls = [
        [(0, 0, nfr, vmax, engine.SKIP_REFERENCE), (0, 0, nfr, vmax, engine.SKIP_NONE)],
        [(0, 0, nfr // 2, 0.5 * vmax, engine.SKIP_NONE)],
    ]
    # the allocating entry point first (it also builds the trajectories)
    refs = [rt.trace(items, want_image=True) for items in calls]
    nrays = int(rt.nrays[0])
    outs = []
    for items in calls:              # two no-allocation calls in a row
        nflux = sum(it[2] for it in items)
        fd = torch.full((nflux,), float("nan"), dtype=torch.float64, device="cuda")
        im = torch.full((nrays * nflux,), float("nan"), dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        warn = rt.trace_noalloc(items, fd.data_ptr(), im.data_ptr())
        outs.append((fd.cpu().numpy(), im.cpu().numpy(), warn))
    for items, (rfl, rim, rwarn), (fl, im, warn) in zip(calls, refs, outs):
        o = io = 0
        for k, it in enumerate(items):
            n = it[2]
            assert np.array_equal(fl[o:o + n], rfl[k]), (it, np.abs(fl[o:o + n] - rfl[k]).max())
            assert np.array_equal(im[io:io + nrays * n].reshape(nrays, n), rim[k]), it
            o += n
            io += nrays * n
        assert np.array_equal(warn, rwarn)
        say(f"    {len(items)} item(s), {sum(it[2] for it in items)} channels: bit-identical")
    # flux only (the image kept in the workspace)
    fd = torch.full((nfr // 2,), float("nan"), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    rt.trace_noalloc(calls[1], fd.data_ptr())
    assert np.array_equal(fd.cpu().numpy(), refs[1][0][0])
    rt.close()


def test_noalloc_refusals():
    ri = _run("t1")
    say("[gpu] noalloc refusals t1")
    rt = engine.Raytracer(0, "fp32")
    rt.load_run(ri)
    nfr, vmax = engine.passband(ri.vmax, ri.dv)
    rt.setup_rays([ri.incl_deg * 0.0174532925199], ri.nphi, ri.b_extra, ri.dbdr)
    item = (0, 0, nfr, vmax, engine.SKIP_NONE)
    fd = torch.zeros(64 * nfr, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    with pytest.raises(engine.RadliteError) as e:     # no trajectories yet
        rt.trace_noalloc([item], fd.data_ptr())
    assert e.value.code == 13 and "not resident" in str(e.value)
    ref, _, _ = rt.trace([item])
    # more channels than the workspace holds images and partial sums for
    big = (0, 0, 1 << 22, vmax, engine.SKIP_NONE)
    with pytest.raises(engine.RadliteError) as e:
        rt.trace_noalloc([big], fd.data_ptr())
    assert e.value.code == 13 and "workspace" in str(e.value)
    rt.trace_noalloc([item], fd.data_ptr())           # still usable
    assert np.array_equal(fd[:nfr].cpu().numpy(), ref[0])
    rt.close()
