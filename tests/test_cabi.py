"""The C-ABI library without a GPU: it loads, exports every function the
public header declares, rejects calls before a device context exists, and
its host pre-pass (line frequencies, grey dust, star, aksmax, passband)
reproduces the oracle's -- i.e. the reference's -- arithmetic exactly."""
import ctypes as C
import os
import re
import tempfile

import numpy as np
import pytest

import pyorc
from cases import build_case
from radlite_amd import engine, l4io

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    with open(os.path.join(ROOT, "include", "radlite.h")) as f:
        text = f.read()
    return sorted(set(re.findall(r"\b(rl_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = engine.load_library()
    names = declared_functions()
    assert len(names) >= 20
    for n in names:
        assert hasattr(L, n), n


def test_create_without_device_fails_cleanly():
    L = engine.load_library()
    ctx = C.c_void_p()
    rc = L.rl_create(C.byref(ctx), 0, engine.RL_FP64)
    if rc == 0:              # a GPU is visible: this is not the CPU-only case
        L.rl_destroy(ctx)
        pytest.skip("GPU present")
    assert rc == 900 and not ctx.value
    assert L.rl_create(C.byref(ctx), 0, 7) == 13          # bad precision


@pytest.mark.parametrize("name", ["t2", "t6", "t8", "c1"])
def test_host_prepass_matches_oracle(name):
    with tempfile.TemporaryDirectory() as d:
        build_case(name, d)
        ri = l4io.read_run_dir(d)
    tb = engine.prepare_tables(ri)
    p = pyorc.Prepared(ri)
    np.testing.assert_array_equal(tb.nu0, p.nu0)
    np.testing.assert_array_equal(tb.jd, p.jd)
    np.testing.assert_array_equal(tb.ad, p.ad)
    np.testing.assert_array_equal(tb.istar, p.istar)
    assert tb.aksmax == p.aksmax
    nfr, vmax = engine.passband(ri.vmax, ri.dv)
    assert nfr == p.nfr and vmax == p.vmax


def test_passband_rule():
    # main.F:209-210: nfr = int(2 vmax/dv + 1), vmax := (nfr-1) dv / 2
    assert engine.passband(30.0, 60.0 / 49 * (1 - 1e-9))[0] == 50
    assert engine.passband(30.0, 60.0 / 49 * (1 + 1e-9))[0] == 49
    with pytest.raises(engine.RadliteError):
        engine.passband(1.0, 5.0)                             # nfr = 1
