"""Outputs beside the spectrum: the integrated-cell mask (imcir_cmask,
telescope.F:550-613) against the reference's golden masks, and the
circular image cube files (lineposvelcirc, telescope.F:1348-1373) written
by run_directory against the golden per-ray intensities."""
import os
import shutil
import sys
import tempfile

import numpy as np
import pytest

from cases import CASES, build_case, input_hash, load_golden
from radlite_amd import engine, l4io

pytestmark = pytest.mark.gpu


def _ri(name, d):
    g = load_golden(name)
    build_case(name, d)
    assert str(g["input_sha256"]) == input_hash(d)
    return g, l4io.read_run_dir(d)


@pytest.mark.parametrize("name", sorted(CASES))
def test_cmask_matches_reference(name):
    print(f"[gpu] cmask {name}", file=sys.stderr, flush=True)
    with tempfile.TemporaryDirectory() as d:
        g, ri = _ri(name, d)
    rt = engine.Raytracer(0, "fp64")
    rt.load_run(ri)
    nfr, vmax = engine.passband(ri.vmax, ri.dv)
    rt.setup_rays([ri.incl_deg * 0.0174532925199], ri.nphi, ri.b_extra, ri.dbdr)
    nl = len(g["flux"])
    fl, im, _, cm = rt.trace([(il, 0, nfr, vmax, engine.SKIP_REFERENCE) for il in range(nl)],
                             want_image=True, want_cmask=True)
    rt.close()
    cum = np.zeros_like(cm[0])
    for il in range(nl):
        cum |= cm[il]                     # the reference never clears the mask
        np.testing.assert_array_equal(cum.astype(bool), g["cmask"][il].astype(bool))


@pytest.mark.parametrize("name", ["t1", "t6"])
def test_image_cube_files(name):
    print(f"[gpu] image cube {name}", file=sys.stderr, flush=True)
    with tempfile.TemporaryDirectory() as d:
        g, ri = _ri(name, d)
        blocks, _ = engine.run_directory(d, precision="fp64", save_image=True)
        nl = len(g["flux"])
        for il in range(nl):
            c = l4io.read_image_cube(os.path.join(d, f"lineposvelcirc_{ri.molname}_{il + 1}.dat"))
            assert np.array_equal(g["image_rows"], np.arange(len(g["image_rows"])))
            img = g["image"][il]                      # [ray, channel], every ray
            nphi, nr = c["nphi"], c["nr"]
            assert c["nfr"] == img.shape[1] and nphi == ri.nphi
            assert np.allclose(c["ri"], g["ring_ri"][1:nr + 2], rtol=1e-5)
            assert np.allclose(c["center"], img[0], rtol=1e-14)
            for ch in range(c["nfr"]):
                ring = img[1:, ch].reshape(nr, nphi).T
                assert np.allclose(c["image"][ch], ring, rtol=5.01e-4, atol=0)   # E10.4
                cm = g["cmask"][il][1:, ch].reshape(nr, nphi).T
                np.testing.assert_array_equal(c["cmask"][ch], cm)
            assert np.allclose(c["velo"], g["dnu"][il] / c["nu0"] * 2.99792458e5, rtol=1e-14)
