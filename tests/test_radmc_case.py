"""Case r1: the reference's own tutorial RADMC inputs (dust density and
temperature, frequency grid, opacities, star, line.inp on a 120 x 100
grid) with pyradlite's gas-side files, rendered by the compiled reference.

* The oracle reproduces the reference's spectra, per-ray intensities,
  REAL*4 windows and integrated-channel masks (as for the other goldens,
  1e-12 relative; bit-exact in practice), from the inputs stored with the
  golden.
* Where the reference is mounted, reading the run directory made of its
  tutorial files gives exactly the stored inputs: the readers parse files
  the reference distributes, not only files this package writes.
"""
import tempfile

import numpy as np
import pytest

import pyorc
import radmc_case
from cases import load_golden
from radlite_amd import l4io

TOL = 1e-12


def test_oracle_matches_reference_on_tutorial_inputs():
    g = load_golden("r1")
    ri = radmc_case.radmc_inputs(g)
    orc = pyorc.Oracle(pyorc.Prepared(ri))
    assert orc.p.aksmax == float(g["aksmax"])
    cum_mask = np.zeros_like(g["cmask"][0], dtype=bool)
    for il in range(len(g["flux"])):
        o = orc.line_spectrum(il)
        rel = np.abs(o["flux"] - g["flux"][il]) / np.abs(g["flux"][il])
        assert rel.max() <= TOL, (il, rel.max())
        img = o["image"][g["image_rows"]]
        rel = np.abs(img - g["image"][il]) / np.maximum(np.abs(g["image"][il]), 1e-300)
        assert rel.max() <= TOL, (il, rel.max())
        np.testing.assert_array_equal(o["minvel"], g["minvel"][il])
        np.testing.assert_array_equal(o["maxvel"], g["maxvel"][il])
        cum_mask |= o["cmask"].astype(bool)
        np.testing.assert_array_equal(cum_mask, g["cmask"][il].astype(bool))
    np.testing.assert_array_equal(o["x0"], g["x0"])
    np.testing.assert_array_equal(o["ring_ri"], g["ring_ri"])


@pytest.mark.skipif(not radmc_case.available(), reason="the reference's tutorial files are not "
                    "mounted here")
def test_readers_parse_the_tutorial_files():
    g = load_golden("r1")
    with tempfile.TemporaryDirectory() as d:
        radmc_case.build_radmc_case("r1", d)
        ri = l4io.read_run_dir(d)
    stored = radmc_case.radmc_inputs(g)
    for k in radmc_case.ARRAYS + radmc_case.SCALARS:
        np.testing.assert_array_equal(np.asarray(getattr(ri, k)), np.asarray(getattr(stored, k)),
                                      err_msg=k)
    assert ri.molname == stored.molname
    # the tutorial grid and its RADMC headers as the reference reads them
    assert ri.rho_dust.shape == (1, 120, 100) and len(ri.cont_nu) == 130
    assert ri.fix_lw == 0.0 and ri.umass_av == 2.3
