"""FP64 spectra of the large workloads computed by the oracle, kept as
fixtures so that the GPU's FP64 mode is held to 1e-10 per channel on the
whole of them (the oracle is pinned bit-exact against the compiled
reference on every golden case; the reference build itself cannot hold
these grids: 300 radii at most, configure.h FRSIZE_X).

* c2_oracle.npz: the 400x200 disk (133,633 rays, 200 channels) at 30 and
  60 degrees, every channel integrated and with the reference's
  NONREDUNDANT skipping: flux[incl][mode][channel].
* c3_oracle.npz: 16 items of the 64-line x 8-inclination workload (lines
  0-7 at the first and last inclinations, their own channel counts),
  every channel integrated: flux concatenated in item order.

Runs in the build container (CPU, a few minutes on 8 cores):
    python tests/golden/make_oracle_spectra.py
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, "oracle"), HERE]

import pyorc  # noqa: E402
from cases import inputs_digest  # noqa: E402
from radlite_amd import engine, synth  # noqa: E402

C3_ITEMS = [(il, ii) for ii in (0, 7) for il in range(8)]


def c2():
    ri, w = synth.workload_inputs("c2", seed=0)
    nfr, vmax = engine.passband(ri.vmax, ri.dv)
    orc = pyorc.Oracle(pyorc.Prepared(ri))
    flux = np.zeros((2, 2, nfr))
    for i, inc in enumerate(w["incl"]):
        for mode in (1, 0):               # every channel; NONREDUNDANT skipping
            t0 = time.time()
            o = orc.line_spectrum(0, incl_rad=inc * 0.0174532925199, skip_mode=mode,
                                  want_image=False)
            flux[i, mode] = o["flux"]
            print(f"c2 incl {inc} mode {mode}: {time.time() - t0:.1f} s", flush=True)
    np.savez_compressed(os.path.join(HERE, "c2_oracle.npz"), flux=flux,
                        incl_deg=np.array(w["incl"]), nfr=np.array(nfr), vmax=np.array(vmax),
                        inputs_sha256=np.array(inputs_digest(ri)))


def c3():
    ri, items, incl = synth.c3_inputs(0)
    orc = pyorc.Oracle(pyorc.Prepared(ri))
    fl, meta = [], []
    bykey = {(it[0], it[1]): it for it in items}
    for il, ii in C3_ITEMS:
        _, _, n, v = bykey[(il, ii)]
        t0 = time.time()
        o = orc.line_spectrum(il, incl_rad=incl[ii] * 0.0174532925199, nfr=n, vmax=v,
                              skip_mode=1, want_image=False)
        fl.append(o["flux"])
        meta.append((il, ii, n, v))
        print(f"c3 line {il} incl {incl[ii]} nfr {n}: {time.time() - t0:.1f} s", flush=True)
    np.savez_compressed(os.path.join(HERE, "c3_oracle.npz"), flux=np.concatenate(fl),
                        items=np.array([m[:3] for m in meta]), vmax=np.array([m[3] for m in meta]),
                        inputs_sha256=np.array(inputs_digest(ri)))


if __name__ == "__main__":
    which = sys.argv[1:] or ["c2", "c3"]
    if "c3" in which:
        c3()
    if "c2" in which:
        c2()
