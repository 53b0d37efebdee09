"""The oracle-spectra fixtures (tests/golden/c2_oracle.npz, c3_oracle.npz)
were made from the workload generators' output; the generators still
produce the same inputs (sha256 of every array), so the GPU tests that
compare against the fixtures compare like with like."""
import os

import numpy as np

from cases import inputs_digest
from radlite_amd import synth

G = os.path.join(os.path.dirname(__file__), "golden")


def test_c2_fixture_inputs_current():
    ri, _ = synth.workload_inputs("c2", seed=0)
    z = np.load(os.path.join(G, "c2_oracle.npz"))
    assert inputs_digest(ri) == str(z["inputs_sha256"])
    assert z["flux"].shape == (2, 2, 200)


def test_c3_fixture_inputs_current():
    ri, items, incl = synth.c3_inputs(0)
    z = np.load(os.path.join(G, "c3_oracle.npz"))
    assert inputs_digest(ri) == str(z["inputs_sha256"])
    by = {(a, b): (n, v) for a, b, n, v in items}
    for (a, b, n), v in zip(z["items"], z["vmax"]):
        assert by[(int(a), int(b))] == (int(n), float(v))
