"""The Fortran host (radlite_amd/fortran/rl_host, ISO_C_BINDING module
radlite_gpu) on golden run directories: it reads the L4 files with
Fortran I/O, prepares the tables through the library's host pre-pass,
traces through librl.so and writes linespectrum_<mol>.dat.  Its file must
be byte-identical to the Python host's (same library, same formats) and
its spectra must match the reference's golden spectra to the 7 digits of
the E13.6 format."""
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np
import pytest

from cases import build_case, input_hash, load_golden
from radlite_amd import build, engine, l4io

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["t1", "t5", "t6", "t8", "t9", "c1"])
def test_fortran_host_writes_the_python_hosts_file(name):
    print(f"[gpu] fortran host {name}", file=sys.stderr, flush=True)
    assert os.path.exists(build.FORTRAN_HOST), "build the Fortran host first"
    g = load_golden(name)
    with tempfile.TemporaryDirectory() as d:
        fdir, pdir = os.path.join(d, "f"), os.path.join(d, "p")
        build_case(name, fdir)
        assert str(g["input_sha256"]) == input_hash(fdir)
        shutil.copytree(fdir, pdir)
        r = subprocess.run([build.FORTRAN_HOST, fdir, "fp64"], capture_output=True, text=True,
                           timeout=120)
        assert r.returncode == 0, r.stdout + r.stderr
        engine.run_directory(pdir, precision="fp64")
        ri = l4io.read_run_dir(pdir)
        fname = f"linespectrum_{ri.molname}.dat"
        ftext = open(os.path.join(fdir, fname)).read()
        ptext = open(os.path.join(pdir, fname)).read()
        if ftext != ptext:
            import difflib
            diff = list(difflib.unified_diff(ptext.splitlines(), ftext.splitlines(), "python",
                                             "fortran", lineterm="", n=0))
            raise AssertionError("\n".join(diff[:40]))
        blocks = l4io.read_line_spectrum_file(os.path.join(fdir, fname))
    for il, b in enumerate(blocks):
        ref = g["flux"][il][::-1]              # file rows run from the last channel
        assert np.allclose(b["flux"], ref, rtol=5.01e-6, atol=0), (name, il)   # E13.6
    # against the file the reference itself wrote for this run: identical
    # except where a flux within 1e-10 of an E13.6 rounding boundary
    # rounds the other way (the FP64 bound of the trace)
    refl = open(os.path.join(os.path.dirname(__file__), "golden", f"spec_{name}.dat")
                ).read().splitlines()
    got = ftext.splitlines()
    assert len(got) == len(refl)
    for a, b in zip(got, refl):
        if a != b:
            va, vb = [float(x) for x in a.split()], [float(x) for x in b.split()]
            # one unit of the sixth digit of the 0.dddddd mantissa
            assert va[0] == vb[0] and abs(va[1] - vb[1]) <= 1.01e-5 * abs(vb[1]), (a, b)


@pytest.mark.parametrize("name", ["p1", "p2", "p3"])
def test_fortran_host_writes_the_references_posvel_files(name):
    """A run directory asking for position-velocity cubes (linespectrum.inp
    command 2): the Fortran host's lineposvel files, written with the
    reference's own format statements from the library's FP64 cubes, are
    the files the reference wrote for the same run, byte for byte."""
    print(f"[gpu] fortran host posvel {name}", file=sys.stderr, flush=True)
    assert os.path.exists(build.FORTRAN_HOST), "build the Fortran host first"
    g = load_golden(name)
    with tempfile.TemporaryDirectory() as d:
        build_case(name, d)
        assert str(g["input_sha256"]) == input_hash(d)
        r = subprocess.run([build.FORTRAN_HOST, d, "fp64"], capture_output=True, text=True,
                           timeout=120)
        assert r.returncode == 0, r.stdout + r.stderr
        ri = l4io.read_run_dir(d)
        for il in range(len(g["nu0"])):
            iline = ri.iline_start + il
            got = open(os.path.join(d, l4io.posvel_file_name(ri.molname, iline))).read()
            ref = open(os.path.join(os.path.dirname(__file__), "golden",
                                    f"posvel_{name}_{ri.molname}_{iline}.dat")).read()
            assert got == ref, (name, iline)
