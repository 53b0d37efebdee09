"""Spectrum files written by the compiled reference (oracle/_ref) for
golden cases, kept as expected-output fixtures: spec_<case>.dat is the
reference's own linespectrum_moldata.dat (header_line_spectrum /
write_line_spectrum, telescope.F:1662-1804) for that run directory.

Runs in the build container only (it needs the reference build):
    python tests/golden/make_spectrum_files.py t1 t5 t6 c1
"""
import os
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [HERE, ROOT]
from cases import build_case  # noqa: E402

PROBE = os.path.join(ROOT, "oracle", "_ref", "RADlite_probe")


def main(names):
    for name in names:
        with tempfile.TemporaryDirectory() as d:
            build_case(name, d)
            with open(os.path.join(d, "probe_rays.inp"), "w") as f:
                f.write("1\n1\n")
            r = subprocess.run([PROBE], cwd=d, capture_output=True, text=True, timeout=3600)
            if r.returncode != 0 or not os.path.exists(os.path.join(d, "radlite.success")):
                raise SystemExit(f"reference failed on {name}")
            shutil.copy(os.path.join(d, "linespectrum_moldata.dat"),
                        os.path.join(HERE, f"spec_{name}.dat"))
            print(name, "ok")


if __name__ == "__main__":
    main(sys.argv[1:] or ["t1", "t5", "t6", "c1"])
