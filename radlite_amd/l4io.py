"""RADLite input/output files ("L4" formats): writers and readers.

These keep RADLite's run-directory contract, so the same directory can
be rendered by the reference ``RADlite`` executable or by this package.
Every reader states the reference routine whose semantics it follows;
the writers emit full-precision numbers wherever the reference reads
list-directed, and the reference's fixed formats where it reads
formatted (moldata.dat), so both sides see identical doubles.

Array convention used throughout the package: cell fields are numpy
arrays of shape ``[nr, nth_half]`` (theta index fastest), the order in
which the reference reads them (``setup.F:778-785``, ``line.F:915-919``).
"""
from __future__ import annotations

import os
from dataclasses import dataclass, field

import numpy as np

F17 = "{:.17e}"


def _fmt(x: float) -> str:
    return F17.format(float(x))


# ----------------------------------------------------------------------
# writers
# ----------------------------------------------------------------------
def write_radlite_inp(path, nphi, dbdr, b_extra, outer_bc=0, incl_deg=45.0):
    """radlite.inp, format 104 (main.F:535-597; setup.F:1117-1137).

    ``strip_comments`` (tools.F:6-19) keeps only columns 1-16, so every
    value is left-justified in a 16-column field.
    """
    rows = [
        "104", "=" * 40, "-1", "2", "0", "1.000e-06", "0", "1", "=" * 40,
        "41", "16", "2", "0.1", "1.0", "3", "=" * 40,
        str(outer_bc), "2", "0", "=" * 40,
        repr(float(incl_deg))[:15], str(nphi), str(dbdr), str(b_extra), "=" * 40,
        "1", "0", "-1", "0", "-1", "0", "0", "0", "=" * 40,
        "-551", "0", "1", "1", "1", "=" * 40,
    ]
    with open(os.path.join(path, "radlite.inp"), "w") as f:
        for r in rows:
            if r.startswith("="):
                f.write(r + "\n")
            else:
                assert len(r) <= 15, r
                f.write(f"{r:<16s}\n")


def write_linespectrum_inp(path, vmax, dv, incl_deg, nlines, ilinestart=1,
                           molfile="moldata.dat", vrad=0.0, posvel=None):
    """linespectrum.inp, object format 2 (telescope.F:104-238): a line
    spectrum, or with ``posvel`` (dict nx, ny, size_x, size_y [cm, the whole
    image], phioff [rad], xoff, yoff [cm], addstar) a position-velocity cube
    (linespec_command 2, telescope.F:196-224)."""
    with open(os.path.join(path, "linespectrum.inp"), "w") as f:
        f.write("1\n1\n" + "-" * 63 + "\n2\n" + "-" * 63 + "\n")
        f.write(_fmt(vmax) + "\n" + _fmt(dv) + "\n" + "-" * 65 + "\n")
        f.write("2\n" + molfile + "\n" + ("2" if posvel else "0") + "\n1.0\n")
        f.write(_fmt(incl_deg) + "\n" + _fmt(vrad) + "\n")
        f.write(f"{nlines}\n{ilinestart}\n")
        if posvel:
            pv = posvel
            f.write(f"{int(pv['nx'])}\n{int(pv['ny'])}\n1\n")
            for k in ("size_x", "size_y", "phioff", "xoff", "yoff"):
                f.write(_fmt(pv[k]) + "\n")
            f.write(f"{int(pv['addstar'])}\n")


def write_line_inp(path, nlines, umass_av=2.3, nfreq=40, width=80.0, fix_lw=0.0):
    """line.inp, format 2 (line.F:168-279). The telescope passband
    overrides nfreq/width; ``fix_lw`` is the fallback width that
    line.F:3020-3033 uses for aksmax where the local width is <= 0."""
    with open(os.path.join(path, "line.inp"), "w") as f:
        f.write(f"2\n1\n{umass_av!r}\n" + "=" * 60 + "\n")
        f.write("-" * 60 + "\n1\n0\n" + f"{nfreq}\n{_fmt(width)}\n1\n")
        if fix_lw > 0:
            f.write("1\n" + _fmt(fix_lw) + "\n")
        f.write("0\n")
        for _ in range(1, nlines):
            f.write("-" * 60 + "\n0\n")


def quantize_moldata(E_cm, g, A):
    """Values as the reference sees them after its formatted reads:
    energies F12.4, weights F7.1, Einstein A E12.3 (line.F:1902,1934)."""
    Eq = np.array([float(f"{e:12.4f}") for e in E_cm])
    gq = np.array([float(f"{x:7.1f}") for x in g])
    Aq = np.array([float(f"{a:12.3e}") for a in A])
    return Eq, gq, Aq


def write_moldata(path, name, weight, E_cm, g, lev_up, lev_low, A,
                  fname="moldata.dat"):
    """LAMDA-like molecule file read by read_species_lambda
    (line.F:1865-1935). Levels 1-based; lev_up > lev_low required
    (line.F:1756-1762)."""
    E_cm = np.asarray(E_cm, float)
    with open(os.path.join(path, fname), "w") as f:
        f.write(f"!MOLECULE\n{name}\n!MOLECULAR WEIGHT\n{weight:4.1f}\n")
        f.write(f"!NUMBER OF ENERGY LEVELS\n{len(E_cm):6d}\n")
        f.write("!LEVEL + ENERGIES(cm^-1) + WEIGHT\n")
        for i, (e, gg) in enumerate(zip(E_cm, g)):
            f.write(f"{i + 1:5d}{e:12.4f}{gg:7.1f}\n")
        f.write(f"!NUMBER OF RADIATIVE TRANSITIONS\n{len(A):6d}\n")
        f.write("!TRANS + UP + LOW + EINSTEINA(s^-1)\n")
        for i, (u, l, a) in enumerate(zip(lev_up, lev_low, A)):
            f.write(f"{i + 1:5d}{int(u):5d}{int(l):5d}{a:12.3e}\n")


def write_levelpop(path, pops, E_erg=None, g=None, molname="moldata"):
    """levelpop.info + levelpop_<mol>.dat (line.F:842-919).
    ``pops`` has shape [nr, nth_half, nlev]."""
    nr, nth, nlev = pops.shape
    with open(os.path.join(path, "levelpop.info"), "w") as f:
        f.write(f"-3\nlevelpop_{molname}.dat\n0\n")
    E_erg = np.zeros(nlev) if E_erg is None else E_erg
    g = np.ones(nlev) if g is None else g
    with open(os.path.join(path, f"levelpop_{molname}.dat"), "w") as f:
        f.write(f"{nr} {nth} {nlev} 1\n")
        f.write(" ".join(_fmt(e) for e in E_erg) + "\n")
        f.write(" ".join(_fmt(x) for x in g) + "\n")
        flat = pops.reshape(nr * nth, nlev)
        f.write("\n".join(" ".join(_fmt(x) for x in row) for row in flat))
        f.write("\n")


def _write_scalar(fname, header, arr):
    with open(fname, "w") as f:
        f.write(header + "\n")
        f.write("\n".join(_fmt(x) for x in np.asarray(arr).ravel()))
        f.write("\n")


def write_fields(path, r, theta_half, rho_gas, abund, vel, lw_kms,
                 rho_dust, t_dust):
    """radius/theta/density/abundance/velocity/turbulence/dustdens/
    dusttemp files (grid.F:716-727,1103-1165; setup.F:775-786,824-844,
    902-911,1484-1485,701-725,159-250)."""
    nr, nth = rho_gas.shape
    _write_scalar(os.path.join(path, "radius.inp"), f"{nr}", r)
    _write_scalar(os.path.join(path, "theta.inp"), f"{nth} 1", theta_half)
    _write_scalar(os.path.join(path, "density.inp"), f"{nr} {nth} 1", rho_gas)
    with open(os.path.join(path, "abundance.inp"), "w") as f:
        f.write(f"{nr} {nth}\n")
        f.write("\n".join(_fmt(a) + " 0.0" for a in abund.ravel()) + "\n")
    with open(os.path.join(path, "velocity.inp"), "w") as f:
        f.write(f"{nr} {nth}\n")
        f.write("\n".join(" ".join(_fmt(v) for v in row)
                          for row in vel.reshape(nr * nth, 3)) + "\n")
    with open(os.path.join(path, "turbulence.inp"), "w") as f:
        f.write(f"1\n{nr} {nth}\n")
        f.write("\n".join(_fmt(x) for x in lw_kms.ravel()) + "\n")
    _write_scalar(os.path.join(path, "dustdens.inp"), f"1 {nr} {nth} 1",
                  rho_dust)
    with open(os.path.join(path, "dusttemp.info"), "w") as f:
        f.write("-2\n1\n1\n1\n1\n")
    _write_scalar(os.path.join(path, "dusttemp_final.dat"),
                  f"1 {nr} {nth} 1\n1", t_dust)


def write_dust_and_star(path, nu, kappa_abs, kappa_scat, rstar, fnu_1pc):
    """frequency.inp, dustopac.inp + dustopac_1.inp, starinfo.inp,
    starspectrum.inp (grid.F:1451-1469; dust.F:92-245,317-520;
    star.F:449-620)."""
    _write_scalar(os.path.join(path, "frequency.inp"), f"{len(nu)}", nu)
    with open(os.path.join(path, "dustopac.inp"), "w") as f:
        f.write("2\n1\n" + "=" * 40 + "\n-1\n0\n1\n" + "-" * 40 + "\n")
    with open(os.path.join(path, "dustopac_1.inp"), "w") as f:
        f.write(f"{len(nu)} 1\n\n")
        f.write("\n".join(_fmt(k) for k in kappa_abs) + "\n")
        f.write("\n".join(_fmt(k) for k in kappa_scat) + "\n")
    with open(os.path.join(path, "starinfo.inp"), "w") as f:
        f.write("1\n" + _fmt(rstar) + "\n")
    with open(os.path.join(path, "starspectrum.inp"), "w") as f:
        f.write(f"{len(nu)}\n")
        f.write("\n".join(_fmt(a) + " " + _fmt(b) for a, b in zip(nu, fnu_1pc)))
        f.write("\n")


def write_interstellfield(path, isrf):
    """interstellfield.inp: count, then one intensity per line."""
    _write_scalar(os.path.join(path, "interstellfield.inp"), f"{len(isrf)}", isrf)


def write_scatsource(path, scat, name="scatsource.dat"):
    """scatsource.dat (or quantsource.dat): ``nfreq nr nth 1``, then one
    value per line, frequency outer, theta, radius inner; scat is
    [nfreq, nr, nth]."""
    nf, nr, nth = scat.shape
    with open(os.path.join(path, name), "w") as f:
        f.write(f"{nf} {nr} {nth} 1\n")
        f.write("\n".join(_fmt(v) for v in np.asarray(scat).transpose(0, 2, 1).ravel()))
        f.write("\n")


# ----------------------------------------------------------------------
# readers (semantics of the reference's own readers)
# ----------------------------------------------------------------------
def _tokens(fname):
    with open(fname) as f:
        return f.read().split()


@dataclass
class RunInputs:
    """Everything the reference reads for a line spectrum, as doubles."""
    r: np.ndarray
    theta_half: np.ndarray
    rho_gas: np.ndarray            # [nr, nth]
    abund: np.ndarray              # [nr, nth]
    vel: np.ndarray                # [nr, nth, 3] cm/s
    lw_kms: np.ndarray             # [nr, nth] km/s
    pops: np.ndarray               # [nr, nth, nlev]
    rho_dust: np.ndarray           # [nspec, nr, nth]
    t_dust: np.ndarray             # [nspec, nr, nth]
    cont_nu: np.ndarray            # frequency.inp
    kappa_abs: np.ndarray          # [nspec, nfreq]
    kappa_scat: np.ndarray         # [nspec, nfreq]
    rstar: float
    star_fnu: np.ndarray           # F_nu at 1 pc on cont_nu
    umass_av: float
    fix_lw: float
    E_cm: np.ndarray
    g: np.ndarray
    lev_up: np.ndarray
    lev_low: np.ndarray
    A: np.ndarray
    molname: str
    nphi: int
    dbdr: int
    b_extra: int
    outer_bc: int
    vmax: float
    dv: float
    incl_deg: float
    vrad: float
    nlines_run: int
    iline_start: int
    extra: dict = field(default_factory=dict)
    # interstellar radiation field on the continuum grid (interstellfield.inp,
    # read for outer boundary type 3, star.F:675-720), or None
    isrf: np.ndarray | None = None
    # dust scattering/PAH source on the continuum grid, [nfreq, nr, nth]:
    # scatsource.dat plus quantsource.dat (source.F:494-546), or None when
    # neither file is present
    scat_src: np.ndarray | None = None


class _ListDirected:
    """Fortran list-directed READ of one file: each read takes the values
    it asks for from the next records (blank records and the ends of
    records only separate values), and the rest of its last record is
    skipped, so a following read starts on a new record."""

    def __init__(self, fname):
        with open(fname) as f:
            self.lines = f.read().splitlines()
        self.i = 0

    def firsts(self, n):
        """n reads of one value each: the first value of each of the next
        n non-blank records."""
        out = []
        while len(out) < n:
            if self.i >= len(self.lines):
                raise EOFError(f"end of file after {len(out)} of {n} records")
            toks = self.lines[self.i].replace(",", " ").split()
            self.i += 1
            if toks:
                out.append(toks[0])
        return np.array([v.replace("d", "e").replace("D", "e") for v in out], float)

    def read(self, n):
        vals = []
        while len(vals) < n:
            if self.i >= len(self.lines):
                raise EOFError(f"end of file after {len(vals)} of {n} values")
            toks = self.lines[self.i].replace(",", " ").split()
            self.i += 1
            vals.extend(toks[:n - len(vals)])
        return [v.replace("d", "e").replace("D", "e") for v in vals]


def read_interstellfield(path, nfreq):
    """interstellfield.inp (read_interstellfield, star.F:675-720): a count
    equal to frequency.inp's, then one intensity per record.  A missing
    file or a different count is the reference's stop 13."""
    fname = os.path.join(path, "interstellfield.inp")
    if not os.path.exists(fname):
        raise InputError(13, "outer BC type 3 without interstellfield.inp")
    rd = _ListDirected(fname)
    nf = int(rd.read(1)[0])
    if nf != nfreq:
        raise InputError(13, f"interstellfield.inp has {nf} frequencies, frequency.inp {nfreq}")
    return rd.firsts(nf)


def read_scatsource(path, nfreq, nr, nth):
    """scatsource.dat and quantsource.dat (read_scatsource,
    source.F:494-546), each optional: a header of four integers, then one
    value per record for every frequency of frequency.inp (outer), theta
    cell of the half grid, radius (inner); the two files are summed.
    Returns [nfreq, nr, nth] or None when neither file is present."""
    out = None
    for name in ("scatsource.dat", "quantsource.dat"):
        fname = os.path.join(path, name)
        if not os.path.exists(fname):
            continue
        rd = _ListDirected(fname)
        rd.read(4)
        a = rd.firsts(nfreq * nth * nr).reshape(nfreq, nth, nr).transpose(0, 2, 1)
        out = a.copy() if out is None else out + a
    return out


class InputError(RuntimeError):
    """An input condition the reference stops on, with its stop code."""

    def __init__(self, code, msg):
        super().__init__(f"stop {code}: {msg}")
        self.code = code


def read_run_dir(path) -> RunInputs:
    """Read a RADLite run directory the way main.F:79-238 does."""
    P = lambda n: os.path.join(path, n)  # noqa: E731
    # radlite.inp: 16-column fields (tools.F:6-19), order main.F:535-597
    with open(P("radlite.inp")) as f:
        rows = [ln[:16].split()[0] for ln in f
                if ln.strip() and ln[0] not in "#;%="]
    outer_bc, nphi, dbdr, b_extra = int(rows[13]), int(rows[17]), int(rows[18]), int(rows[19])
    t = _tokens(P("radius.inp"))
    nr = int(t[0]); r = np.array(t[1:1 + nr], float)
    t = _tokens(P("theta.inp"))
    nth = int(t[0]); theta = np.array(t[2:2 + nth], float)
    t = _tokens(P("density.inp"))
    rho = np.array(t[3:3 + nr * nth], float).reshape(nr, nth)
    t = _tokens(P("abundance.inp"))
    ab = np.array(t[2:2 + 2 * nr * nth], float).reshape(nr * nth, 2)[:, 0].reshape(nr, nth)
    t = _tokens(P("velocity.inp"))
    vel = np.array(t[2:2 + 3 * nr * nth], float).reshape(nr, nth, 3)
    t = _tokens(P("turbulence.inp"))
    lw = np.array(t[3:3 + nr * nth], float).reshape(nr, nth)
    # dust density: nspec nr nt imirt, then per species (setup.F:558-630)
    t = _tokens(P("dustdens.inp"))
    nsp = int(t[0])
    rhod = np.array(t[4:4 + nsp * nr * nth], float).reshape(nsp, nr, nth)
    t = _tokens(P("dusttemp_final.dat"))
    nsp_t = int(t[0]); k = 4; td = []
    for _ in range(nsp_t):
        nsize = int(t[k]); k += 1
        assert nsize == 1
        td.append(np.array(t[k:k + nr * nth], float).reshape(nr, nth)); k += nr * nth
    td = np.array(td)
    t = _tokens(P("frequency.inp"))
    nf = int(t[0]); nu = np.array(t[1:1 + nf], float)
    with open(P("dustopac.inp")) as f:
        dl = [ln.split()[0] for ln in f if ln.strip()]
    nspec = int(dl[1]); ka, ks = [], []
    for isp in range(nspec):
        idx = int(dl[5 + 4 * isp])
        t = _tokens(P(f"dustopac_{idx}.inp"))
        assert int(t[0]) == nf and int(t[1]) == 1
        ka.append(np.array(t[2:2 + nf], float))
        ks.append(np.array(t[2 + nf:2 + 2 * nf], float))
    t = _tokens(P("starinfo.inp"))
    rstar = float(t[1])
    t = _tokens(P("starspectrum.inp"))
    sp = np.array(t[1:1 + 2 * nf], float).reshape(nf, 2)
    # line.inp (line.F:168-279)
    with open(P("line.inp")) as f:
        ll = [ln.split()[0] for ln in f if ln.strip()]
    # format 2 has a species count after the format number, format 1 not
    # (line.F:173-180)
    o = 1 if int(float(ll[0])) == 2 else 0
    umass = float(ll[1 + o].replace("d", "e").replace("D", "e"))
    fix_lw = 0.0
    # first line block: comment, iinfo=1, symm, nfreq, width, proftype, imore...
    j = 4 + o
    j += 5
    while True:
        imore = int(ll[j]); j += 1
        if imore == 0:
            break
        fix_lw = float(ll[j].replace("d", "e").replace("D", "e")); j += 1
    # linespectrum.inp (telescope.F:104-238)
    with open(P("linespectrum.inp")) as f:
        ls = [ln.split()[0] for ln in f if ln.strip()]
    vmax, dv = float(ls[5]), float(ls[6])
    molfile = ls[9]
    incl_deg, vrad = float(ls[12]), float(ls[13])
    nlines_run, istart = int(ls[14]), int(ls[15])
    extra = {}
    if int(ls[10]) == 2:                  # position-velocity cube (telescope.F:196-224)
        extra["posvel"] = dict(nx=int(ls[16]), ny=int(ls[17]), size_x=float(ls[19]),
                               size_y=float(ls[20]), phioff=float(ls[21]),
                               xoff=float(ls[22]), yoff=float(ls[23]), addstar=int(ls[24]))
    molname = molfile[:molfile.rfind(".dat")]
    # moldata (line.F:1865-1935), fixed columns
    with open(P(molfile)) as f:
        ml = f.read().splitlines()
    nlev = int(ml[5][:6])
    E_cm, g = [], []
    for i in range(nlev):
        s = ml[7 + i]
        E_cm.append(float(s[5:17])); g.append(float(s[17:24]))
    nlin = int(ml[8 + nlev][:6])
    up, lo, A = [], [], []
    for i in range(nlin):
        s = ml[10 + nlev + i]
        up.append(int(s[5:10])); lo.append(int(s[10:15])); A.append(float(s[15:27]))
    # level populations (line.F:842-919)
    t = _tokens(P(f"levelpop_{molname}.dat"))
    nlp = int(t[2])
    pops = np.array(t[4 + 2 * nlp:4 + 2 * nlp + nr * nth * nlp], float).reshape(nr, nth, nlp)
    # optional continuum inputs (setup.F:1149 read_scatsource; main.F:929-941
    # read_interstellfield for outer boundary type 3)
    scat = read_scatsource(path, nf, nr, nth)
    isrf = read_interstellfield(path, nf) if outer_bc == 3 else None
    return RunInputs(
        r=r, theta_half=theta, rho_gas=rho, abund=ab, vel=vel, lw_kms=lw,
        pops=pops, rho_dust=rhod, t_dust=td, cont_nu=nu,
        kappa_abs=np.array(ka), kappa_scat=np.array(ks), rstar=rstar,
        star_fnu=sp[:, 1], umass_av=umass, fix_lw=fix_lw,
        E_cm=np.array(E_cm), g=np.array(g), lev_up=np.array(up),
        lev_low=np.array(lo), A=np.array(A), molname=molname, nphi=nphi,
        dbdr=dbdr, b_extra=b_extra, outer_bc=outer_bc, vmax=vmax, dv=dv,
        incl_deg=incl_deg, vrad=vrad, nlines_run=nlines_run,
        iline_start=istart, extra=extra, isrf=isrf, scat_src=scat)


def write_line_spectrum_file(fname, molname, nfr_max, incl_rad, blocks,
                             distance_pc=1.0, vrad=0.0, style=1, beamsize=0.0):
    """Spectrum output as the reference writes it (header_line_spectrum /
    write_line_spectrum, telescope.F:1662-1804): (i2) format version and
    style; (a80) molecule name and data file (read through the data
    directory, line.F:1840-1848, hence ./<name>.dat); (i10) reverting over
    the line count and the channel count (two records); (e12.4,e12.4,f7.3)
    distance [pc], radial velocity, inclination [deg] (anginf *
    57.2957795132); per line an empty record, (i5,i5) levels, (e14.9) nu0,
    (e10.5) beam size, (i5) channels, a ' ' record, and rows
    (E13.6,1X,E13.6) v, F_nu for inu = nfr..1.

    ``blocks`` is a list of dicts with keys lev_up, lev_low, nu0, nu
    (channel frequencies, ascending), flux (same length)."""
    def e(x, w, d):
        return _fortran_e(x, w, d)
    with open(fname, "w") as f:
        f.write(f"{1:2d}\n{style:2d}\n{molname:<80.80s}\n{'./' + molname + '.dat':<80.80s}\n")
        f.write(f"{len(blocks):10d}\n{nfr_max:10d}\n")
        f.write(e(distance_pc, 12, 4) + e(vrad, 12, 4) + f"{incl_rad * 57.2957795132:7.3f}\n")
        for b in blocks:
            nu0 = b["nu0"]
            f.write("\n" + f"{b['lev_up']:5d}{b['lev_low']:5d}\n")
            f.write(e(nu0, 14, 9) + "\n" + e(beamsize, 10, 5) + "\n")
            f.write(f"{len(b['flux']):5d}\n  \n")
            for k in range(len(b["flux"]) - 1, -1, -1):
                v = b["nu"][k] - nu0
                v = -2.99792458e5 * v / nu0
                v = v + vrad
                f.write(e(v, 13, 6) + " " + e(b["flux"][k], 13, 6) + "\n")


def _fortran_e(x: float, w: int, d: int) -> str:
    """Fortran Ew.d edit descriptor: 0.ddddE+xx, right-justified; a
    three-digit exponent drops the E (0.dddd-119)."""
    x = float(x)
    if x == 0.0:
        s = "0." + "0" * d + "E+00"
    else:
        m, ex = f"{abs(x):.{d - 1}e}".split("e")
        digits = m.replace(".", "")
        ex = int(ex) + 1
        tail = ("E%+03d" % ex) if abs(ex) <= 99 else ("%+04d" % ex)
        s = ("-" if x < 0 else "") + "0." + digits + tail
    if len(s) > w:                    # the leading zero is optional
        s = s.replace("0.", ".", 1)
    if len(s) > w:
        return "*" * w
    return s.rjust(w)


def _fortran_es(x: float, w: int, d: int) -> str:
    """Fortran ESw.d edit descriptor: d.ddddE+xx, right-justified; a
    three-digit exponent drops the E (d.dddd-119)."""
    m, ex = f"{float(x):.{d}E}".split("E")
    e = int(ex)
    s = m + (("E%+03d" % e) if abs(e) <= 99 else ("%+04d" % e))
    return s.rjust(w) if len(s) <= w else "*" * w


def posvel_file_name(molname: str, iline: int) -> str:
    """lineposvel_<mol>_<line>.dat as calc_write_line_posvel names it
    (telescope.F:1940-1962, the non-Intel branch): the line number is
    written (i2) into a 12-character buffer, of which characters 2:2 are
    taken below line 10 and 2:3 (the second digit and a blank) from 10 to
    99; line 100 and above stop the reference (code 177)."""
    if iline < 1 or iline >= 100:
        raise ValueError("the reference names posvel files for lines 1-99 only")
    ch = f"{iline:2d}".ljust(12)
    return f"lineposvel_{molname}_" + (ch[1:2] if iline < 10 else ch[1:3]) + ".dat"


def write_posvel_file(fname, molname, lev_up, lev_low, nu0, nu, incl_rad, pix, image, tau,
                      distance_pc=1.0, vrad=0.0):
    """Position-velocity cube as calc_write_line_posvel writes it
    (telescope.F:1966-2019): an empty record; (I2) format version 1;
    (A80) molecule name and data file; (ES12.4,ES12.4,F7.3) distance
    [pc], radial velocity, inclination [deg]; (I5,I5) levels; (ES12.4)
    nu0; (I5) channels; (I5,I5,ES12.4,ES12.4,F7.3,ES12.4,ES12.4) nx, ny,
    pixel sizes [cm], phioff, xoff, yoff; a ' ' record; the velocities
    (E10.3, one per record); a ' ' record; then per channel, iy outer
    and ix inner, (ES12.5,1X,ES12.5) the Rayleigh-Jeans brightness
    temperature 3.25465503368e36 / nu0^2 * I and the optical depth,
    closed by a ' ' record.

    nu0: the line's signed rest frequency; nu: channel frequencies [nfr];
    pix: dict nx, ny, spx, spy (pixel sizes), phioff, xoff, yoff;
    image, tau: [ny, nx, nfr]."""
    es = _fortran_es
    ny, nx, nfr = image.shape
    out = ["", " 1", f"{molname:<80.80s}", f"{'./' + molname + '.dat':<80.80s}",
           es(distance_pc, 12, 4) + es(vrad, 12, 4) + f"{incl_rad * 57.2957795132:7.3f}",
           f"{int(lev_up):5d}{int(lev_low):5d}", es(nu0, 12, 4), f"{nfr:5d}",
           f"{int(pix['nx']):5d}{int(pix['ny']):5d}" + es(pix["spx"], 12, 4) +
           es(pix["spy"], 12, 4) + f"{pix['phioff']:7.3f}" + es(pix["xoff"], 12, 4) +
           es(pix["yoff"], 12, 4), "  "]
    for k in range(nfr):
        v = float(nu[k]) - nu0
        v = -2.99792458e5 * v / nu0
        v = v + vrad
        out.append(_fortran_e(v, 10, 3))
    out.append("  ")
    temp = 3.25465503368e36 / (nu0 * nu0)
    for k in range(nfr):
        tk = temp * image[:, :, k]
        for iy in range(ny):
            out.extend(es(a, 12, 5) + " " + es(b, 12, 5)
                       for a, b in zip(tk[iy].tolist(), tau[iy, :, k].tolist()))
        out.append("  ")
    with open(fname, "w") as f:
        f.write("\n".join(out) + "\n")


def _flang_list(x: float) -> str:
    """A REAL(8) as flang's list-directed output writes it: one leading
    blank, the shortest round-trip digits, F form (no leading zero,
    trailing point for integral values) for 0.095 <= |x| < 1e15, else
    d.dddE+xx."""
    from decimal import Decimal
    x = float(x)
    if x == 0.0:
        return " 0."
    d = Decimal(repr(abs(x))).normalize()
    t = d.as_tuple()
    D = "".join(str(v) for v in t.digits)
    k = len(D) + t.exponent            # |x| = 0.D x 10^k
    ax = abs(x)
    if 0.095 <= ax < 1e15 - 0.5:
        if k <= 0:
            body = "." + "0" * (-k) + D
        elif k >= len(D):
            body = D + "0" * (k - len(D)) + "."
        else:
            body = D[:k] + "." + D[k:]
    else:
        e = k - 1
        body = D[0] + "." + D[1:] + "E" + ("+" if e >= 0 else "-") + "%02d" % abs(e)
    return " " + ("-" if x < 0 else "") + body


def write_image_cube(fname, nu0, nphi, nrings, ring_edge, ring_r, velo_kms, image, cmask):
    """Circular image cube lineposvelcirc_<mol>_<line>.dat (SAVE_IMCIR,
    telescope.F:1595-1621, 1348-1373; read by PRO/read_imcir.pro):
    nfr; nu0; nphi nr; ring edges ri(1..nr+1) and radii r(1..nr+1) in
    E12.6; then per channel a blank record, the velocity dnu/nu0*c, a
    blank record, the central intensity, a blank record and for every
    phi, every ring (E10.4,2X,i1) intensity and integrated-cell mask.

    ring_edge: [nrings+2] edges ri(0..nr+1); ring_r: [nrings+1] radii
    r(0..nr) (r(nr+1) is never set by the reference and is written as 0);
    image, cmask: [nrays, nfr] with ray 0 the central beam and ray
    1 + (ir-1)*nphi + (ip-1) the ring ir, angle ip."""
    nfr = image.shape[1]
    r_ext = list(ring_r[1:nrings + 1]) + [0.0]
    lines = [" %d" % nfr, _flang_list(nu0), " %d %d" % (nphi, nrings)]
    lines += [_fortran_e(ring_edge[ir], 12, 6) for ir in range(1, nrings + 2)]
    lines += [_fortran_e(v, 12, 6) for v in r_ext]
    rings = image[1:].reshape(nrings, nphi, nfr)
    masks = cmask[1:].reshape(nrings, nphi, nfr)
    blank = "        "
    for ch in range(nfr):
        lines += [blank, _flang_list(velo_kms[ch]), blank, _flang_list(image[0, ch]), blank]
        a = rings[:, :, ch].T.ravel()       # phi outer, ring inner
        m = masks[:, :, ch].T.ravel()
        lines += [_fortran_e(v, 10, 4) + "  %1d" % f for v, f in zip(a, m)]
    with open(fname, "w") as f:
        f.write("\n".join(lines) + "\n")


def read_image_cube(fname):
    """Parse a circular image cube the way read_imcir.pro does."""
    with open(fname) as f:
        rec = f.read().split("\n")
    k = 0
    nfr = int(rec[k]); k += 1
    nu0 = float(rec[k]); k += 1
    nphi, nr = (int(v) for v in rec[k].split()); k += 1
    ri = np.array([_f(rec[k + i]) for i in range(nr + 1)]); k += nr + 1
    r = np.array([_f(rec[k + i]) for i in range(nr + 1)]); k += nr + 1
    velo = np.zeros(nfr); center = np.zeros(nfr)
    img = np.zeros((nfr, nphi, nr)); cm = np.zeros((nfr, nphi, nr), np.int8)
    for ch in range(nfr):
        k += 1
        velo[ch] = _f(rec[k]); k += 2
        center[ch] = _f(rec[k]); k += 2
        for ip in range(nphi):
            for ir in range(nr):
                a, b = rec[k].rsplit(None, 1)
                img[ch, ip, ir] = _f(a); cm[ch, ip, ir] = int(b); k += 1
    return dict(nfr=nfr, nu0=nu0, nphi=nphi, nr=nr, ri=ri, r=r, velo=velo, center=center,
                image=img, cmask=cm)


def _f(s: str) -> float:
    """A Fortran real field, including E-less three-digit exponents."""
    s = s.strip()
    if "E" not in s.upper() and ("-" in s[1:] or "+" in s[1:]):
        i = max(s.rfind("-"), s.rfind("+"))
        s = s[:i] + "E" + s[i:]
    return float(s)


def read_line_spectrum_file(fname):
    """Parse a spectrum file with pyradlite's indices (radlite.py:2825-2850):
    line count in record 4, channel maximum in record 5, inclination in
    record 6; per line the channel count 5 records after the block start
    and the rows from 7 records after it."""
    with open(fname) as f:
        lines = f.read().splitlines()
    nl = int(lines[4])
    out, iloc = [], 6
    for _ in range(nl):
        up, lo = (int(x) for x in lines[iloc + 2].split())
        nu0 = float(lines[iloc + 3])
        nfr = int(lines[iloc + 5])
        rows = np.array([[float(x) for x in ln.split()] for ln in lines[iloc + 7:iloc + 7 + nfr]])
        out.append(dict(lev_up=up, lev_low=lo, nu0=nu0, v=rows[:, 0], flux=rows[:, 1]))
        iloc = iloc + 7 + nfr - 1
    return out
