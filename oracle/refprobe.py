"""Readers for the binary dumps written by oracle/refbuild/probe_dump.F
(test infrastructure: used by tests/golden/make_golden.py only).

Record layout per line (all little-endian, Fortran stream access):
int32[7] header (iline, nfr, rays_amount, nrr, nphi, maserflag,
overshoot), then the arrays listed in probe_dump.F.
"""
import numpy as np


def read_probe_out(fname):
    buf = open(fname, "rb").read()
    off, recs = 0, []

    def take(dtype, n):
        nonlocal off
        a = np.frombuffer(buf, dtype=dtype, count=n, offset=off)
        off += a.nbytes
        return a.copy()

    while off < len(buf):
        il, nfr, nray, nrr, nphi, maser, oversh = take("<i4", 7)
        rec = dict(iline=int(il), nfr=int(nfr), nray=int(nray), nrr=int(nrr),
                   nphi=int(nphi), maser=int(maser), overshoot=int(oversh))
        rec["flux"] = take("<f8", nfr)
        rec["dnu"] = take("<f8", nfr)
        rec["nu0"] = float(take("<f8", 1)[0])
        rec["ring_r"] = take("<f8", nrr + 1)
        rec["ring_ri"] = take("<f8", nrr + 2)
        # imcir_int(1:nfr,1:nphi,0:nrr) -> [ring][phi][ch]
        rec["image"] = take("<f8", nfr * nphi * (nrr + 1)).reshape(nrr + 1, nphi, nfr)
        rec["cmask"] = take("<i4", nfr * nphi * nrr).reshape(nrr, nphi, nfr)
        rec["minvel"] = take("<f4", nray)
        rec["maxvel"] = take("<f4", nray)
        rec["x0"] = take("<f8", nray)
        rec["z0"] = take("<f8", nray)
        rec["theta0"] = float(take("<f8", 1)[0])
        rec["aksmax"] = float(take("<f8", 1)[0])
        recs.append(rec)
    return recs


def read_probe_traj(fname):
    buf = open(fname, "rb").read()
    off, out = 0, []

    def take(dtype, n):
        nonlocal off
        a = np.frombuffer(buf, dtype=dtype, count=n, offset=off)
        off += a.nbytes
        return a.copy()

    while off < len(buf):
        iray, n = (int(x) for x in take("<i4", 2))
        d = take("<f8", 5 * n).reshape(5, n)
        i = take("<i4", 3 * n).reshape(3, n)
        out.append(dict(iray=iray, s=d[0], r=d[1], theta=d[2], mu=d[3], phi=d[4],
                        icross=i[0], ir=i[1], it=i[2]))
    return out


def read_probe_posvel(fname):
    """Records of probe_posvel.bin (probe_dump.F, one per line)."""
    buf = open(fname, "rb").read()
    pos = 0
    out = []

    def take(dtype, n):
        nonlocal pos
        a = np.frombuffer(buf, dtype=dtype, count=n, offset=pos)
        pos += a.nbytes
        return a

    while pos < len(buf):
        il, nfr, nx, ny, nray, unres, maser, ovs = [int(v) for v in take(np.int32, 8)]
        nu0 = float(take(np.float64, 1)[0])
        dnu = take(np.float64, nfr).copy()
        img = take(np.float64, nfr * nx * ny).reshape(ny, nx, nfr).copy()
        tau = take(np.float64, nfr * nx * ny).reshape(ny, nx, nfr).copy()
        x0 = take(np.float64, nray).copy()
        z0 = take(np.float64, nray).copy()
        b = take(np.float64, nray).copy()
        th0 = float(take(np.float64, 1)[0])
        out.append(dict(iline=il, nfr=nfr, nx=nx, ny=ny, nrays=nray, starunres=unres,
                        maser=maser, overshoot=ovs, nu0=nu0, dnu=dnu, image=img, tau=tau,
                        x0=x0, z0=z0, b=b, theta0=th0))
    return out
