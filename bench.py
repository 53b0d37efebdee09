"""Benchmark: ray*channel integrations per second of the circular-image
line raytrace on one MI355X.

Workload (default, one GPU): the nr=400 x nth=200 disk with 200
channels over +-40 km/s, 113+10+399 rings x 256 phi = 133,633 rays per
inclination, inclinations 30 and 60 deg, T with lognormal(0.1) noise, one
CO-like line at 4.7 um.  One step = the whole hot path for both
inclinations: ray set, trajectories (FP64), node records, formal
solution, flux sums (telescope.F:1465 calc_line_spectrum per item).

Prints one JSON line.  value = nominal ray*channels (rays x channels x
inclinations, the reference's own yardstick) / seconds per step.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
FP64_PEAK_TFLOPS = 78.6




def measured_traffic(precision, kernel):
    """HBM bytes per launch of the formal-solution kernel from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (counter values calibrated
    against a plain copy), or None when the summary has no line for this
    kernel and precision."""
    path = os.path.join(ROOT, "profiles", f"hbm_traffic_c2_{precision}_forward.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        f = line.split()
        if (len(f) == 4 and f[0] == "traffic_bytes_per_launch" and f[1] == kernel
                and f[2] == precision):
            return float(f[3])
    return None


def measured_valu_issue(precision):
    """VALU wave-instructions per CU-cycle of the formal kernel over the FP32
    wave64 issue peak (2 per CU-cycle), from the committed rocprofv3 SQ
    counter summary, or None."""
    path = os.path.join(ROOT, "profiles", f"sq_counters_c2_{precision}_forward.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if "VALU wave-instructions per CU-cycle" in line:
            return float(line.split("=")[1].split()[0]) / 2.0
    return None


def bench_adjoint(args):
    """Forward vs forward+adjoint on the 200x100 disk (34,049 rays x 100
    channels, 45 deg, every channel integrated).  Forward: the spectrum
    (rl_trace_lines_forward).  Forward+adjoint, one training step: the
    spectrum with the adjoint's tape kept (rl_trace_lines_forward_tape: I
    checkpoints) and then the gradient of a chi^2 loss
    (rl_trace_lines_adjoint: the reverse sweep from the tape, down to
    per-cell gradients)."""
    from radlite_amd import engine, synth
    ri, w = synth.workload_inputs("c4", seed=0)
    nfr, vmax = engine.passband(ri.vmax, ri.dv)
    rt = engine.Raytracer(0, args.precision)
    rt.load_run(ri)
    rt.setup_rays([w["incl"][0] * 0.0174532925199], ri.nphi, ri.b_extra, ri.dbdr)
    items = [(0, 0, nfr, vmax, engine.SKIP_NONE)]
    fl, _, _ = rt.trace(items)
    target = fl[0] * 1.01
    sigma = 1e-3 * np.abs(target).max()

    def fwd():
        return rt.trace(items)[0]

    def fwd_adj():
        f = rt.trace_tape(items)[0]
        g = 2.0 * (f - target) / sigma ** 2
        return rt.adjoint(items, [g])

    def fwd_tape():
        return rt.trace_tape(items)

    res = {}
    # checkpoint interval chosen by memory (every segment kept when it fits),
    # then fixed at 16 segments, the interval the 200x100 workload states
    for ck, sfx in ((0, ""), (16, "_k16")):
        rt.set_option(engine.OPT_CKPT, ck)
        for name, fn in (("fwd", fwd), ("fwd_tape", fwd_tape), ("fwd_adj", fwd_adj)):
            if ck and name == "fwd":
                continue
            for _ in range(max(1, args.warmup)):
                fn()
            rt.L.rl_sync(rt.ctx)
            t0 = time.perf_counter()
            for _ in range(args.steps):
                fn()
            rt.L.rl_sync(rt.ctx)
            res[name + sfx] = (time.perf_counter() - t0) / args.steps
            if name == "fwd_adj":
                st = rt.stats()
                res["adj_device_ms" + sfx] = st["t_formal_ms"]
                res["kblk" + sfx], res["used_tape" + sfx], _ = rt.adjoint_info()
    rt.set_option(engine.OPT_CKPT, 0)
    nominal = rt.nrays[0] * nfr
    rt.close()
    return {
        "workload": f"c4: nr={w['nr']} x nth={w['nth']}, {nominal // nfr} rays x {nfr} ch x 1 incl",
        "precision": args.precision,
        "fwd_ms": res["fwd"] * 1e3,
        "fwd_tape_ms": res["fwd_tape"] * 1e3,
        "fwd_adj_ms": res["fwd_adj"] * 1e3,
        "adjoint_call_device_ms": res["adj_device_ms"],
        "checkpoint_interval": res["kblk"],
        "adjoint_from_tape": bool(res["used_tape"]),
        "ratio_fwd_adj_over_fwd": res["fwd_adj"] / res["fwd"],
        "value_fwd": nominal / res["fwd"],
        "value_fwd_adj": nominal / res["fwd_adj"],
        "k16": {
            "checkpoint_interval": res["kblk_k16"],
            "adjoint_from_tape": bool(res["used_tape_k16"]),
            "fwd_tape_ms": res["fwd_tape_k16"] * 1e3,
            "fwd_adj_ms": res["fwd_adj_k16"] * 1e3,
            "adjoint_call_device_ms": res["adj_device_ms_k16"],
            "ratio_fwd_adj_over_fwd": res["fwd_adj_k16"] / res["fwd"],
            "value_fwd_adj": nominal / res["fwd_adj_k16"],
        },
        "unit": "ray*channel/s",
    }


def bench_c5_slice(args):
    """A measured slice of the large workload: the 1024x512 disk (529,409
    rays per inclination, 300 channels, non-LTE populations of the
    366-level ro-vibrational ladder) with the first 2 of its 1000 lines at
    all 4 inclinations, FP32.  Forward: the 8 items in one
    rl_trace_lines_forward call (batches cut to the memory budget,
    trajectories of other inclinations evicted and rebuilt as the whole
    job would).  Adjoint: one item's gradient with respect to the
    populations (checkpoint interval and ray chunks chosen to fit)."""
    from radlite_amd import engine, synth
    t0 = time.perf_counter()
    ri, items, incl = synth.c5_inputs(0, keep_lines=2)
    incs = [x * 0.0174532925199 for x in incl]
    its = [(il, ii, n, v, engine.SKIP_NONE) for (il, ii, n, v) in items]
    rt = engine.Raytracer(0, "fp32")
    rt.load_run(ri)
    rt.setup_rays(incs, ri.nphi, ri.b_extra, ri.dbdr)
    t_setup = time.perf_counter() - t0
    print(f"[c5] inputs, tables and ray sets: {t_setup:.1f} s", file=sys.stderr, flush=True)
    fwd = []
    for rep in range(2):                          # the first pass tunes and builds
        t1 = time.perf_counter()
        rt.trace(its)
        fwd.append(time.perf_counter() - t1)
        st = rt.stats()
        print(f"[c5] forward pass {rep}: {fwd[-1]:.2f} s, batches {rt.batch_info()}",
              file=sys.stderr, flush=True)
    nominal = sum(int(rt.nrays[it[1]]) * it[2] for it in its)
    mem = rt.memory_info()
    g = [np.random.default_rng(1).normal(size=its[0][2])]
    adj = []
    for rep in range(2):
        t1 = time.perf_counter()
        _, gr = rt.adjoint([its[0]], g, outputs=("g_pops",))
        adj.append(time.perf_counter() - t1)
        print(f"[c5] adjoint item pass {rep}: {adj[-1]:.2f} s, (K, tape, chunks) "
              f"{rt.adjoint_info()}", file=sys.stderr, flush=True)
    K, _, nch = rt.adjoint_info()
    assert np.isfinite(gr["g_pops"]).all()
    nb = rt.batch_info()[0]
    nrays0 = int(rt.nrays[0])
    rt.close()
    return {
        "workload": f"c5 slice: nr=1024 x nth=512, {int(its[0][2])} ch, 2 of 1000 lines x "
                    f"{len(incl)} incl, {nrays0} rays per incl, fp32",
        "precision": "fp32", "items": len(its), "fwd_s": fwd[-1],
        "fwd_items_per_s": len(its) / fwd[-1], "fwd_value": nominal / fwd[-1],
        "fwd_batches": nb, "fwd_t_traj_ms": st["t_traj_ms"],
        "fwd_nominal_ray_channels": nominal,
        "adj_item_s": adj[-1], "adj_value": nrays0 * its[0][2] / adj[-1],
        "adj_ckpt_interval": K, "adj_chunks": nch,
        "budget_gb": mem["budget"] / 1e9, "unit": "ray*channel/s",
    }


def bench_large(args):
    """The 1024x512 disk with 1000 ro-vibrational-like lines, 300 channels,
    4 inclinations (529,409 rays each): the memory plan of the whole job
    (4000 items, forward and adjoint) without running it, then a measured
    slice: the forward of 2 lines at every inclination (8 items, batches,
    trajectories built per inclination and evicted to fit), and the
    adjoint of one item with gradients with respect to the populations
    (checkpoint interval and ray chunks chosen to fit).  The whole job's
    time is extrapolated from the per-item times."""
    from radlite_amd import engine, synth
    t0 = time.perf_counter()
    ri, items, incl = synth.c5_inputs(0)
    incs = [x * 0.0174532925199 for x in incl]
    its = [(il, ii, n, v, engine.SKIP_NONE) for (il, ii, n, v) in items]
    t1 = time.perf_counter()
    tb = engine.prepare_tables(ri)                # host tables, the reference's arithmetic
    t_host_tables = time.perf_counter() - t1
    rt = engine.Raytracer(0, "fp32")
    rt.load_run(ri, tables=tb)
    t_setup = time.perf_counter() - t0
    print(f"[large] inputs and tables: {t_setup:.1f} s (host dust tables {t_host_tables:.1f} s)",
          file=sys.stderr, flush=True)
    # the same tables from the per-cell fields on the device (LTE populations
    # of the 366 levels, widths, dust at the 1000 line centres)
    from radlite_amd import diff
    rt.set_level_data(np.asarray(ri.E_cm, float) * synth.K_PER_CM, np.asarray(ri.g, float))
    ka = np.array([diff.kappa_at_lines(ri.cont_nu, k, tb.nu0) for k in ri.kappa_abs])
    ks = np.array([diff.kappa_at_lines(ri.cont_nu, k, tb.nu0) for k in ri.kappa_scat])
    rt.set_dust_species(ri.rho_dust, ka, ks)
    vt = np.full(ri.rho_gas.shape, 0.1)
    for rep in range(2):
        t1 = time.perf_counter()
        rt.set_physics(ri.t_dust[0], vt, log_abund=np.log(ri.abund), q_over=0.0)
        t_dev_pre = time.perf_counter() - t1
    print(f"[large] device pre-pass: {t_dev_pre*1e3:.1f} ms", file=sys.stderr, flush=True)
    rt.load_run(ri, tables=tb)                    # back to the run's own tables
    rt.setup_rays(incs, ri.nphi, ri.b_extra, ri.dbdr)
    plan = rt.plan(its)
    print(f"[large] plan of {len(its)} items: {plan}", file=sys.stderr, flush=True)
    sl = [it for it in its if it[0] < 2]
    res = {"workload": f"c5: nr=1024 x nth=512, {ri.nphi} phi, {rt.nrays[0]} rays x "
                       f"{its[0][2]} ch, {len(ri.A)} lines x {len(incl)} incl",
           "precision": "fp32", "plan_full_job": plan, "setup_s": t_setup,
           "host_dust_tables_s": t_host_tables, "device_prepass_s": t_dev_pre}
    for rep in range(2):                          # the first pass tunes and builds
        t1 = time.perf_counter()
        rt.trace(sl)
        dt = time.perf_counter() - t1
        print(f"[large] forward slice pass {rep}: {dt:.2f} s, batches {rt.batch_info()}",
              file=sys.stderr, flush=True)
    nominal = sum(rt.nrays[it[1]] * it[2] for it in sl)
    res.update(fwd_items=len(sl), fwd_s=dt, fwd_batches=rt.batch_info()[0],
               fwd_value=nominal / dt, fwd_mem=rt.memory_info())
    it1 = [its[0]]
    g = [np.random.default_rng(1).normal(size=it1[0][2])]
    for rep in range(2):
        t1 = time.perf_counter()
        _, gr = rt.adjoint(it1, g, outputs=("g_pops",))
        dta = time.perf_counter() - t1
        print(f"[large] adjoint item pass {rep}: {dta:.2f} s, (K, tape, chunks) "
              f"{rt.adjoint_info()}", file=sys.stderr, flush=True)
    K, _, nch = rt.adjoint_info()
    assert np.isfinite(gr["g_pops"]).all()
    res.update(adj_item_s=dta, adj_ckpt_interval=K, adj_chunks=nch, adj_mem=rt.memory_info(),
               adj_value=rt.nrays[0] * it1[0][2] / dta,
               extrapolated_fwd_h=len(its) * dt / len(sl) / 3600,
               extrapolated_fwd_adj_h=len(its) * (dt / len(sl) + dta) / 3600,
               unit="ray*channel/s")
    rt.close()
    return res


def bench_batched(args):
    """The batched workload: 256x128 disk, 64 random lines with 20-400
    channels each, 8 inclinations (512 items, 17,025 rays each); one step
    = ray sets and trajectories for the 8 inclinations + every item's
    spectrum in one rl_trace_lines_forward call (batches sized to HBM,
    persistent formal kernel, graph replay, autotuned tiles)."""
    from radlite_amd import engine, synth
    ri, items, incl = synth.c3_inputs(0)
    incs = [x * 0.0174532925199 for x in incl]
    its = [(il, ii, n, v, engine.SKIP_NONE) for (il, ii, n, v) in items]
    rt = engine.Raytracer(0, args.precision)
    rt.load_run(ri)

    def step():
        rt.setup_rays(incs, ri.nphi, ri.b_extra, ri.dbdr)
        rt.trace(its)

    for _ in range(max(1, args.warmup)):
        step()
    rt.L.rl_sync(rt.ctx)
    steps = max(1, min(args.steps, 3))
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    rt.L.rl_sync(rt.ctx)
    dt = (time.perf_counter() - t0) / steps
    nominal = int(sum(int(rt.nrays[ii]) * n for (_, ii, n, _, _) in its))
    nb, cw, blocks = rt.batch_info()
    st = rt.stats()
    rt.close()
    return {
        "workload": f"c3: nr=256 x nth=128, 64 lines x 8 incl, 17025 rays/incl, "
                    f"{sum(n for (_, _, n, _, _) in its) // 8} channels/incl (20-400 per line)",
        "precision": args.precision, "items": len(its), "steps": steps,
        "ms_per_step": dt * 1e3, "value": nominal / dt, "unit": "ray*channel/s",
        "nominal_ray_channels": nominal, "batches": nb, "tile_cw": cw,
        "persistent_blocks": blocks, "t_traj_ms": st["t_traj_ms"],
    }


def bench_posvel(args):
    """A position-velocity cube: the 400x200 disk at 60 deg, 512x512 pixel
    rays over 2.2 outer radii, 200 channels over +-40 km/s.  One step = the
    rectangular ray set and its trajectories + the cube (intensity and
    optical depth) returned to host memory.  Device time (trajectories,
    node records, formal kernel) is reported beside the whole call, which
    includes the 2 x 420 MB device-to-host copy."""
    import numpy as np
    from radlite_amd import engine, synth
    ri, w = synth.workload_inputs("c2", seed=0)
    nfr, vmax = engine.passband(ri.vmax, ri.dv)
    nx = ny = 512
    size = 2.2 * float(ri.r[-1])
    incl = 60.0 * 0.0174532925199
    rt = engine.Raytracer(0, args.precision)
    rt.load_run(ri)

    def step():
        rt.setup_rays_rectang([incl], nx, ny, size, size, 0.0, 0.0, 0.0, 1)
        img, tau, _ = rt.trace_posvel([(0, 0, nfr, vmax)])
        return img[0], rt.stats()

    for _ in range(max(1, args.warmup)):
        step()
    steps = max(1, min(args.steps, 3))
    dev_ms = []
    t0 = time.perf_counter()
    for _ in range(steps):
        img, st = step()
        dev_ms.append(st["t_total_ms"])
    dt = (time.perf_counter() - t0) / steps
    rt.close()
    npc = nx * ny * nfr
    dms = float(np.mean(dev_ms))
    return {
        "workload": f"pos-vel cube: nr=400 x nth=200 disk, 60 deg, {nx}x{ny} pixels, "
                    f"{nfr} channels", "precision": args.precision, "steps": steps,
        "pixel_channels": npc, "traced_rays": st["rays"], "nodes": st["nodes"],
        "device_ms": dms, "device_value": npc / (dms * 1e-3),
        "call_ms": dt * 1e3, "call_value": npc / dt, "unit": "pixel*channel/s",
        "t_traj_ms": st["t_traj_ms"], "t_trace_ms": st["t_formal_ms"],
    }


def host_cores():
    """CPU threads this process may run on (its affinity set), and the
    cgroup CPU quota when one is set (cpu.max), for the record."""
    cores = len(os.sched_getaffinity(0))
    quota = None
    try:
        q, per = open("/sys/fs/cgroup/cpu.max").read().split()
        if q != "max":
            quota = float(q) / float(per)
    except (OSError, ValueError):
        pass
    return cores, quota


def reference_on_host(cores):
    """The reference's own trace (oracle/_ref/RADlite_probe, the flang -O2
    build of the reference sources) on this host's cores: ``cores``
    processes at once, one per core, each rendering the 100x60 disk
    (7,041 rays x 50 channels, NONREDUNDANT skipping compiled in, as the
    reference always runs), as pyradlite runs one process per work item.
    The probe records the wall time of each calc_line_spectrum call
    (telescope.F:1465), so set-up and file output are not timed.  The
    oracle traces the same run with ``cores`` threads and the same
    skipping.  Returns None where the reference build is absent (it is
    compiled in the build container from the reference sources)."""
    import shutil
    import subprocess
    import tempfile
    probe = os.path.join(ROOT, "oracle", "_ref", "RADlite_probe")
    if not os.access(probe, os.X_OK):
        return None
    from radlite_amd import engine, l4io, synth
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyorc
    base = tempfile.mkdtemp(prefix="rl_refcpu_")
    try:
        src = os.path.join(base, "run")
        synth.write_workload(src, "c1", seed=0)
        dirs = []
        for k in range(cores):
            d = os.path.join(base, f"p{k}")
            shutil.copytree(src, d)
            dirs.append(d)
        t0 = time.perf_counter()
        procs = [subprocess.Popen([probe], cwd=d, stdout=subprocess.DEVNULL,
                                  stderr=subprocess.DEVNULL) for d in dirs]
        rcs = [p.wait(timeout=600) for p in procs]
        wall = time.perf_counter() - t0
        if any(rcs):
            return {"error": f"reference probe exit codes {sorted(set(rcs))}"}
        tref = [sum(float(ln.split()[1]) for ln in open(os.path.join(d, "probe_time.txt")))
                for d in dirs]
        ri = l4io.read_run_dir(src)
        nfr, _ = engine.passband(ri.vmax, ri.dv)
        orc = pyorc.Oracle(pyorc.Prepared(ri))
        torc = []
        for _ in range(3):
            t1 = time.perf_counter()
            o = orc.line_spectrum(0, skip_mode=0, nthreads=cores, want_image=False)
            torc.append(time.perf_counter() - t1)
        nrc = o["nrays"] * nfr
        ref_rate = sum(nrc / t for t in tref)          # all processes at once
        orc_rate = nrc / sorted(torc)[1]
        return {"workload": f"100x60 disk, {o['nrays']} rays x {nfr} ch, one line, 45 deg",
                "processes": cores, "wall_s": wall,
                "trace_s_median": sorted(tref)[len(tref) // 2],
                "reference_value": ref_rate, "oracle_value": orc_rate,
                "ref_over_oracle_rate": ref_rate / orc_rate}
    finally:
        shutil.rmtree(base, ignore_errors=True)


def cpu_baseline(args, ri, incs, skip, nfr, gpu_value):
    """The oracle (C restatement, OpenMP over rays) on every core of this
    host on a sample of the headline workload, and the reference's own
    trace on the same cores (reference_on_host), whose rate ratio to the
    oracle, measured here, converts the sample into reference-equivalent
    throughput.  The reference build cannot hold the 400x200 disk itself:
    its compile-time capacity is 300 radii (configure.h FRSIZE_X)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyorc
    cores, quota = host_cores()
    orc = pyorc.Oracle(pyorc.Prepared(ri))
    t0 = time.perf_counter()
    nrc = 0
    for i in range(len(incs)):
        o = orc.line_spectrum(0, incl_rad=incs[i], skip_mode=1 if skip else 0,
                              nthreads=cores, ray_stride=args.cpu_stride, want_image=False)
        nrc += ((o["nrays"] + args.cpu_stride - 1) // args.cpu_stride) * nfr
    tc = time.perf_counter() - t0
    res = {
        "value": nrc / tc, "unit": "ray*channel/s", "cores": cores, "kind": "port",
        "sample": f"every {args.cpu_stride}th ray of both inclinations of the headline "
                  f"workload ({nrc} ray*channels, {tc:.1f} s), FP64 C restatement, "
                  f"skip={args.skip}, {cores} threads"
                  + (f" (cgroup CPU quota {quota:g})" if quota else ""),
    }
    ref = reference_on_host(cores)
    if ref is not None and "error" not in ref:
        ratio = ref["ref_over_oracle_rate"]
        res["reference_equivalent"] = {
            "value": nrc / tc * ratio, "cores": cores, "ref_over_oracle_rate": ratio,
            "gpu_speedup": gpu_value / (nrc / tc * ratio), "measured": ref}
    elif ref is not None:
        res["reference_equivalent"] = ref
    return res


def copy_probe(nbytes):
    """A plain device-to-device copy of known size: what FETCH_SIZE and
    WRITE_SIZE read for it calibrates the counters against true bytes."""
    import torch
    n = nbytes // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": n * 4, "copies": 3}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--precision", default="fp32", choices=["fp32", "fp64"])
    ap.add_argument("--skip", default="none", choices=["none", "reference"])
    ap.add_argument("--workload", default="c2")
    ap.add_argument("--cpu-stride", type=int, default=8,
                    help="CPU baseline traces every n-th ray of the workload")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true",
                    help="skip the forward+adjoint measurement on the 200x100 disk")
    ap.add_argument("--adjoint-only", action="store_true")
    ap.add_argument("--no-batched", action="store_true",
                    help="skip the 64-line x 8-inclination batched measurement")
    ap.add_argument("--batched-only", action="store_true")
    ap.add_argument("--no-posvel", action="store_true",
                    help="skip the position-velocity cube leg")
    ap.add_argument("--posvel-only", action="store_true")
    ap.add_argument("--no-c5", action="store_true",
                    help="skip the 1024x512-disk slice (2 lines x 4 inclinations)")
    ap.add_argument("--c5-only", action="store_true")
    ap.add_argument("--large-only", action="store_true",
                    help="only the 1024x512-disk, 1000-line plan and measured slice")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                    help="run only a plain device copy of BYTES (HBM counter calibration)")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.gpus != 1:
        print(f"bench.py: the raytracer runs on one GPU; --gpus {args.gpus} is not supported "
              "(several GPUs would each take a share of the (line, inclination) items)",
              file=sys.stderr)
        sys.exit(2)

    if args.adjoint_only:
        print(json.dumps(bench_adjoint(args)))
        return
    if args.batched_only:
        print(json.dumps(bench_batched(args)))
        return
    if args.large_only:
        print(json.dumps(bench_large(args)))
        return
    if args.c5_only:
        print(json.dumps(bench_c5_slice(args)))
        return
    if args.posvel_only:
        print(json.dumps(bench_posvel(args)))
        return
    from radlite_amd import engine, synth
    ri, w = synth.workload_inputs(args.workload, seed=0)
    nfr, vmax = engine.passband(ri.vmax, ri.dv)
    incs = [x * 0.0174532925199 for x in w["incl"]]
    skip = engine.SKIP_NONE if args.skip == "none" else engine.SKIP_REFERENCE
    rt = engine.Raytracer(0, args.precision)
    rt.load_run(ri)
    # direct launches, so that the stage events bracket the formal-solution
    # kernel alone (a replayed graph is timed as a whole); graph replay is
    # measured in the batched workload
    rt.set_option(engine.OPT_GRAPH, 0)
    items = [(0, i, nfr, vmax, skip) for i in range(len(incs))]

    def step():
        rt.setup_rays(incs, ri.nphi, ri.b_extra, ri.dbdr)   # rays + trajectories redone
        fl, _, _ = rt.trace(items)
        return fl, rt.stats()

    for _ in range(args.warmup):
        step()
    rt.L.rl_sync(rt.ctx)
    t_formal, t_total, stats = [], [], None
    t0 = time.perf_counter()
    for _ in range(args.steps):
        fl, stats = step()
        t_formal.append(stats["t_formal_ms"])
        t_total.append(stats["t_total_ms"])
    rt.L.rl_sync(rt.ctx)
    dt = (time.perf_counter() - t0) / args.steps
    nominal = sum(rt.nrays[i] * nfr for i in range(len(incs)))
    value = nominal / dt
    nodes = stats["nodes"]
    rays = stats["rays"]
    nseg = (nodes - rays) / rays                      # mean segments per ray
    # records read per segment and ray-channel without reuse across channels:
    # FP32 the split records (SegHot 16 B + NodeLine 16 B), FP64 the 64-byte
    # NodeRec
    bytes_per_rc = (32.0 if args.precision == "fp32" else 64.0) * nseg
    nbatch, tile_cw, tile_blocks = rt.batch_info()
    formal_s = float(np.mean(t_formal)) / 1e3     # per step: every batch's formal launch
    # per launch (one inclination): algorithmic bytes / mean launch duration
    flops_per_rc = 40.0 * nseg
    peak_tf = FP32_PEAK_TFLOPS if args.precision == "fp32" else FP64_PEAK_TFLOPS
    achieved_tf = flops_per_rc * nominal / formal_s / 1e12
    kname = "k_formal_batch" if tile_blocks else "k_formal_tiles"
    traffic = measured_traffic(args.precision, kname) if args.workload == "c2" else None
    hbm_gbs = traffic * nbatch / formal_s / 1e9 if traffic is not None else None

    out = {
        "metric": "ray*channel integrations/s (forward)",
        "value": value,
        "unit": "ray*channel/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if args.precision == "fp32" else "f64",
        "data": "synthetic",
        "config": {
            "workload": f"{args.workload}: nr={w['nr']} x nth={w['nth']}, {rays // len(incs)} rays"
                        f" x {nfr} ch x {len(incs)} incl, skip={args.skip}",
            "nr": w["nr"], "nth_half": w["nth"], "rays_per_incl": rays // len(incs),
            "channels": nfr, "inclinations": list(w["incl"]), "lines": 1,
            "precision": args.precision, "skip_mode": args.skip,
            "mean_segments_per_ray": nseg,
        },
        "stage_ms": {k: stats[k] for k in ("t_traj_ms", "t_nodes_ms", "t_formal_ms",
                                           "t_flux_ms", "t_total_ms")},
        "roofline": {
            # the formal kernel is bound by VALU issue and latency, not by HBM:
            # algorithmic FP32 operations (40 per segment per ray*channel) per
            # second of the kernel against the MI355X FP32 vector peak
            "bound": "valu",
            "achieved": achieved_tf,
            "peak": peak_tf,
            "unit": "TFLOP/s",
            "frac": achieved_tf / peak_tf,
            "traffic": traffic,
            "kernel": kname,
            "kernel_ms": formal_s * 1e3 / nbatch,        # mean per launch (one batch)
            "launches_per_step": nbatch,
            "items_per_launch": len(items) // nbatch,
            "flops_per_ray_channel": flops_per_rc,
            "tile_cw": tile_cw, "persistent_blocks": tile_blocks,
            # HBM, reported apart: the streaming formulation's bytes (28 B per
            # segment per ray*channel) and the counter-measured bytes per
            # launch over the kernel time, against the 8 TB/s peak
            "hbm_algorithmic_bytes_per_ray_channel": bytes_per_rc,
            "hbm_measured_gbs": hbm_gbs,
            "hbm_measured_frac": (hbm_gbs / HBM_PEAK_GBS) if hbm_gbs is not None else None,
            # measured VALU issue rate over the wave64 FP32 issue peak (SQ counters)
            "valu_issue_frac_measured": (measured_valu_issue(args.precision)
                                         if args.workload == "c2" else None),
        },
    }
    rt.close()

    if not args.no_adjoint:
        out["fwd_adjoint"] = bench_adjoint(args)
    if not args.no_batched:
        out["batched"] = bench_batched(args)
    if not args.no_posvel:
        out["posvel"] = bench_posvel(args)

    if not args.no_c5:
        out["c5_slice"] = bench_c5_slice(args)

    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args, ri, incs, skip, nfr, out["value"])
    print(json.dumps(out))


if __name__ == "__main__":
    main()
