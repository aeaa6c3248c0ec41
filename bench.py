"""Throughput of the fused D3Q19 MRT-LES stream-collide step on one MI355X.

Workload (the metric's configuration): channel 512x128x128, D3Q19 MRT with
d'Humieres rates, s_nu from tau=0.52, Smagorinsky Cs=0.1, FP32 populations,
square cylinder of side 16 at x=128 centred in y/z, halfway bounce-back walls
at the y/z bounds, West_Extrap inlet u=0.05 with a seeded +-1e-3 perturbation,
East_Extrap outlet.  A step is one stream-collide pass over all 8.39M nodes.

Prints one JSON line with GLUPS, the roofline fraction of the stream-collide
kernel (152 algorithmic bytes per node update in FP32, 76 in 16-bit, 304 in
FP64, over its launch time from HIP events on the library's stream) and the
CPU baseline (the oracle, OpenMP, on a bounded sample of the same workload).

Two further workloads are measured for their own profiles (not the metric's
line): --workload refined (the same channel on a 256x128x128 coarse grid
with a 128x96x96 fine patch at 2:1, two fine sub-steps per coarse step;
GLUPS counts coarse plus fine node updates) and --workload closure (the
256x64x64 channel with the learned closure fused into the collide; its
roofline is the matrix cores: 17408 algorithmic FLOP per node for the
7-64-64-64-1 network, against the dense BF16 peak).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "GLUPS and % HBM roofline (D3Q19 MRT-LES); learned vs Smagorinsky error/time"
HBM_PEAK_GBS = 8000.0
MFMA_BF16_PEAK_TFLOPS = 2500.0
MLP_FLOP_PER_NODE = 2 * (7 * 64 + 64 * 64 + 64 * 64 + 64)  # algorithmic, padding excluded
BYTES_PER_NODE = {"f32": 152, "f16": 76, "bf16": 76, "f64": 304}
# L2-to-memory request counters of one pass each (rocprofv3 --pmc allows 4
# TCC counters per pass): FETCH_SIZE's three, then WRITE_SIZE's two
TRAFFIC_PASSES = (("read", ("TCC_EA0_RDREQ_sum", "TCC_EA0_RDREQ_32B_sum", "TCC_BUBBLE_sum")),
                  ("write", ("TCC_EA0_WRREQ_sum", "TCC_EA0_WRREQ_64B_sum")))
ESIZE = {"f32": 4, "f16": 2, "bf16": 2, "f64": 8}
Q_DIRS = 19


WORKLOADS = {
    "channel": "channel {c.nx}x{c.ny}x{c.nz} D3Q19 MRT (d'Humieres rates, tau=0.52) + Smagorinsky Cs=0.1, "
               "square cylinder side 16, halfway walls, seeded inlet",
    "refined": "channel {c.nx}x{c.ny}x{c.nz} MRT-LES with a 2:1 fine patch {c.lr_size[0]}x{c.lr_size[1]}x{c.lr_size[2]} "
               "(temporal interpolation), square cylinder side 16",
    "closure": "channel {c.nx}x{c.ny}x{c.nz} MRT (tau=0.51) with the learned 7-64-64-64-1 closure fused into "
               "the collide, square cylinder side 8",
}


def cpu_baseline(cfg, budget_s=12.0):
    """The oracle on the same workload for a bounded number of steps."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import numpy as np

    try:
        import oracle_lib as O
        O.lib()
    except OSError:
        return None
    import lbm_multigpu_amd as L
    from test_gpu_parity import NODE_TO_ORC, orc_params

    c = cfg.to_c()
    with L.Solver(c) as s:
        s.set_geometry(None)
        m = s.get_mask()
        s.init_equilibrium(L.INIT_BGK)
        f = s.populations()
    lut = np.zeros(256, np.uint8)
    for k, v in NODE_TO_ORC.items():
        lut[k] = v
    mask = lut[m]
    u_in = O.inlet_profile(c.ny, c.nz, c.umax, c.inlet_amp, c.seed)
    p = orc_params(c)

    def timed(max_steps, budget):
        nonlocal f
        steps, t0 = 0, time.perf_counter()
        while True:
            f, _ = O.step(p, f, mask=mask, u_in=u_in)
            steps += 1
            if time.perf_counter() - t0 > budget or steps >= max_steps:
                break
        return steps, time.perf_counter() - t0

    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    cells = c.nx * c.ny * c.nz
    steps, dt = timed(50, budget_s)
    out = {"value": cells * steps / dt / 1e9, "unit": "GLUPS", "cores": threads, "kind": "port",
           "sample": f"oracle (C, OpenMP {threads} threads) on the full {c.nx}x{c.ny}x{c.nz} workload, "
                     f"{steps} steps in {dt:.1f} s"}
    # one thread as well (the reference's CPU program is single-threaded), and the host CPU
    try:
        import ctypes

        gomp = ctypes.CDLL("libgomp.so.1")
        gomp.omp_set_num_threads(1)
        s1, d1 = timed(3, budget_s / 2)
        gomp.omp_set_num_threads(threads)
        out["single_thread_value"] = cells * s1 / d1 / 1e9
        out["sample"] += f"; one thread: {s1} steps in {d1:.1f} s"
    except OSError:
        pass
    cpu = host_cpu()
    out["cpu_model"] = cpu.get("model name")
    out["cpu"] = cpu
    out["nproc"] = cpu["nproc"]
    return out


def host_cpu():
    """The host CPU as /proc/cpuinfo names it, and the threads this process may use."""
    info = {}
    try:
        with open("/proc/cpuinfo") as fh:
            for ln in fh:
                k, _, v = ln.partition(":")
                k, v = k.strip(), v.strip()
                if k in ("vendor_id", "model name", "cpu family", "model", "stepping", "cpu MHz", "cache size") \
                        and k not in info:
                    info[k] = v
                if k == "processor":
                    info["processors"] = info.get("processors", 0) + 1
    except OSError:
        pass
    info["nproc"] = os.cpu_count()
    info["affinity"] = len(os.sched_getaffinity(0))
    info["omp_threads"] = int(os.environ.get("OMP_NUM_THREADS", info["affinity"]))
    return info


def cpu_plan(out_path=None):
    """The CPU baseline plan: the oracle (C, OpenMP) on each workload, on all
    the host threads this process may use and on one thread: the 64^3
    Taylor-Green case in full (1000 steps), the 512x128x128 channel, the
    refined channel (256x128x128 + 128x96x96 fine, two fine sub-steps per
    coarse step) and the 1024x256x256 channel for 100 steps on all threads
    (fewer on one thread), each with its full run time extrapolated from the
    measured rate (5000, 2000 and 2000 steps).  Prints one JSON line per
    measurement and the whole record at the end."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import ctypes
    import numpy as np
    import oracle_lib as O
    import lbm_multigpu_amd as L
    from lbm_multigpu_amd import config as C
    from test_gpu_parity import NODE_TO_ORC, orc_params

    gomp = ctypes.CDLL("libgomp.so.1")
    cpu = host_cpu()
    nthr = cpu["omp_threads"]
    lut = np.zeros(256, np.uint8)
    for k, v in NODE_TO_ORC.items():
        lut[k] = v
    rec = {"cpu": cpu, "runs": []}
    print(json.dumps({"cpu": cpu}), flush=True)

    def state(cfg, tgv=False):
        with L.Solver(cfg.to_c()) as s:
            s.set_geometry(None)
            if tgv:
                s.init_taylor_green(0.05)
            else:
                s.init_equilibrium(L.INIT_BGK)
            out = [lut[s.get_mask(0)], s.populations(0)]
            if cfg.refinement:
                out += [lut[s.get_mask(1)], s.populations(1)]
        return out

    def runner(p, f, m, u_in=None):
        """The oracle's time loop on preallocated lattices (no copies timed)."""
        fa = np.ascontiguousarray(f).copy()
        fb = np.empty_like(fa)
        fn = O.lib().orc_run_f32
        uin = None if u_in is None else u_in.ctypes.data
        return lambda k: fn(ctypes.byref(p), m.ctypes.data, uin, fa, fb, k, None)

    def timed(fn, steps, threads):
        gomp.omp_set_num_threads(threads)
        t0 = time.perf_counter()
        fn(steps)
        dt = time.perf_counter() - t0
        gomp.omp_set_num_threads(nthr)
        return dt

    def report(name, updates, steps, threads, dt, full_steps):
        r = {"workload": name, "threads": threads, "steps": steps, "seconds": round(dt, 3),
             "mlups": round(updates * steps / dt / 1e6, 2)}
        if full_steps != steps:
            r["full_steps"] = full_steps
            r["extrapolated_full_run_s"] = round(dt / steps * full_steps, 1)
        rec["runs"].append(r)
        print(json.dumps(r), flush=True)

    # config 1: Taylor-Green 64^3, BGK tau 0.6, 1000 steps, in full
    cfg = C.taylor_green()
    m, f = state(cfg, tgv=True)
    p = orc_params(cfg.to_c())
    n1 = cfg.nx * cfg.ny * cfg.nz
    for th in (nthr, 1):
        report("taylor_green 64^3 BGK", n1, 1000, th, timed(runner(p, f, m), 1000, th), 1000)
    # configs 2 and 4: the channel at 512x128x128 and 1024x256x256
    for name, cfg, full, one in (("channel 512x128x128", C.channel(), 5000, 10),
                                 ("channel 1024x256x256", C.channel(nx=1024, ny=256, nz=256, side=32, x_center=256), 2000, 2)):
        m, f = state(cfg)
        c = cfg.to_c()
        p = orc_params(c)
        u_in = O.inlet_profile(c.ny, c.nz, c.umax, c.inlet_amp, c.seed)
        n = cfg.nx * cfg.ny * cfg.nz
        for th, k in ((nthr, 100), (1, one)):
            report(name, n, k, th, timed(runner(p, f, m, u_in), k, th), full)
        del f
    # config 3: the refined channel, coarse plus both fine sub-steps per step
    cfg = C.refined_channel()
    cm, cf, fm, ff = state(cfg)
    c = cfg.to_c()
    pc = orc_params(c)
    r = O.refine([int(round(x - 0.25)) for x in cfg.lr_loc], cfg.lr_size, float(np.float32(c.omega)),
                 cs=float(np.float32(c.cs)), temporal=1, eq_kind=O.EQ_BGK_MEQ)
    pf = O.params(*cfg.lr_size, bc=(O.CLAMP,) * 3, collision=pc.collision, halfway=1, omega=r.omega_f,
                  cs=pc.cs, dpdy=0.5 * pc.dpdy)
    u_in = O.inlet_profile(c.ny, c.nz, c.umax, c.inlet_amp, c.seed)
    n3 = cfg.nx * cfg.ny * cfg.nz + 2 * int(np.prod(cfg.lr_size))
    for th, k in ((nthr, 100), (1, 3)):
        report("refined channel 256x128x128 + 128x96x96", n3, k, th,
               timed(lambda s: O.refine_run(pc, pf, r, cf, ff, s, cmask=cm, fmask=fm, u_in=u_in), k, th), 2000)
    print(json.dumps(rec), flush=True)
    if out_path:
        with open(out_path, "w") as fh:
            json.dump(rec, fh, indent=1)


def copy_calibration(nbytes=637_534_208, reps=20):
    """Device-to-device copy of nbytes (the 19 x 4 B of FP32 populations of
    the workload's 512x128x128 nodes: the stream-collide kernel reads and
    writes this much per launch), timed with
    HIP events: the practical HBM ceiling, and the reference kernel for
    scaling rocprof's FETCH_SIZE/WRITE_SIZE counters."""
    import ctypes as C

    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)) == 0
    hip.hipMemset(a, 1, C.c_size_t(nbytes))
    e0, e1 = C.c_void_p(), C.c_void_p()
    hip.hipEventCreate(C.byref(e0))
    hip.hipEventCreate(C.byref(e1))
    for _ in range(3):
        hip.hipMemcpyAsync(b, a, C.c_size_t(nbytes), 3, None)
    hip.hipEventRecord(e0, None)
    for _ in range(reps):
        hip.hipMemcpyAsync(b, a, C.c_size_t(nbytes), 3, None)
    hip.hipEventRecord(e1, None)
    hip.hipEventSynchronize(e1)
    ms = C.c_float()
    hip.hipEventElapsedTime(C.byref(ms), e0, e1)
    hip.hipFree(a)
    hip.hipFree(b)
    per = ms.value / reps
    gbs = 2 * nbytes / (per * 1e-3) / 1e9
    print(json.dumps({"copy_bytes": nbytes, "copy_ms": round(per, 4), "copy_GBps_read_plus_write": round(gbs, 1),
                      "frac_of_8TBps": round(gbs / HBM_PEAK_GBS, 4)}), flush=True)


def traffic_probe(storage):
    """The program the counter passes run (under rocprofv3 --pmc): three
    device copies of the bytes the kernel reads per launch (the counters'
    reference of known size), then the channel's stream-collide in the given
    storage, 3 steps after 5 of warm-up."""
    import lbm_multigpu_amd as L
    from lbm_multigpu_amd import config as C

    cfg = C.channel(storage=storage)
    copy_calibration(Q_DIRS * ESIZE[storage] * cfg.nx * cfg.ny * cfg.nz, reps=3)
    with L.Solver(cfg.to_c()) as s:
        s.set_geometry(None)
        s.init_equilibrium(L.INIT_BGK)
        s.step(5)
        s.step_timed(3)


def _counter_rows(out_dir):
    """(kernel name, dispatch id, counter, value) rows of a rocprofv3
    counter-collection CSV under out_dir."""
    import csv
    import glob

    files = glob.glob(os.path.join(out_dir, "**", "*counter_collection.csv"), recursive=True)
    if not files:
        raise RuntimeError("no counter_collection.csv under " + out_dir)
    rows = []
    with open(files[0]) as fh:
        for r in csv.DictReader(fh):
            rows.append((r["Kernel_Name"], int(r["Dispatch_Id"]), r["Counter_Name"], float(r["Counter_Value"])))
    return rows


def _per_dispatch(rows, match):
    """{counter: value} of each dispatch whose kernel name contains match, in dispatch order."""
    d = {}
    for name, disp, ctr, val in rows:
        if match in name:
            d.setdefault(disp, {})[ctr] = d.get(disp, {}).get(ctr, 0.0) + val
    return [d[k] for k in sorted(d)]


def measure_traffic(storage, timeout_s=180):
    """HBM bytes per stream-collide launch from the L2-to-memory request
    counters, in two rocprofv3 --pmc passes over traffic_probe() run as child
    processes (before this process touches the GPU).

    Reads: FETCH_SIZE's terms, 32 B x TCC_EA0_RDREQ_32B + 128 B x TCC_BUBBLE
    + b x the remaining TCC_EA0_RDREQ, where FETCH_SIZE takes b = 64 B; b is
    instead solved from the copy of known size in the same pass (it reads
    128 B per request on this part, where FETCH_SIZE reads half the bytes).
    Writes: WRITE_SIZE's terms, 64 B x TCC_EA0_WRREQ_64B + 32 B x the rest,
    checked against the copy's known bytes (the ratio is reported).
    Returns (result dict, None) or (None, reason)."""
    import shutil
    import subprocess
    import tempfile

    prof = shutil.which("rocprofv3")
    if not prof:
        return None, "rocprofv3 not on PATH"
    nodes = 512 * 128 * 128
    copy_bytes = Q_DIRS * ESIZE[storage] * nodes
    res = {"kernel": f"k_stream_collide ({storage} channel 512x128x128)", "copy_bytes": copy_bytes,
           "algorithmic_bytes": BYTES_PER_NODE[storage] * nodes, "commands": []}
    with tempfile.TemporaryDirectory() as tmp:
        for tag, ctrs in TRAFFIC_PASSES:
            out = os.path.join(tmp, tag)
            cmd = [prof, "--pmc", *ctrs, "--output-format", "csv", "-d", out, "-o", tag, "--",
                   sys.executable, os.path.abspath(__file__), "--traffic-probe", "--storage", storage]
            res["commands"].append(" ".join(["rocprofv3", "--pmc", *ctrs, "--output-format", "csv", "-d", "<dir>",
                                             "-o", tag, "--", "python3", "bench.py", "--traffic-probe",
                                             "--storage", storage]))
            try:
                r = subprocess.run(cmd, capture_output=True, text=True, timeout=timeout_s, cwd=ROOT)
            except subprocess.TimeoutExpired:
                return None, f"{tag} pass timed out"
            try:
                rows = _counter_rows(out)
            except (RuntimeError, KeyError, ValueError) as e:
                return None, f"{tag} pass (rc {r.returncode}): {e}: {r.stderr[-300:]}"
            copies = _per_dispatch(rows, "copyBuffer")
            kern = _per_dispatch(rows, "k_stream_collide")
            if not copies or len(kern) < 3:
                return None, f"{tag} pass: {len(copies)} copy and {len(kern)} kernel dispatches"
            c, k = copies[-1], kern[-1]
            if tag == "read":
                R, R32, B = (ctrs[0], ctrs[1], ctrs[2])
                other_c = c[R] - c[R32] - c[B]
                b = (copy_bytes - 32 * c[R32] - 128 * c[B]) / other_c if other_c > 0 else 64.0
                rd = 32 * k[R32] + 128 * k[B] + b * (k[R] - k[R32] - k[B])
                res.update(read_bytes=round(rd), read_request_bytes=round(b, 2),
                           fetch_size_kib=round((32 * k[R32] + 128 * k[B] + 64 * (k[R] - k[R32] - k[B])) / 1024, 1),
                           read_counters={n: k[n] for n in ctrs}, copy_read_counters={n: c[n] for n in ctrs})
            else:
                W, W64 = ctrs
                wr = 64 * k[W64] + 32 * (k[W] - k[W64])
                wc = 64 * c[W64] + 32 * (c[W] - c[W64])
                res.update(write_bytes=round(wr), write_size_kib=round(wr / 1024, 1),
                           copy_write_ratio=round(wc / copy_bytes, 4),
                           write_counters={n: k[n] for n in ctrs}, copy_write_counters={n: c[n] for n in ctrs})
    res["bytes_per_launch"] = res["read_bytes"] + res["write_bytes"]
    res["ratio_to_algorithmic"] = round(res["bytes_per_launch"] / res["algorithmic_bytes"], 4)
    return res, None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--storage", default="f32", choices=["f32", "f16", "bf16", "f64"])
    ap.add_argument("--workload", default="channel", choices=["channel", "refined", "closure"])
    ap.add_argument("--closure-every", type=int, default=1,
                    help="closure workload: evaluate the network every N steps")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-secondary", action="store_true", help="skip the FP16-storage secondary figure")
    ap.add_argument("--forces", action="store_true",
                    help="record the momentum-exchange force every step (timed with the steps)")
    ap.add_argument("--traffic-probe", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--no-traffic", action="store_true",
                    help="skip the two rocprofv3 counter passes that measure the kernel's HBM traffic")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="time a plain device copy of one lattice instead (HBM ceiling / counter scale)")
    ap.add_argument("--cpu-plan", metavar="OUT", nargs="?", const="",
                    help="run only the CPU baseline plan over every workload (several minutes)")
    a = ap.parse_args()
    if a.traffic_probe:
        traffic_probe(a.storage)
        return
    if a.calibrate_copy:
        copy_calibration()
        return
    if a.cpu_plan is not None:
        cpu_plan(a.cpu_plan or None)
        return
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {a.gpus}); the sharded path is a design, not a build.",
              file=sys.stderr)
        sys.exit(2)

    # the counter passes run first, as child processes, before this process
    # initialises the GPU
    traffic, traffic_err = None, None
    if a.workload == "channel" and not a.no_traffic:
        traffic, traffic_err = measure_traffic(a.storage)

    import lbm_multigpu_amd as L
    from lbm_multigpu_amd import config as C

    def make_cfg(storage, workload=None):
        workload = workload or a.workload
        if workload == "refined":
            return C.refined_channel(storage=storage)
        if workload == "closure":
            return C.channel(nx=256, ny=64, nz=64, side=8, x_center=64, tau=0.51, storage=storage)
        return C.channel(storage=storage)

    def measure(cfg, forces, warmup=None, workload=None, closure_every=None):
        """(ms per step from HIP events on the library's stream, wall s, FP64
        mass sum, fine-level sizes or None)"""
        warmup = a.warmup if warmup is None else warmup
        workload = workload or a.workload
        closure_every = a.closure_every if closure_every is None else closure_every
        if forces:
            # record the momentum-exchange force of every step, as the
            # reference does from STARTF on
            cfg.force_steps = a.warmup + a.steps + 2
            cfg.startf = 0
        with L.Solver(cfg.to_c()) as s:
            s.set_geometry(None)
            s.init_equilibrium(L.INIT_BGK)
            mlp = None
            if workload == "closure":
                # a seeded network over the whole input range: every fluid node
                # runs the network (the cost does not depend on the weights)
                import numpy as np
                mlp = L.Mlp(seed=1)
                big = np.full(7, 3e38, np.float32)
                mlp.set_norm(np.zeros(7, np.float32), np.ones(7, np.float32), -big, big, 1e-4)
                s.set_closure(mlp, every=closure_every)
            s.step(warmup)
            s.step_timed(1)  # drains the warm-up
            t0 = time.perf_counter()
            ms = s.step_timed(a.steps)
            wall = time.perf_counter() - t0
            mm = s.mass_momentum()
            fine = s.sizes(1) if workload == "refined" else None
            if forces:
                assert len(s.forces()) >= a.steps
        return ms / a.steps, wall, mm, fine

    cfg = make_cfg(a.storage)
    ms_step, wall, mm, fine = measure(cfg, a.forces)
    cells = cfg.nx * cfg.ny * cfg.nz
    updates = cells + (2 * fine.nx * fine.ny * fine.nz if fine is not None else 0)
    glups = updates / (ms_step * 1e-3) / 1e9
    if a.workload == "closure":
        # the network runs on one step in closure_every: its flops per step
        ach = MLP_FLOP_PER_NODE * cells / max(a.closure_every, 1) / (ms_step * 1e-3) / 1e12
        roof = {"bound": "mfma", "achieved": round(ach, 2), "peak": MFMA_BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": round(ach / MFMA_BF16_PEAK_TFLOPS, 4), "traffic": None}
    else:
        ach = BYTES_PER_NODE[a.storage] * updates / (ms_step * 1e-3) / 1e9
        roof = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": traffic["bytes_per_launch"] if traffic else None}
    # the 16-bit storage variants of the same workload and the refinement and
    # closure workloads, measured in the same run as secondary figures
    secondary = secondary_wl = None
    if a.workload == "channel" and a.storage == "f32" and not a.no_secondary and not a.forces:
        sec = {}
        for st in ("f16", "bf16"):
            # past the start-up transient: the 16-bit kernels run up to 25 %
            # slower over the first ~150 steps from the uniform start (127
            # us/step after it on the 512x128x128 channel, 133-160 us within
            # it); FP32 does not vary
            ms2, _, _, _ = measure(make_cfg(st), False, warmup=max(a.warmup, 200))
            g2 = cells / (ms2 * 1e-3) / 1e9
            ach2 = BYTES_PER_NODE[st] * cells / (ms2 * 1e-3) / 1e9
            sec[st] = {"value": round(g2, 3), "unit": "GLUPS", "ms_per_step": round(ms2, 4),
                       "roofline_frac": round(ach2 / HBM_PEAK_GBS, 4)}
        secondary = sec
        wl = {}
        # refined channel: coarse plus both fine sub-steps per coarse step
        rc = make_cfg("f32", "refined")
        ms3, _, _, f3 = measure(rc, False, workload="refined")
        up3 = rc.nx * rc.ny * rc.nz + 2 * f3.nx * f3.ny * f3.nz
        wl["refined"] = {"workload": WORKLOADS["refined"].format(c=rc), "value": round(up3 / (ms3 * 1e-3) / 1e9, 3),
                         "unit": "GLUPS", "ms_per_step": round(ms3, 4),
                         "roofline_frac": round(BYTES_PER_NODE["f32"] * up3 / (ms3 * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
        # learned closure fused into the collide, every step and every 10 steps
        cc = make_cfg("f32", "closure")
        n4 = cc.nx * cc.ny * cc.nz
        for every in (1, 10):
            ms4, _, _, _ = measure(make_cfg("f32", "closure"), False, workload="closure", closure_every=every)
            ach4 = MLP_FLOP_PER_NODE * n4 / every / (ms4 * 1e-3) / 1e12
            wl[f"closure_every_{every}"] = {"workload": WORKLOADS["closure"].format(c=cc), "closure_every": every,
                                            "value": round(n4 / (ms4 * 1e-3) / 1e9, 3), "unit": "GLUPS",
                                            "ms_per_step": round(ms4, 4),
                                            "mfma_frac": round(ach4 / MFMA_BF16_PEAK_TFLOPS, 4)}
        secondary_wl = wl
    cpu = None if (a.no_cpu or a.workload != "channel") else cpu_baseline(cfg)
    line = {
        "metric": METRIC, "value": round(glups, 3), "unit": "GLUPS", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": round(ms_step, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": a.storage, "data": "synthetic",
        "config": {"workload": WORKLOADS[a.workload].format(c=cfg), "cells": cells, "node_updates_per_step": updates,
                   "storage": a.storage, "wall_s": round(wall, 3), "mass_sum": mm[0],
                   **({"closure_every": a.closure_every} if a.workload == "closure" else {})},
        "roofline": roof, "cpu_baseline": cpu,
    }
    if secondary:
        line["secondary_storage"] = secondary
    if secondary_wl:
        line["secondary_workloads"] = secondary_wl
    if traffic or traffic_err:
        line["traffic_counters"] = traffic if traffic else {"error": traffic_err}
    if a.forces:
        line["config"]["forces_recorded_every_step"] = True
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
