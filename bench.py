"""Benchmark of the per-timestep pseudospectral update on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload ensemble]

Workload (default): the batched stochastic ensemble -- 512 independent 2D
instances (side uniform in {64,128,256,512}, Edwards-Wilkinson or KPZ with
p = 1/2, nu ~ U[0.5,2], lambda ~ U[0,2], D = 0.01, dt = 0.005, h0 = 0,
seeded), advanced together by the batched fused step with K steps per
hipGraph, against the plain per-solve path (each system advanced alone):
the configuration the batched-vs-plain throughput and HBM-roofline figures
are quoted on.  Other workloads: ch2d (2D Cahn-Hilliard 1024^2), ch3d
(256^3), diffusion (256^2), modelh (2D model H 1024^2), serve (Poisson
stream of model-H jobs: p50/p99 time-to-solution).

A "step" is one advanceTime of every instance; the timed region holds
exactly K steps with all fields resident in HBM, timed with HIP events on
the stream the kernels run on.  Printed: one JSON line with
grid-point-steps/s, the roofline entry of the dominant kernel (the column
launch that updates the state: 8 B per grid point, average duration from
HIP events around its launches in the timed run; the whole-step figure
beside it) and the CPU baseline (the oracle's FP32 restatement of the
reference CPU scheme on one core, on a bounded sample).
"""
import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
BYTES_PER_POINT_STEP = 8.0


def cahn_hilliard(S, P, dims, x0, dt=0.01, **kw):
    s = S(dims[0], dims[1], dims[2], 1.0, 1.0, 1.0, dt, **kw)
    phi = s.add_field(True)
    s.add_implicit(phi, P(1.0, q2n=1))     # "dt phi + (a*q^2 + k*q^4)*phi = -b*q^2*phi^3"
    s.add_implicit(phi, P(-1.0, q2n=2))    # a=-1, k=1, b=1; the parser negates LHS pres
    s.add_term(phi, [phi, phi, phi], [P(-1.0, q2n=1)])
    s.set_real(phi, x0)
    s.prepare()
    return s, phi


def diffusion(S, P, dims, x0, dt=0.01, **kw):
    s = S(dims[0], dims[1], dims[2], 1.0, 1.0, 1.0, dt, **kw)
    phi = s.add_field(True)
    s.add_implicit(phi, P(-1.0, q2n=1))    # "dt phi + D*q^2*phi = 0", D = 1
    s.add_term(phi, [], [P(0.0)])
    s.set_real(phi, x0)
    s.prepare()
    return s, phi


def ensemble_members(count=512, seed=0):
    """The ensemble: 512 independent 2D instances, side uniform in {64,128,256,512},
    Edwards-Wilkinson for even index (lambda = 0) and KPZ for odd index
    (lambda ~ U[0, 2]), nu ~ U[0.5, 2], D = 0.01, dt = 0.005, h0 = 0."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        n = int(rng.choice([64, 128, 256, 512]))
        nu = float(rng.uniform(0.5, 2.0))
        lam = float(rng.uniform(0.0, 2.0)) if i % 2 == 1 else 0.0
        out.append((n, nu, lam))
    return out


def kpz_system(S, P, n, nu, lam, D=0.01, dt=0.005, seed=0, **kw):
    """dt h + nu*q^2*h = 0.5*lambda*iqxh^2 + 0.5*lambda*iqyh^2, iqxh = iqx*h,
    iqyh = iqy*h, addNoise(h, "2*D"); lambda = 0 is Edwards-Wilkinson with
    the parser's "= 0" constant term."""
    s = S(n, n, 1, 1.0, 1.0, 1.0, dt, **kw)
    h = s.add_field(True)
    ix = s.add_field(False)
    iy = s.add_field(False)
    s.add_implicit(h, P(-nu, q2n=1))
    if lam != 0.0:
        s.add_term(h, [ix, ix], [P(0.5 * lam)])
        s.add_term(h, [iy, iy], [P(0.5 * lam)])
    else:
        s.add_term(h, [], [P(0.0)])
    s.add_term(ix, [h], [P(1.0, iqx=1)])
    s.add_term(iy, [h], [P(1.0, iqy=1)])
    s.set_noise(h, P(2 * D), seed)
    s.set_real(h, np.zeros(n * n, np.float32))
    s.prepare()
    return s, h


WORKLOADS = {
    "ch2d": dict(build=cahn_hilliard, dims=(1024, 1024, 1), name="2D Cahn-Hilliard 1024^2 (single instance)",
                 cpu_steps=50),
    "ch3d": dict(build=cahn_hilliard, dims=(256, 256, 256), name="3D Cahn-Hilliard 256^3 (single instance)",
                 cpu_steps=3),
    "diffusion": dict(build=diffusion, dims=(256, 256, 1), name="2D diffusion 256^2 (reference example 01)",
                      cpu_steps=2000),
    "ensemble": dict(name="batched EW/KPZ ensemble, 512 instances of 64^2..512^2", cpu_steps=0),
    "serve": dict(name="served model-H stream, Poisson arrivals, sides 128..1024", cpu_steps=0),
    "modelh": dict(build=None, dims=(1024, 1024, 1), name="2D model H 1024^2 (single instance)", cpu_steps=1),
    "grayscott": dict(build=None, dims=(1024, 1024, 1), name="2D Gray-Scott 1024^2, two dynamic fields (single instance)",
                      cpu_steps=15, dt=0.5, dynamic=2),
}


def bench_ensemble(args, C):
    """Batched (one row + one column launch per size class per step, hipGraph
    replay) against the plain per-solve path (each system advanced on its own,
    one after the other)."""
    members = ensemble_members()
    systems = []
    for i, (n, nu, lam) in enumerate(members):
        s, h = kpz_system(C.System, C.pres, n, nu, lam, seed=1000 + i)
        s.set_noise_mode(0 if args.noise == "real" else 1)
        systems.append(s)
    npts = sum(n * n for n, _, _ in members)
    # plain path: a sample of 8 systems per size, each alone
    plain_t = 0.0
    k_plain = max(10, args.steps // 10)
    for size in (64, 128, 256, 512):
        idx = [i for i, m in enumerate(members) if m[0] == size][:8]
        for i in idx:
            systems[i].step(2)
            systems[i].synchronize()
            ms = systems[i].time_steps(k_plain)
            plain_t += ms / 1e3 * (sum(1 for m in members if m[0] == size) / len(idx))
    plain_value = npts * k_plain / plain_t
    batch = C.Batch(graph_steps=args.graph)
    for s in systems:
        batch.add(s)
    batch.profile_kernel(True)
    print("bench: ensemble of %d instances, %.1fM points, noise %s, graph %d" %
          (len(members), npts / 1e6, args.noise, args.graph), file=sys.stderr, flush=True)
    setup = _batch_setup(batch, args)
    batch.step(args.warmup)
    batch.synchronize()
    ms = batch.time_steps(args.steps)
    value = npts * args.steps / (ms / 1e3)
    kernel = batch.kernel_time()
    print(batch.schedule(), file=sys.stderr, end="", flush=True)
    extra = {"plain_value": plain_value, "speedup_vs_plain": value / plain_value,
             "instances": len(members), "points": npts, "noise": args.noise,
             "graph_steps": args.graph, "setup_steps": setup}
    # the same ensemble with the reference's real-space noise draw (Philox
    # per cell, then the forward transform), timed the same way
    other = "real" if args.noise == "spectral" else "spectral"
    for s in systems:
        s.set_noise_mode(0 if other == "real" else 1)
    _batch_setup(batch, args)
    batch.step(args.warmup)
    batch.synchronize()
    ms2 = batch.time_steps(args.steps)
    extra["value_%s_noise" % other] = npts * args.steps / (ms2 / 1e3)
    extra["ms_per_step_%s_noise" % other] = ms2 / args.steps
    batch.close()
    for s in systems:
        s.close()
    return value, ms, npts, extra, kernel


def _batch_setup(batch, args, min_s=0.3):
    """Untimed set-up before the warm-up steps: the batch's first run builds
    its tables, runs the autotuner, and (graph_steps >= 2) captures and
    instantiates the K-step hipGraph, which only a run of at least K steps
    does; then runs of K steps until min_s seconds have passed, so the timed
    steps find the device at its sustained clocks rather than ramping up
    from idle.  Returns the number of steps it advanced."""
    k = args.graph if args.graph >= 2 else 1
    batch.step(k)
    batch.synchronize()
    setup = k
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < min_s and setup < 4000:
        batch.step(k)
        batch.synchronize()
        setup += k
    return setup


# HBM bytes per launch of the dominant kernel of each workload (read + written),
# measured with rocprofv3 FETCH_SIZE and WRITE_SIZE passes scaled by the same
# counters on a 256 MiB device copy (profiles/hbm_traffic_pmc.md has the
# commands and the per-pass figures); keyed by workload and kernel template
PMC_TRAFFIC = {
    ("ensemble", "k_fcols<512,"): 1.47e8 + 1.70e8,   # 69 KPZ instances of 512^2 (seed 0)
    ("ch2d", "k_fcols<1024,"): 8.79e6 + 6.31e6,
    ("ch3d", "k_f3z<256,"): 1.56e8 + 1.00e8,
}


def roofline(value, kernel, per, workload=None, dynamic=1):
    """The dominant kernel's roofline entry: the launch that performs the
    state update (one read and one write of each grid point's spectrum:
    8 B per point and dynamic field) of the class with the most points, its
    average duration from the wall-clock stamps its workgroups record inside
    the timed run."""
    ms, pts, name = kernel
    achieved = dynamic * BYTES_PER_POINT_STEP * pts / (ms / 1e3) / 1e9
    whole = dynamic * BYTES_PER_POINT_STEP * value / 1e9
    traffic = None
    for (wl, prefix), b in PMC_TRAFFIC.items():
        if wl == workload and name.replace(" ", "").startswith(prefix):
            traffic = b
    return {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "kernel": name,
            "kernel_avg_us": ms * 1e3, "kernel_points": pts,
            "whole_step_achieved": whole, "whole_step_frac": whole / HBM_PEAK_GBS, "per": per}


SERVE_SIZES = (128, 256, 512, 1024)
SERVE_WEIGHTS = (0.4, 0.3, 0.2, 0.1)
SERVE_STEPS = (1e3, 1e5)


def serve_jobs(count, seed=0):
    """The served stream's jobs: model H, side in {128,256,512,1024} with
    weights 0.4/0.3/0.2/0.1, steps log-uniform in [1e3, 1e5], phi0 ~ U(+-0.1)
    from a per-job seed."""
    rng = np.random.default_rng(seed)
    sizes = rng.choice(SERVE_SIZES, size=count, p=SERVE_WEIGHTS)
    steps = np.exp(rng.uniform(np.log(SERVE_STEPS[0]), np.log(SERVE_STEPS[1]), size=count))
    return [(int(n), int(k), 1000 + i) for i, (n, k) in enumerate(zip(sizes, steps))]


def bench_serve(args, C):
    """Poisson arrivals of model-H jobs at 80% of the measured capacity,
    served with continuous batching under a device-memory budget."""
    from cupss_amd.systems import build, spec_model_h
    est = {}
    for n in SERVE_SIZES:
        s, _ = build(spec_model_h(n, seed=1))
        est[n] = s.device_bytes()
        s.close()
    budget = int(args.budget_gib * (1 << 30))

    def submit(srv, jobs, arrivals, sink=None):
        for (n, k, seed), t in zip(jobs, arrivals):
            srv.submit(lambda n=n, seed=seed: build(spec_model_h(n, seed=seed))[0], k, arrival=t,
                       bytes_estimate=est[n], done=sink)

    # capacity: the batched step's throughput on a saturated mix (two jobs of
    # each size advanced together, steps replayed from a hipGraph)
    batch = C.Batch(graph_steps=args.quantum & ~1)
    cal = []
    for i, n in enumerate(SERVE_SIZES * 2):
        s, _ = build(spec_model_h(n, seed=7 + i))
        batch.add(s)
        cal.append(s)
    batch.step(2 * args.quantum)
    batch.synchronize()
    k = 4 * args.quantum
    ms = batch.time_steps(k)
    capacity = sum(n * n for n in SERVE_SIZES * 2) * k / (ms / 1e3)
    batch.close()
    for s in cal:
        s.close()
    mean_points = sum(w * n * n for n, w in zip(SERVE_SIZES, SERVE_WEIGHTS))
    a, b = SERVE_STEPS
    mean_steps = (b - a) / np.log(b / a)
    rate = args.load * capacity / (mean_points * mean_steps)
    jobs = serve_jobs(args.jobs)
    arrivals = np.cumsum(np.random.default_rng(1).exponential(1.0 / rate, len(jobs)))
    print("bench: serve %d model-H jobs, capacity %.3g point-steps/s, rate %.3g jobs/s (load %.2f), "
          "budget %.1f GiB" % (len(jobs), capacity, rate, args.load, args.budget_gib), file=sys.stderr, flush=True)
    srv = C.Server(budget, quantum=args.quantum, max_active=args.max_active)
    submit(srv, jobs, arrivals)
    srv.run()
    st = srv.stats()
    tts = np.array([srv.times(j)[2] - srv.times(j)[0] for j in range(len(jobs))])
    wait = np.array([srv.times(j)[1] - srv.times(j)[0] for j in range(len(jobs))])
    srv.close()
    value = st["point_steps"] / st["wall_s"]
    extra = {"jobs": len(jobs), "arrival_rate_per_s": rate, "load": args.load, "capacity": capacity,
             "p50_tts_s": float(np.percentile(tts, 50)), "p99_tts_s": float(np.percentile(tts, 99)),
             "p50_queue_s": float(np.percentile(wait, 50)), "p99_queue_s": float(np.percentile(wait, 99)),
             "wall_s": st["wall_s"], "budget_bytes": budget, "peak_bytes": st["peak_bytes"],
             "peak_active": st["peak_active"], "quantum": args.quantum}
    return value, st["wall_s"] * 1e3, extra, jobs


def cpu_baseline_serve(jobs):
    """The oracle (FP32 restatement of the reference CPU scheme, 1 core):
    one model-H step per size, extrapolated to the served jobs (the
    reference runs one job at a time)."""
    from cupss_amd.systems import spec_model_h
    from oracle import oracle as O
    from cupss_amd.systems import build
    per = {}
    for n in SERVE_SIZES:
        spec = dict(spec_model_h(n, seed=1))
        spec["precision"] = "f32"
        s, _ = build(spec, O.System, O.pres)
        s.step(1)
        t0 = time.perf_counter()
        s.step(1)
        per[n] = time.perf_counter() - t0
    total_t = sum(per[n] * k for n, k, _ in jobs)
    pts = sum(n * n * k for n, k, _ in jobs)
    return {"value": pts / total_t, "unit": "grid-point-steps/s", "cores": 1, "kind": "port",
            "sample": "one model-H step per size (128..1024) after 1 warm-up step, oracle FP32 (C2C, radix-2), "
                      "extrapolated to the served jobs' sizes and step counts, %s" % _cpu_name()}


def cpu_baseline_ensemble(members, budget_s=20.0):
    """The oracle (FP32 restatement of the reference CPU scheme, 1 core) on
    one EW and one KPZ instance of each size, a few steps each; the
    ensemble's aggregate rate is extrapolated from the per-instance step
    times (the reference advances one instance at a time)."""
    per = _cpu_sample(budget_s)
    total_t = sum(per[(n, lam != 0.0)] for n, _, lam in members)
    npts = sum(n * n for n, _, _ in members)
    return {"value": npts / total_t, "unit": "grid-point-steps/s", "cores": 1, "kind": "port",
            "sample": "one EW and one KPZ instance per size (64..512), a few steps each, oracle FP32 "
                      "(C2C, radix-2); ensemble rate = sum of points / sum of per-instance step times, %s" %
                      _cpu_name()}


def _cpu_sample(budget_s):
    """Per-instance step times of the oracle on one EW and one KPZ instance of
    each ensemble size (a worker of the all-cores baseline: a spawned process
    that loads only the oracle, never the GPU library)."""
    from oracle import oracle as O
    per = {}
    for n in (64, 128, 256, 512):
        for kpz in (False, True):
            s, h = kpz_system(O.System, O.pres, n, 1.0, 1.0 if kpz else 0.0, precision="f32")
            s.step(1)
            k = min(200, max(1, int(budget_s / 8 / (n * n * (6.0e-7 if kpz else 2.5e-7)))))
            t0 = time.perf_counter()
            s.step(k)
            per[(n, kpz)] = (time.perf_counter() - t0) / k
    return per


def cpu_baseline_ensemble_all_cores(members, budget_s=8.0):
    """The same oracle sample on every core at once (instances dealt
    round-robin over W worker processes): per-instance times measured while
    all W workers run, ensemble time = their sum over the instances / W.
    Runs before the GPU is touched, from spawned processes."""
    import multiprocessing as mp
    # at most 16 worker processes: the host-CPU share of one GPU on the
    # machines this runs on, whatever nproc reports there
    nproc = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    workers = max(1, min(16, nproc))
    with mp.get_context("spawn").Pool(workers) as pool:
        samples = pool.map(_cpu_sample, [budget_s] * workers)
    per = {key: sum(p[key] for p in samples) / workers for key in samples[0]}
    total_t = sum(per[(n, lam != 0.0)] for n, _, lam in members) / workers
    npts = sum(n * n for n, _, _ in members)
    return {"value": npts / total_t, "unit": "grid-point-steps/s", "cores": workers, "kind": "port",
            "nproc": nproc, "cpu_count": os.cpu_count(),
            "sample": "the 1-core sample run by %d processes at once (nproc %d), instances dealt round-robin, %s" %
                      (workers, nproc, _cpu_name())}


def _cpu_name():
    cpu = platform.processor() or platform.machine()
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return cpu


def cpu_baseline(wl, x0):
    """The oracle (FP32 restatement of the reference CPU scheme: full C2C
    spectra, 1 core) on a bounded sample of the same workload."""
    from oracle import oracle as O
    s, phi = wl["build"](O.System, O.pres, wl["dims"], x0, precision="f32")
    s.step(1)
    k = wl["cpu_steps"]
    t0 = time.perf_counter()
    s.step(k)
    dt = time.perf_counter() - t0
    n = wl["dims"][0] * wl["dims"][1] * wl["dims"][2]
    cpu = _cpu_name()
    return {"value": n * k / dt, "unit": "grid-point-steps/s", "cores": 1, "kind": "port",
            "sample": "%d steps of the same %s system after 1 warm-up step, oracle FP32 (C2C, radix-2), %s" %
                      (k, "x".join(map(str, wl["dims"])), cpu)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=40)
    ap.add_argument("--workload", default="ensemble", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--path", default="auto", choices=["auto", "general"],
                    help="auto: fused two-kernel step when the system qualifies; general: step program")
    ap.add_argument("--noise", default="spectral", choices=["real", "spectral"],
                    help="ensemble noise draw: real-space Philox + R2C, or spectral-space Philox")
    ap.add_argument("--graph", type=int, default=20, help="ensemble: steps per hipGraph (0 = direct launches)")
    ap.add_argument("--jobs", type=int, default=32, help="serve: number of jobs")
    ap.add_argument("--serve-jobs", type=int, default=32,
                    help="ensemble: also serve this many model-H jobs and report p50/p99 (0 = skip)")
    ap.add_argument("--load", type=float, default=0.8, help="serve: arrival rate / measured capacity")
    ap.add_argument("--budget-gib", type=float, default=4.0, help="serve: device-memory budget")
    ap.add_argument("--quantum", type=int, default=100, help="serve: steps per scheduling round")
    ap.add_argument("--max-active", type=int, default=64, help="serve: resident jobs at most")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="first copy MIB MiB device-to-device (the known-size kernel that HBM counters are scaled by)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one MI355X: this library runs on a single GPU (--gpus 1); "
              "sharding across GPUs is not implemented.", file=sys.stderr)
        sys.exit(2)

    # the all-cores CPU sample spawns worker processes: before this process
    # touches the GPU
    all_cores = None
    if args.workload == "ensemble" and not args.no_cpu:
        all_cores = cpu_baseline_ensemble_all_cores(ensemble_members())

    import cupss_amd as C
    wl = WORKLOADS[args.workload]
    if args.copy_calibration > 0:
        nb = args.copy_calibration << 20
        src, dst = C.DeviceBuffer(nb), C.DeviceBuffer(nb)
        C.check(C.lib().cupss_memcpy_d2d(dst.ptr, src.ptr, nb))
        C.check(C.lib().cupss_device_synchronize())
        del src, dst
    if args.workload == "serve":
        value, ms, extra, jobs = bench_serve(args, C)
        achieved = BYTES_PER_POINT_STEP * value / 1e9
        out = {"metric": "grid-point-steps/s", "value": value, "unit": "grid-point-steps/s", "n_gpus": 1,
               "steps": extra["jobs"], "warmup": 0, "ms_per_step": None,
               "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
               "data": "synthetic", "config": dict({"workload": wl["name"], "dt": 0.01,
                                                     "path": "server: fused batch + general program graphs"},
                                                    **extra),
               "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                            "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                            "per": "aggregate over the served run, 8 B/point-step algorithmic"}}
        if not args.no_cpu:
            out["cpu_baseline"] = cpu_baseline_serve(jobs)
        print(json.dumps(out), flush=True)
        return
    if args.workload == "ensemble":
        value, ms, npts, extra, kernel = bench_ensemble(args, C)
        out = {"metric": "grid-point-steps/s", "value": value, "unit": "grid-point-steps/s", "n_gpus": 1,
               "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms / args.steps,
               "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
               "data": "synthetic", "config": dict({"workload": wl["name"], "dt": 0.005,
                                                     "path": "batched fused step"}, **extra),
               "roofline": roofline(value, kernel, "8 B per grid point of the timed launch (the state update)",
                                    "ensemble")}
        if args.serve_jobs > 0:
            # the metric's latency half: a served model-H stream (Poisson
            # arrivals at 80% of measured capacity) run after the ensemble
            sargs = argparse.Namespace(**vars(args))
            sargs.jobs = args.serve_jobs
            svalue, sms, sextra, _ = bench_serve(sargs, C)
            out["serve"] = {"value": svalue, "unit": "grid-point-steps/s", "workload": WORKLOADS["serve"]["name"],
                            "p50_tts_s": sextra["p50_tts_s"], "p99_tts_s": sextra["p99_tts_s"],
                            "jobs": sextra["jobs"], "load": sextra["load"], "capacity": sextra["capacity"],
                            "wall_s": sextra["wall_s"]}
        if not args.no_cpu:
            out["cpu_baseline"] = cpu_baseline_ensemble(ensemble_members())
            out["cpu_baseline_all_cores"] = all_cores
        print(json.dumps(out), flush=True)
        return
    dims = wl["dims"]
    n = dims[0] * dims[1] * dims[2]
    x0 = np.random.default_rng(1).uniform(-0.1, 0.1, n).astype(np.float32)
    if args.workload in ("modelh", "grayscott"):
        from cupss_amd.systems import build, spec_gray_scott, spec_model_h
        if args.workload == "modelh":
            spec = spec_model_h(dims[0], seed=1)
            spec["init"] = {"phi": x0}
        else:
            spec = spec_gray_scott(dims[0], seed=1)
        wl["build"] = lambda S, P, dims, x0, **kw: build(dict(spec, **({"precision": kw["precision"]} if kw else {})),
                                                        S, P)
    s, phi = wl["build"](C.System, C.pres, dims, x0)
    s.set_path(0 if args.path == "auto" else 1)
    if args.path == "auto":
        s.profile_kernel(True)  # the fused paths time their dominant launch; the general program does not
    print("bench: %s, path %s, warmup %d, steps %d" % (wl["name"], args.path, args.warmup, args.steps),
          file=sys.stderr, flush=True)
    s.step(args.warmup)
    s.synchronize()
    ms = s.time_steps(args.steps)
    s.synchronize()
    kernel = None
    if args.path == "auto":
        try:
            kernel = s.kernel_time()
        except C.CupssError:
            kernel = None
    # the C++ evolver's advanceTime() loop: one step per call, host clock
    calls = max(10, min(args.steps, 200))
    t0 = time.perf_counter()
    for _ in range(calls):
        s.step(1)
    s.synchronize()
    per_call_ms = (time.perf_counter() - t0) * 1e3 / calls
    fused = kernel is not None or (args.path == "auto" and s.fused_eligible)
    if kernel is not None:
        path = "fused four-pass 3D step" if dims[2] > 1 else "fused two-kernel step"
    elif fused:
        path = "fused elementwise step (no transform, steps in registers)"
    else:
        path = "general step program"
    sec = ms / 1e3
    value = n * args.steps / sec
    ndyn = wl.get("dynamic", 1)
    achieved = ndyn * BYTES_PER_POINT_STEP * value / 1e9
    out = {
        "metric": "grid-point-steps/s",
        "value": value,
        "unit": "grid-point-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": wl["name"], "grid": list(dims), "dt": wl.get("dt", 0.01), "path": path,
                   "ms_per_step_one_step_per_call": per_call_ms},
        "roofline": (roofline(value, kernel, "%d B per grid point of the timed launch (the state update of %d "
                              "dynamic field%s)" % (8 * ndyn, ndyn, "s" if ndyn > 1 else ""), args.workload, ndyn)
                     if kernel is not None else
                     {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                      "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                      "per": "whole step (all kernels of one advanceTime), %d B/point algorithmic" % (8 * ndyn)}),
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(wl, x0)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
