"""Benchmark of the per-timestep pseudospectral update on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload ch2d]

Workload (default): 2D Cahn-Hilliard, 1024x1024, FP32, a=-1, b=1, k=1,
dt=0.01, reference dealiasing (|n| <= N/4 for phi^3), phi0 ~ U(-0.1, 0.1)
from seed 1 -- the single-instance configuration the headline metric is
quoted on.  A "step" is one evolver::advanceTime of that system; the timed
region holds exactly K steps with the field resident in HBM, bracketed by
device synchronisation, timed with HIP events on the stream the kernels run
on.

Printed: one JSON line with grid-point-steps/s, the HBM-roofline fraction
of the algorithmic traffic (8 B per grid point per step: one FP32 read and
one FP32 write of the dynamic field) and the CPU baseline (the oracle's FP32
restatement of the reference CPU scheme on one core, on a bounded sample).
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


WORKLOADS = {
    "ch2d": dict(build=cahn_hilliard, dims=(1024, 1024, 1), name="2D Cahn-Hilliard 1024^2 (single instance)",
                 cpu_steps=30),
    "ch3d": dict(build=cahn_hilliard, dims=(256, 256, 256), name="3D Cahn-Hilliard 256^3 (single instance)",
                 cpu_steps=1),
    "diffusion": dict(build=diffusion, dims=(256, 256, 1), name="2D diffusion 256^2 (reference example 01)",
                      cpu_steps=2000),
}


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
    cpu = platform.processor() or platform.machine()
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    cpu = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"value": n * k / dt, "unit": "grid-point-steps/s", "cores": 1, "kind": "port",
            "sample": "%d steps of the same %s system after 1 warm-up step, oracle FP32 (C2C, radix-2), %s" %
                      (k, "x".join(map(str, wl["dims"])), cpu)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--workload", default="ch2d", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--path", default="auto", choices=["auto", "general"],
                    help="auto: fused two-kernel step when the system qualifies; general: step program")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one MI355X: this build runs on a single GPU (--gpus 1); "
              "multi-GPU sharding is a design note only.", file=sys.stderr)
        sys.exit(2)

    import cupss_amd as C
    wl = WORKLOADS[args.workload]
    dims = wl["dims"]
    n = dims[0] * dims[1] * dims[2]
    x0 = np.random.default_rng(1).uniform(-0.1, 0.1, n).astype(np.float32)
    s, phi = wl["build"](C.System, C.pres, dims, x0)
    s.set_path(0 if args.path == "auto" else 1)
    path = "fused two-kernel step" if (args.path == "auto" and s.fused_eligible) else "general step program"
    print("bench: %s, %s, warmup %d, steps %d" % (wl["name"], path, args.warmup, args.steps), file=sys.stderr,
          flush=True)
    s.step(args.warmup)
    s.synchronize()
    ms = s.time_steps(args.steps)
    s.synchronize()
    sec = ms / 1e3
    value = n * args.steps / sec
    achieved = BYTES_PER_POINT_STEP * value / 1e9
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
        "config": {"workload": wl["name"], "grid": list(dims), "dt": 0.01, "path": path},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "per": "whole step (all kernels of one advanceTime), 8 B/point algorithmic"},
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(wl, x0)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
