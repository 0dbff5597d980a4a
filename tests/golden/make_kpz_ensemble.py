"""Writes tests/golden/kpz_ensemble_64.npz: the structure factor of an
ensemble of KPZ runs of the oracle (the FP32 restatement of the reference
CPU scheme), the statistics the GPU's KPZ runs are checked against.

System (examples/06_kpz/06_kpz.cpp:39-62 with these parameters):
dt h + nu q^2 h = lambda/2 (iqxh^2 + iqyh^2), iqxh = iqx h, iqyh = iqy h,
addNoise(h, "2*D"); 64^2, dx = 1, nu = 1, lambda = 3, D = 0.5, dt = 0.01,
h0 = 0, real-space noise from the Philox stream keyed by the instance's
seed (SEED0 + i).  Each of the INSTANCES runs BURN steps, then SAMPLES
snapshots EVERY steps apart; S_i(q) = sum over snapshots of |rfft2(h)|^2 / N.

    python tests/golden/make_kpz_ensemble.py      (about 3 CPU-minutes)
"""
import multiprocessing as mp
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
OUT = os.path.join(HERE, "kpz_ensemble_64.npz")

N, NU, LAM, D, DT = 64, 1.0, 3.0, 0.5, 0.01
INSTANCES, BURN, SAMPLES, EVERY, SEED0 = 16, 1000, 300, 10, 5000


def spec(i):
    sys.path.insert(0, ROOT)
    from cupss_amd.systems import spec_kpz
    return spec_kpz(N, nu=NU, lam=LAM, D=D, dt=DT, seed=SEED0 + i)


def run_instance(i, samples=SAMPLES):
    """Accumulated |rfft2(h)|^2 / N of instance i over `samples` snapshots."""
    sys.path.insert(0, ROOT)
    from cupss_amd.systems import build
    from oracle import oracle as O
    sp = dict(spec(i))
    sp["precision"] = "f32"
    s, ids = build(sp, O.System, O.pres)
    s.step(BURN)
    acc = np.zeros((N, N // 2 + 1))
    for _ in range(samples):
        s.step(EVERY)
        h = s.get_real(ids["h"]).astype(np.float64).reshape(N, N)
        acc += np.abs(np.fft.rfft2(h)) ** 2 / (N * N)
    return acc


def main():
    with mp.get_context("spawn").Pool(min(8, os.cpu_count() or 1)) as pool:
        acc = pool.map(run_instance, range(INSTANCES))
    np.savez_compressed(OUT, S_sum=np.array(acc, np.float32), samples=SAMPLES,
                        params=np.array([N, NU, LAM, D, DT, INSTANCES, BURN, SAMPLES, EVERY, SEED0], np.float64))
    print("wrote", OUT)


if __name__ == "__main__":
    main()
