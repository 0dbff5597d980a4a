"""GPU: the batched fused path.

* Every system advanced inside a batch of mixed sizes, models and
  parameters ends bit-identical to the same system advanced alone (same
  kernels, same per-system parameters and noise streams).
* hipGraph replay equals direct launches bit for bit.
* The full 512-instance ensemble (configuration C4): sampled members of
  every size and model equal the same systems advanced alone.
* Edwards-Wilkinson: the stationary structure factor of the scheme,
  S(q) = <|h_q|^2>/N = D / (dV nu q^2 (1 + dt nu q^2 / 2)), is met by both
  noise draws (real-space Philox + R2C, and the spectral-space draw).
* KPZ: the two noise draws give the same structure factor, and both match
  the structure factor of an oracle ensemble within 1 %.
"""
import os

import numpy as np
import pytest

import cupss_amd as C
from tests.helpers import build_system, qgrid, rel_l2, spec_cahn_hilliard, spec_kpz, spec_model_h

pytestmark = pytest.mark.gpu


def make(spec, noise_mode=0):
    g, ids = build_system(C.System, C.pres, spec)
    g.set_noise_mode(noise_mode)
    return g, ids


def mixed_specs():
    rng = np.random.default_rng(4)
    specs = []
    for i, n in enumerate([64, 128, 64, 256, 128, 32, 512, 64]):
        if i % 3 == 0:
            specs.append(spec_cahn_hilliard((n, n, 1), seed=i))
        else:
            lam = 0.0 if i % 3 == 1 else float(rng.uniform(0, 2))
            specs.append(spec_kpz(n, nu=float(rng.uniform(0.5, 2)), lam=lam, D=0.01, dt=0.005, seed=100 + i))
    # model H: stored constraint fields (sigma, v, w) beside derived ones
    specs.append(spec_model_h(64, seed=41))
    specs.append(spec_model_h(128, seed=42))
    return specs


@pytest.mark.parametrize("noise_mode", [0, 1])
@pytest.mark.parametrize("graph", [0, 4])
def test_batch_equals_alone(noise_mode, graph):
    specs = mixed_specs()
    print("batch", len(specs), "noise mode", noise_mode, "graph", graph, flush=True)
    alone = []
    for sp in specs:
        g, ids = make(sp, noise_mode)
        g.step(23)
        alone.append(g.get_real(ids["phi" if "phi" in ids else "h"]))
    batch = C.Batch(graph_steps=graph)
    members = []
    for sp in specs:
        g, ids = make(sp, noise_mode)
        assert g.fused_eligible
        batch.add(g)
        members.append((g, ids))
    batch.step(10)
    batch.step(13)
    bad = []
    for i, ((g, ids), ref) in enumerate(zip(members, alone)):
        got = g.get_real(ids["phi" if "phi" in ids else "h"])
        d = float(np.max(np.abs(got - ref)))
        print("  member %d %s: max|batch - alone| = %.3e" % (i, g.dims, d))
        if not np.array_equal(got, ref):
            bad.append(i)
        assert g.step_count == 23
    assert not bad


def test_odd_runs_around_graph_replays_equal_alone():
    """Runs of odd length swap the roles of each member's two spectrum
    buffers (and of the stored fields' pairs) on the device; graph replays,
    direct launches and the settle that follow must find the current state
    in the buffer they start from.  5 (replay + 1 direct), 8 (two replays
    from the swapped pair), 3 (direct only), 7 (replay + 3 direct) steps
    against 23 steps alone, every field of every member bit for bit."""
    specs = mixed_specs()
    alone = []
    for sp in specs:
        g, ids = make(sp, 1)
        g.step(23)
        alone.append({k: g.get_real(v) for k, v in ids.items()})
    batch = C.Batch(graph_steps=4)
    members = []
    for sp in specs:
        g, ids = make(sp, 1)
        batch.add(g)
        members.append((g, ids))
    for n in (5, 8, 3, 7):
        print("batch run of", n, "steps", flush=True)
        batch.step(n)
    bad = []
    for i, ((g, ids), ref) in enumerate(zip(members, alone)):
        assert g.step_count == 23
        for k, v in ids.items():
            got = g.get_real(v)
            if not np.array_equal(got, ref[k]):
                bad.append((i, k, float(np.max(np.abs(got - ref[k])))))
    print("mismatches", bad)
    assert not bad


def test_split_runs_equal_one_run():
    """Advancing 10 + 13 steps, or one step per call (the C++ evolver's
    advanceTime), equals advancing 23: runs that resume from the previous
    run's column-major state and real fields, and runs after a state read
    or a state write (which start over from the natural spectrum)."""
    for sp in (spec_cahn_hilliard((64, 64, 1), seed=2), spec_kpz(64, nu=1.1, lam=1.3, D=0.01, dt=0.005, seed=5),
               spec_model_h(64, seed=6)):
        key = "phi" if "phi" in dict(sp["fields"]) else "h"
        a, ia = make(sp)
        a.step(23)
        ref = a.get_real(ia[key])
        b, ib = make(sp)
        b.step(10)
        b.step(13)
        assert np.array_equal(ref, b.get_real(ib[key]))
        c, ic = make(sp)
        for i in range(23):
            c.step(1)
            if i == 6:
                c.get_real(ic[key])
            if i == 15:  # write the state back unchanged: the next run starts over
                c.set_real(ic[key], c.get_real(ic[key]))
        got = c.get_real(ic[key])
        assert float(np.max(np.abs(got - ref))) <= 1e-5 * max(1.0, float(np.max(np.abs(ref))))
        d, idd = make(sp)
        for i in range(23):
            d.step(1)
        assert np.array_equal(ref, d.get_real(idd[key]))


def test_members_stepped_alone_and_removed():
    """A member advanced on its own between batch runs, a member read and
    one removed in the middle: every state ends bit-identical to the same
    system advanced alone.  (Batch runs defer the transpose of the state
    back to the natural layout until something reads or advances a member
    elsewhere; these are the paths that must write it back first.)"""
    specs = [spec_kpz(64, nu=1.2, lam=0.9, D=0.01, dt=0.005, seed=61),
             spec_cahn_hilliard((128, 128, 1), seed=62), spec_model_h(64, seed=63)]
    keys = ["h", "phi", "phi"]
    for graph in (0, 4):
        batch = C.Batch(graph_steps=graph)
        mem = []
        for sp in specs:
            g, ids = make(sp)
            batch.add(g)
            mem.append((g, ids))
        batch.step(6)
        mem[0][0].step(3)                      # alone, while still a member
        batch.step(4)
        mid = mem[1][0].get_real(mem[1][1]["phi"])
        batch.step(2)
        batch.remove(mem[2][0])
        mem[2][0].step(5)                      # alone after leaving
        batch.step(4)
        counts = [6 + 3 + 4 + 2 + 4, 6 + 4 + 2 + 4, 6 + 4 + 2 + 5]
        for (g, ids), sp, key, n in zip(mem, specs, keys, counts):
            ref, rids = make(sp)
            ref.step(n)
            got = g.get_real(ids[key])
            print("  graph %d %s: %d steps, max|diff| %.3e" % (graph, g.dims, n, float(np.max(np.abs(got - ref.get_real(rids[key]))))))
            assert g.step_count == n
            assert np.array_equal(got, ref.get_real(rids[key]))
        ref, rids = make(specs[1])
        ref.step(10)
        assert np.array_equal(mid, ref.get_real(rids["phi"]))
        batch.close()


def ew_structure_factor(n, nu, D, dt, noise_mode, steps, burn, every=1):
    sp = spec_kpz(n, nu=nu, lam=0.0, D=D, dt=dt, seed=11)
    g, ids = make(sp, noise_mode)
    g.step(burn)
    acc = np.zeros((n, n // 2 + 1))
    m = 0
    for _ in range(steps // every):
        g.step(every)
        h = g.get_real(ids["h"]).astype(np.float64).reshape(n, n)
        acc += np.abs(np.fft.rfft2(h)) ** 2
        m += 1
    return acc / m / (n * n)


@pytest.mark.parametrize("noise_mode", [0, 1])
def test_edwards_wilkinson_structure_factor(noise_mode):
    n, nu, D, dt = 64, 1.0, 0.01, 0.05
    S = ew_structure_factor(n, nu, D, dt, noise_mode, steps=4000, burn=1500)
    qx, qy, _ = qgrid(n, n)
    q2 = (qx ** 2 + qy ** 2)[0][:, : n // 2 + 1]
    dtf = float(np.float32(dt))
    theory = np.where(q2 > 0, D / (nu * np.maximum(q2, 1e-30) * (1 + dtf * nu * q2 / 2)), 0.0)
    fast = q2 > 0.3       # modes that relax within the burn-in
    ratio = S[fast] / theory[fast]
    print("  EW S(q)/theory over %d modes: mean %.4f" % (fast.sum(), ratio.mean()))
    assert abs(ratio.mean() - 1.0) < 0.01


def test_kpz_structure_factor_same_for_both_noise_draws():
    n = 128
    res = []
    for mode in (0, 1):
        sp = spec_kpz(n, nu=1.0, lam=1.0, D=0.05, dt=0.02, seed=21)
        g, ids = make(sp, mode)
        g.step(1000)
        acc = np.zeros((n, n // 2 + 1))
        for _ in range(1500):
            g.step(2)
            h = g.get_real(ids["h"]).astype(np.float64).reshape(n, n)
            acc += np.abs(np.fft.rfft2(h)) ** 2
        res.append(acc / 1500 / (n * n))
    qx, qy, _ = qgrid(n, n)
    q2 = (qx ** 2 + qy ** 2)[0][:, : n // 2 + 1]
    fast = q2 > 0.5
    r = res[1][fast].sum() / res[0][fast].sum()
    print("  KPZ S(q) spectral / real-space draw: %.4f" % r)
    assert abs(r - 1.0) < 0.01


@pytest.mark.parametrize("noise_mode", [0, 1])
def test_kpz_structure_factor_matches_oracle_ensemble(noise_mode):
    """KPZ S(q) of a GPU ensemble against the oracle's ensemble
    (tests/golden/kpz_ensemble_64.npz: 16 instances of 64^2, nu 1, lambda
    3, D 0.5, dt 0.01, 300 snapshots 10 steps apart after 1000), the same
    instances advanced in one batch.  The aggregate over the modes with
    q^2 > 0.5 agrees within 1 % and each shell within 3 %.  Real-space draw
    (mode 0): the oracle's own Philox stream.  Spectral draw (mode 1): an
    independent stream of the same law, so the agreement is statistical
    (standard error of either ensemble about 0.2 %); at these parameters
    KPZ lies 6 % above Edwards-Wilkinson, so the check can fail."""
    from tests.golden import make_kpz_ensemble as M
    z = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "kpz_ensemble_64.npz"))
    ref = z["S_sum"].astype(np.float64).sum(axis=0) / (int(z["samples"]) * M.INSTANCES)
    batch = C.Batch(graph_steps=0)
    members = []
    for i in range(M.INSTANCES):
        g, ids = make(M.spec(i), noise_mode)
        batch.add(g)
        members.append((g, ids))
    print("kpz ensemble", M.INSTANCES, "x", M.N, "noise mode", noise_mode, flush=True)
    batch.step(M.BURN)
    acc = np.zeros((M.N, M.N // 2 + 1))
    for _ in range(M.SAMPLES):
        batch.step(M.EVERY)
        for g, ids in members:
            h = g.get_real(ids["h"]).astype(np.float64).reshape(M.N, M.N)
            acc += np.abs(np.fft.rfft2(h)) ** 2 / (M.N * M.N)
    got = acc / (M.SAMPLES * M.INSTANCES)
    batch.close()
    qx, qy, _ = qgrid(M.N, M.N)
    q2 = (qx ** 2 + qy ** 2)[0][:, : M.N // 2 + 1]
    fast = q2 > 0.5
    r = got[fast].sum() / ref[fast].sum()
    print("  S(q) GPU / oracle ensemble over %d modes: %.4f" % (fast.sum(), r))
    assert abs(r - 1.0) < 0.01
    for lo, hi in ((0.5, 1.0), (1.0, 2.0), (2.0, 4.0), (4.0, 1e9)):
        band = (q2 > lo) & (q2 <= hi)
        rb = got[band].sum() / ref[band].sum()
        print("  shell %.1f < q^2 <= %.1f (%d modes): %.4f" % (lo, hi, band.sum(), rb))
        assert abs(rb - 1.0) < 0.03


def full_ensemble_specs(count=512, seed=0):
    """Configuration C4: 512 independent 2D instances, side uniform in
    {64, 128, 256, 512}; Edwards-Wilkinson at even index (lambda = 0), KPZ at
    odd index (lambda ~ U[0, 2]); nu ~ U[0.5, 2], D = 0.01, dt = 0.005, h0 = 0."""
    rng = np.random.default_rng(seed)
    specs = []
    for i in range(count):
        n = int(rng.choice([64, 128, 256, 512]))
        nu = float(rng.uniform(0.5, 2.0))
        lam = float(rng.uniform(0.0, 2.0)) if i % 2 == 1 else 0.0
        specs.append(spec_kpz(n, nu=nu, lam=lam, D=0.01, dt=0.005, seed=1000 + i))
    return specs


def test_full_ensemble_members_equal_alone():
    """Configuration C4 at its full size: all 512 instances in one batch
    (hipGraphs of 20 steps, spectral noise), 40 steps in two runs; two EW
    and two KPZ members of every size then equal the same systems advanced
    alone, bit for bit, and every member's state is finite."""
    specs = full_ensemble_specs()
    batch = C.Batch(graph_steps=20)
    members = []
    for sp in specs:
        g, ids = make(sp, 1)
        batch.add(g)
        members.append((g, ids))
    print("C4 ensemble: %d members, 40 steps" % len(members), flush=True)
    batch.step(20)
    batch.step(20)
    picked = []
    for n in (64, 128, 256, 512):
        for kpz in (False, True):
            idx = [i for i, sp in enumerate(specs) if sp["dims"][0] == n and (len(sp["terms"]) > 3) == kpz]
            picked += idx[:2]
    for i in picked:
        g, ids = members[i]
        got = g.get_real(ids["h"])
        a, ia = make(specs[i], 1)
        a.step(40)
        ref = a.get_real(ia["h"])
        a.close()
        print("  member %3d (%d^2, %s): max|h| %.3e" % (i, specs[i]["dims"][0], "KPZ" if len(specs[i]["terms"]) > 3 else "EW",
                                                     float(np.max(np.abs(got)))), flush=True)
        assert np.array_equal(got, ref), i
    for g, ids in members[::37]:
        assert np.all(np.isfinite(g.get_real(ids["h"])))
    batch.close()
    for g, _ in members:
        g.close()


def test_model_h_split_runs_match_one_run():
    """Runs of 10 + 13 steps against one run of 23.  A run continues from
    the stored fields' next phase that the previous fused run left: equal bit
    for bit.  After a change of state in between (here set_real with the
    same values) the next run rederives that phase with the general program:
    equal to rounding."""
    sp = spec_model_h(64, seed=43)
    a, ids = make(sp)
    b, _ = make(sp)
    c, _ = make(sp)
    assert a.fused_eligible
    a.step(23)
    b.step(10)
    b.step(13)
    c.step(10)
    c.set_real(ids["phi"], c.get_real(ids["phi"]))
    c.step(13)
    for name, f in ids.items():
        x, y, z = a.get_real(f), b.get_real(f), c.get_real(f)
        err = rel_l2(z, x)
        print("  %-7s split: max|diff| %.3e; rederived: rel L2 %.3e" % (name, float(np.max(np.abs(y - x))), err))
        assert np.array_equal(x, y), name
        assert err < 1e-5, name
