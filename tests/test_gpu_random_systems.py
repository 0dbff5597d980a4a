"""GPU: seeded random 2D systems through the fused planner, against the
oracle's FP32 scheme (the reference's phase order, src/evolver.cpp:199-226).

Each case draws 1-3 dynamic fields (implicit q^2 and sometimes q^4
operators) and a random mix of the other field kinds the fused step
distinguishes:

* derived fields, linear in one dynamic field;
* a constraint of dynamic-field products (solved at the start of a pass);
* a constraint of derived-field products divided by (1 + q^2), which the
  others read one step late;
* a coefficient field without terms.

The dynamic fields get random product, single-field and constant terms with
small coefficients.  The test steps each system 20 times on its default path
and checks every field against the oracle (relative L2 <= 1e-4, or absolute
<= 1e-6 for fields near zero).  Most cases must take the fused step, so the
planner's classification is what gets tested.  Cases that draw a shape the
fused kernels do not take (more than 3 real fields at sx = 32) run on the
general program and are checked the same way.

A second test advances 24 of the fused cases together in one batch.  Their
size classes mix row-kernel widths and the three column-kernel variants
(no stored fields, stored constraints, stored dynamic or fresh fields).
Each member must be bit-identical to the same system advanced alone.

A third test adds noise to random fields (dynamic and constraint) of 30
cases and compares the fused step with the general program, which draws
from the same Philox counters, in both noise modes (relative L2 <= 1e-5).
"""
import numpy as np
import pytest

import cupss_amd as C
from oracle import oracle as O
from tests.helpers import build_system

pytestmark = pytest.mark.gpu

MULTS = [(1.0,), (1.0, 0, 1), (1.0, 0, 0, 1), (-1.0, 1)]  # 1, iqx, iqy, -q^2


def random_spec(seed):
    rng = np.random.default_rng(seed)
    sx, sy = (int(rng.choice([32, 64, 128])) for _ in range(2))
    n = sx * sy
    nd = int(rng.integers(1, 4))
    dyn = ["d%d" % i for i in range(nd)]
    fields = [(d, True) for d in dyn]
    implicit = {d: [(-float(rng.uniform(0.2, 1.0)), 1)] + ([(-0.05, 2)] if rng.random() < 0.5 else [])
                for d in dyn}
    terms, init = [], {d: rng.uniform(-0.1, 0.1, n).astype(np.float32) for d in dyn}

    def coef():
        return float(rng.uniform(-0.3, 0.3))

    derived = []
    for i in range(int(rng.integers(0, 3))):
        name = "g%d" % i
        fields.append((name, False))
        terms.append((name, [str(rng.choice(dyn))], [MULTS[int(rng.integers(1, 4))]]))
        derived.append(name)
    fresh = []
    if rng.random() < 0.5:
        fields.append(("mu", False))
        terms.append(("mu", [str(rng.choice(dyn)), str(rng.choice(dyn))], [(coef(),)]))
        if rng.random() < 0.5:
            terms.append(("mu", [str(rng.choice(dyn))], [(coef(),)]))
        if rng.random() < 0.5:
            terms.append(("mu", [], [(coef(),)]))
        fresh.append("mu")
    lagged = []
    if derived and rng.random() < 0.5:
        fields.append(("c", False))
        a, b = str(rng.choice(derived)), str(rng.choice(derived))
        terms.append(("c", [a, b], [(coef(),)]))
        if fresh and rng.random() < 0.5:
            terms.append(("c", ["mu"], [(coef(),)]))
        implicit["c"] = [(1.0,), (1.0, 1)]
        lagged.append("c")
    kept = []
    if rng.random() < 0.4:
        fields.append(("k", False))
        init["k"] = rng.uniform(0.5, 1.5, n).astype(np.float32)
        kept.append("k")
    readable = dyn + derived + lagged + kept  # in real space (products)
    for d in dyn:
        for _ in range(int(rng.integers(1, 4))):
            kind = rng.random()
            if kind < 0.5:
                fac = [str(rng.choice(readable)) for _ in range(int(rng.integers(2, 4)))]
                terms.append((d, fac, [(coef(),) if rng.random() < 0.6 else (coef(), 1)]))
            elif kind < 0.85:
                src = str(rng.choice(dyn + derived + fresh + lagged + kept))
                terms.append((d, [src], [(coef(),) + MULTS[int(rng.integers(0, 4))][1:]]))
            else:
                terms.append((d, [], [(coef(),)]))
    return {"dims": (sx, sy, 1), "dt": 0.01, "fields": fields, "implicit": implicit, "terms": terms,
            "init": init}


SEEDS = list(range(100))


def test_random_systems_against_oracle():
    fused = 0
    for seed in SEEDS:
        spec = random_spec(1000 + seed)
        names = [f for f, _ in spec["fields"]]
        print("case seed %d: %s, fields %s, %d terms" % (1000 + seed, spec["dims"][:2], names, len(spec["terms"])),
              flush=True)
        g, gid = build_system(C.System, C.pres, spec)
        fused += int(g.fused_eligible)
        o, oid = build_system(O.System, O.pres, dict(spec, precision="f32"))
        g.step(20)
        o.step(20)
        for f in names:
            got, ref = g.get_real(gid[f]), o.get_real(oid[f])
            assert np.all(np.isfinite(ref)), (seed, f)
            err = float(np.linalg.norm(got - ref))
            scale = float(np.linalg.norm(ref))
            ok = err <= 1e-4 * scale or err / np.sqrt(ref.size) <= 1e-6
            if not ok:
                print("  %s: |diff| %.3e |ref| %.3e (fused %d)" % (f, err, scale, int(g.fused_eligible)), flush=True)
            assert ok, (1000 + seed, f)
        g.close()
    print("fused-eligible cases: %d of %d" % (fused, len(SEEDS)), flush=True)
    assert fused >= len(SEEDS) // 2


def test_random_systems_batched_equal_alone():
    specs = []
    for seed in SEEDS:
        spec = random_spec(1000 + seed)
        g, _ = build_system(C.System, C.pres, spec)
        ok = g.fused_eligible
        g.close()
        if ok:
            specs.append((1000 + seed, spec))
        if len(specs) == 24:
            break
    batch = C.Batch(graph_steps=4)
    members = []
    for seed, sp in specs:
        g, ids = build_system(C.System, C.pres, sp)
        batch.add(g)
        members.append((seed, g, ids, sp))
    print("batch of %d random systems" % len(members), flush=True)
    batch.step(18)
    for seed, g, ids, sp in members:
        a, ia = build_system(C.System, C.pres, sp)
        a.step(18)
        for f in ids:
            assert np.array_equal(a.get_real(ia[f]), g.get_real(ids[f])), (seed, f)
        a.close()


@pytest.mark.parametrize("noise_mode", [0, 1])
def test_random_noisy_systems_fused_equal_general(noise_mode):
    checked = 0
    for seed in SEEDS[:30]:
        spec = random_spec(1000 + seed)
        rng = np.random.default_rng(5000 + seed)
        # noise on dynamic fields and, less often, on constraint fields (a
        # noisy derived field becomes a stored constraint, a noisy
        # coefficient field a pure-noise field)
        spec["noise"] = {f: ((float(rng.uniform(0.001, 0.01)),), 7000 + 10 * seed + i)
                         for i, (f, d) in enumerate(spec["fields"]) if rng.random() < (0.6 if d else 0.3)}
        if not spec["noise"]:
            continue
        out = []
        for path in (0, 1):
            g, ids = build_system(C.System, C.pres, spec)
            g.set_path(path)
            g.set_noise_mode(noise_mode)
            if path == 0 and not g.fused_eligible:
                g.close()
                break
            g.step(20)
            out.append({f: g.get_real(ids[f]) for f in ids})
            g.close()
        if len(out) < 2:
            continue
        checked += 1
        for f in out[0]:
            a, b = out[0][f], out[1][f]
            err, scale = float(np.linalg.norm(a - b)), float(np.linalg.norm(b))
            assert err <= 1e-5 * scale or err / np.sqrt(b.size) <= 1e-7, (1000 + seed, f, err, scale)
    print("noise mode %d: %d noisy fused cases equal to the general program" % (noise_mode, checked), flush=True)
    assert checked >= 10
