"""CPU: the product parser (cupss_amd/csrc/host/parser.cpp) against parses
derived by hand from the reference grammar (src/parser.cpp:98-846).

The harness tests/cpp/parser_harness.cpp links parser.cpp with recording
stand-ins of evolver::createField/createTerm and prints the parsed system."""
import json
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "bin", "parser_harness")


def harness():
    srcs = [os.path.join(ROOT, "tests", "cpp", "parser_harness.cpp"),
            os.path.join(ROOT, "cupss_amd", "csrc", "host", "parser.cpp")]
    if not os.path.exists(BIN) or any(os.path.getmtime(s) > os.path.getmtime(BIN) for s in srcs):
        os.makedirs(os.path.dirname(BIN), exist_ok=True)
        subprocess.run(["g++", "-std=c++17", "-O1", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
                        "-I" + os.path.join(ROOT, "include"), "-o", BIN] + srcs, check=True)
    return BIN


def parse(lines, tmp_path, expect_fail=False):
    f = tmp_path / "sys.txt"
    f.write_text("\n".join(lines) + "\n")
    r = subprocess.run([harness(), str(f)], capture_output=True, text=True, timeout=60)
    if expect_fail:
        assert r.returncode == 1, r.stdout
        return r.stdout
    assert r.returncode == 0, r.stdout + r.stderr
    return {d["field"]: d for d in map(json.loads, r.stdout.strip().splitlines())}


def P(pre, q2n=0, iqx=0, iqy=0, iqz=0, invq=0):
    return [pytest.approx(pre, rel=1e-6, abs=1e-12), q2n, iqx, iqy, iqz, invq]


def test_reference_test_equations(tmp_path):
    # tests/tests.cpp:293-306
    s = parse(["field phi 1", "field lapphi 0", "field iqxphi 0", "field iqzphi 0", "field invqphi 0",
               "eq dt phi +q^2*phi = iqxphi^2", "eq lapphi = -q^2*phi", "eq iqxphi =  iqx*phi",
               "eq iqzphi =  iqz*phi", "eq invqphi =  1/q*phi"], tmp_path)
    assert s["phi"]["dynamic"] == 1
    assert s["phi"]["implicit"] == [P(-1, 1)]
    assert s["phi"]["terms"] == [{"product": ["iqxphi", "iqxphi"], "pres": [P(1)], "strings": ["iqxphi^2"],
                                  "used": []}]
    assert s["lapphi"]["implicit"] == []            # bare LHS stores no implicit
    assert s["lapphi"]["terms"][0]["pres"] == [P(-1, 1)]
    assert s["iqxphi"]["terms"][0]["pres"] == [P(1, 0, 1)]
    assert s["iqzphi"]["terms"][0]["pres"] == [P(1, 0, 0, 0, 1)]
    assert s["invqphi"]["terms"][0]["pres"] == [P(1, 0, 0, 0, 0, 1)]


def test_cahn_hilliard_and_noise(tmp_path):
    # examples/02_cahn_hilliard/02_cahn-hilliard.cpp:28-35
    s = parse(["field phi 1", "param a -1.0", "param b 1.0", "param k 4.0", "param D 1.0",
               "eq dt phi + q^2*(a + k*q^2)*phi= - b*q^2*phi^3", "noise phi 2*D*q^2"], tmp_path)
    phi = s["phi"]
    # q^2*(a+k*q^2)*phi -> q^2*a*phi, q^2*k*q^2*phi; negated on a dynamic LHS
    assert phi["implicit"] == [P(1.0, 1), P(-4.0, 2)]
    assert phi["terms"] == [{"product": ["phi"] * 3, "pres": [P(-1.0, 1)], "strings": ["-b*q^2*phi^3"],
                             "used": ["b"]}]
    assert phi["used"] == ["a", "k"]
    assert phi["noise"] == P(2.0, 1)


def test_diffusion_constant_term_and_kpz(tmp_path):
    s = parse(["field phi 1", "param D 1.0", "eq dt phi + D * q^2 * phi = 0"], tmp_path)
    assert s["phi"]["implicit"] == [P(-1.0, 1)]
    # the "0" right-hand side is an empty-product (constant) term of prefactor 0
    assert s["phi"]["terms"] == [{"product": [], "pres": [P(0.0)], "strings": ["0"], "used": []}]
    s = parse(["field h 1", "field iqxh 0", "field iqyh 0", "param nu 0.7", "param lambda 1.5",
               "eq dt h + nu*q^2*h = 0.5*lambda*iqxh^2 + 0.5*lambda*iqyh^2", "eq iqxh = iqx*h",
               "eq iqyh = iqy*h", "noise h 2*0.01"], tmp_path)
    assert s["h"]["implicit"] == [P(-0.7, 1)]
    assert [t["product"] for t in s["h"]["terms"]] == [["iqxh", "iqxh"], ["iqyh", "iqyh"]]
    assert [t["pres"] for t in s["h"]["terms"]] == [[P(0.75)], [P(0.75)]]
    assert s["h"]["noise"] == P(0.02)


def test_model_h_explicit_velocity(tmp_path):
    # examples/04_model_h/modelh_explicit_velocity.cpp:14-39
    s = parse(["field phi 1", "field iqxphi 0", "field iqyphi 0", "field sigxx 0", "field sigxy 0",
               "field vx 0", "field vy 0", "field w 0", "param a -1", "param b 1", "param k 4", "param eta 1",
               "param friction 0", "param ka -4",
               "eq dt phi + ( a *q^2 + k*q^4)*phi= - b* q^2* phi^3 -vx*iqxphi - vy*iqyphi",
               "eq sigxx = - 0.5*ka *iqxphi * iqxphi + 0.5*ka*iqyphi*iqyphi",
               "eq vx * (friction + eta*q^2) = (iqx + iqx^3*1/q^2 - iqx*iqy^2*1/q^2) * sigxx + "
               "(iqy + iqx^2* iqy*1/q^2 + iqx^2*iqy*1/q^2) * sigxy",
               "eq vy * (friction + eta*q^2) = (iqx + iqx*iqy^2*1/q^2 + iqx*iqy^2*1/q^2) * sigxy + "
               "(-iqy - iqy^3*1/q^2 + iqx^2*iqy*1/q^2) * sigxx"], tmp_path)
    assert s["phi"]["implicit"] == [P(1, 1), P(-4, 2)]
    assert [t["product"] for t in s["phi"]["terms"]] == [["phi"] * 3, ["iqxphi", "vx"], ["iqyphi", "vy"]]
    assert [t["pres"] for t in s["sigxx"]["terms"]] == [[P(2.0)], [P(-2.0)]]
    assert s["vx"]["implicit"] == [P(0.0), P(1.0, 1)]       # constraint LHS: not negated
    assert s["vx"]["used"] == ["eta", "friction"]
    assert s["vx"]["terms"][0]["pres"] == [P(1, 0, 1), P(1, 0, 3, 0, 0, 2), P(-1, 0, 1, 2, 0, 2)]
    assert s["vx"]["terms"][1]["pres"] == [P(1, 0, 0, 1), P(1, 0, 2, 1, 0, 2), P(1, 0, 2, 1, 0, 2)]
    assert s["vy"]["terms"][1]["pres"] == [P(-1, 0, 0, 1), P(-1, 0, 0, 3, 0, 2), P(1, 0, 2, 1, 0, 2)]


def test_grammar_details(tmp_path):
    s = parse(["field u 1", "field v 0", "field p 0", "param c 4",
               # q^3 -> q2n += 3/2 = 1 (integer division); a/b -> a*1/b
               "eq dt u = q^3*u + c/2*v + 3/c*u*v - (2*v - (q^2*u))",
               "eq 2*v = u",                      # non-bare LHS: implicit [2]
               "eq -q^2*p = 1/q^2*u"], tmp_path)
    u = s["u"]
    assert u["implicit"] == []
    groups = {tuple(t["product"]): t["pres"] for t in u["terms"]}
    assert groups[("u",)] == [P(1, 1), P(1, 1)]            # q^3*u and -(-(q^2*u)) -> +q^2*u
    assert groups[("v",)] == [P(2.0), P(-2.0)]             # c/2*v ; -(2*v)
    assert groups[("u", "v")] == [P(0.75)]
    assert s["v"]["implicit"] == [P(2.0)]
    assert s["p"]["implicit"] == [P(-1.0, 1)]
    assert s["p"]["terms"][0]["pres"] == [P(1, 0, 0, 0, 0, 2)]


def test_empty_rhs_is_a_unit_constant(tmp_path):
    # "dt u = " leaves one empty term whose prefactor is 1 (src/parser.cpp:547-589)
    s = parse(["field u 1", "eq dt u = "], tmp_path)
    assert s["u"]["terms"] == [{"product": [], "pres": [P(1.0)], "strings": [""], "used": []}]


@pytest.mark.parametrize("eq", ["dt u = u = u", "dt u = zz*u", "dt u = 2/(1+q^2)*u", "u*u = u",
                                "dt u + u^2 = 0", "dt u = q^x*u"])
def test_errors_exit_like_the_reference(eq, tmp_path):
    parse(["field u 1", "eq " + eq], tmp_path, expect_fail=True)
