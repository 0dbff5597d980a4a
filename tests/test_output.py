"""The background field-output writer (cupss_amd/csrc/host/output.cpp)
writes byte-identical files to a synchronous write, in the format of the
reference's field::writeToFile (src/field.cpp:350-402): header
"x, [y, [z, ]]name", then one "%i, ... %.<precision>f" line per point with x
fastest.  Host code only: runs without a device."""
import math
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def build(tmp_path):
    exe = str(tmp_path / "output_harness")
    subprocess.run(["g++", "-O2", "-std=c++17", "-pthread", os.path.join(ROOT, "tests", "cpp", "output_harness.cpp"),
                    os.path.join(ROOT, "cupss_amd", "csrc", "host", "output.cpp"), "-o", exe], check=True)
    return exe


def run(exe, out, sync):
    os.makedirs(out, exist_ok=True)
    env = dict(os.environ)
    env["CUPSS_SYNC_OUTPUT"] = "1" if sync else "0"
    r = subprocess.run([exe, str(out)], env=env, capture_output=True, text=True, check=True)
    assert r.stdout.strip() == "36"


def test_background_writer_matches_synchronous_and_reference_format(tmp_path):
    exe = build(tmp_path)
    run(exe, tmp_path / "async", False)
    run(exe, tmp_path / "sync", True)
    names = sorted(os.listdir(tmp_path / "sync"))
    assert len(names) == 36 and names == sorted(os.listdir(tmp_path / "async"))
    for n in names:
        a = (tmp_path / "async" / n).read_bytes()
        b = (tmp_path / "sync" / n).read_bytes()
        assert a == b, n
    # format: 2D field f2, step 3 (precision 6)
    lines = (tmp_path / "async" / "f2.csv.3").read_text().splitlines()
    assert lines[0] == "x, y, f2"
    assert len(lines) == 1 + 8 * 4
    i, j = 5, 2
    v = np.float32(math.sin(0.37 * (j * 8 + i) + 3)) * np.float32(3.25)
    assert lines[1 + j * 8 + i] == "%i, %i, %.6f" % (i, j, v)
    lines3 = (tmp_path / "async" / "f3.csv.4").read_text().splitlines()
    assert lines3[0] == "x, y, z, f3" and lines3[1].startswith("0, 0, 0, ")
    assert lines3[1].count(".") == 1 and len(lines3[1].split(".")[1]) == 3
