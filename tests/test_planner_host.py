"""CPU test: the batch planner and plan packing under AddressSanitizer.

tests/planner_check.cpp compiles the engine's host code (host pass only, every
kernel launcher stubbed) with ASan, plans batches from stdin into an image of
the plan arena exactly as large as its layout, and checks every index the
kernels form from the plan (rows, T/M entries, replicas, parts, descriptors,
tiles, group members) against the arena capacities.  Cases: every single-term
mask at x = 1e11 (the V-only plan among them), the 512-query workload, slices
of the served stream, and single queries to 2^64 + 1 and 1e20.  The bytes the
admission estimate charges for one query must cover what its plan uses."""
import json
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")
HIPCC = "/opt/rocm/bin/hipcc"

# term masks (engine.hpp)
ORD, TRIV, EASY, HARD, P2, ALL, EASY_A, EASY_B = 1, 2, 4, 8, 16, 31, 32, 64


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    if not os.path.exists(HIPCC) or not shutil.which("g++"):
        pytest.skip("no hipcc")
    exe = tmp_path_factory.mktemp("planner") / "planner_check"
    cmd = [HIPCC, "-x", "hip", "--offload-host-only", "-std=c++17", "-O1", "-g",
           "-Xarch_host", "-fsanitize=address", "-fno-omit-frame-pointer",
           os.path.join(ROOT, "tests", "planner_check.cpp"), "-o", str(exe), "-lpthread"]
    subprocess.run(cmd, check=True, capture_output=True)
    return exe


def run(exe, lines):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=0")
    p = subprocess.run([str(exe)], input="\n".join(lines) + "\n", capture_output=True, text=True,
                       env=env, timeout=600)
    assert p.returncode == 0, p.stdout[-2000:] + p.stderr[-4000:]
    out = [ln.split() for ln in p.stdout.splitlines()]
    assert len(out) == len(lines) and all(o[0] == "ok" for o in out), p.stdout
    return [dict(zip(("queries", "tiles", "plan", "out", "estimate", "passes"), map(int, o[1:]))) for o in out]


def test_single_term_plans(checker, oracle):
    x = 10 ** 11
    y = oracle.choose_y(x, 2.0)
    masks = [EASY, ORD, TRIV, HARD, P2, ALL, EASY | EASY_A, EASY | EASY_B]
    run(checker, ["%d 0 %d:%d" % (m, x, y) for m in masks])


def test_batch_plans(checker):
    with open(os.path.join(GOLD, "queries_c3.json")) as f:
        c3 = json.load(f)["x"]
    with open(os.path.join(GOLD, "queries_c4.json")) as f:
        c4 = json.load(f)["streams"]["2000"]["x"]
    lines = ["%d 0 %s" % (ALL, " ".join(map(str, c3))),
             "%d 2 %s" % (ALL, " ".join(map(str, c3[:64]))),
             "%d 1 %s" % (ALL, " ".join(map(str, c4[:256]))),
             "%d 1 %s" % (ALL, " ".join(map(str, c4[256:300]))),
             "%d 0 5 16383 16384 100000 1000000000" % ALL]
    r = run(checker, lines)
    assert r[0]["queries"] == 512 and r[0]["tiles"] > 512


def test_single_query_plans_and_estimate(checker):
    xs = [10 ** 9, 10 ** 12, 10 ** 14, 10 ** 16, 10 ** 17, 2 ** 64 - 1, 2 ** 64 + 1, 10 ** 20]
    r = run(checker, ["%d 0 %d" % (ALL, x) for x in xs])
    for x, v in zip(xs, r):
        if x <= 10 ** 17:
            assert v["estimate"] >= v["plan"] + v["out"], (x, v)


def test_pass_plans(checker):
    # a single query's hard segments in passes over segment ranges: forced
    # small pass sizes (down to one segment per pass) and 1e20 at 1 GB; the
    # checker verifies that the ranges cover the segments, every tile sits in
    # its range's list once, and every T and M row a pass touches lies in the
    # arrays sized for the largest range
    cases = [(10 ** 10, 4), (10 ** 12, 2048), (10 ** 14, 65536), (10 ** 16, 1 << 20),
             (2 ** 64 + 1, 1 << 28), (10 ** 20, 10 ** 9)]
    r = run(checker, ["pass=%d %d 0 %d" % (b, ALL, x) for x, b in cases])
    assert all(v["passes"] > 1 for v in r), r
    # the default pass size leaves these in one pass
    r = run(checker, ["%d 0 %d" % (ALL, x) for x, _ in cases[:4]])
    assert all(v["passes"] == 1 for v in r), r
