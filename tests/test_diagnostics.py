"""Diagnostics parity with the reference binary (feyntrop.cpp:165-231,
tropical_sampling.hpp error paths): for every named problem, 64 points of
the batched scan and a few special kinematics, the lines our binary writes
before "Started integrating" equal the reference's; they cover:
  * regime lines (Euclidean / Pseudo-Euclidean / Minkowski, generic or
    exceptional);
  * the GP-property line, the exceptional-kinematics warnings and the
    analytic-continuation line.
Where the reference rejects the input (subdivergences, ...), our message
contains its what() text.

Runs on the host: our binary is pointed at a device that does not exist, so
it prepares the problem host-only, prints the diagnostics and stops.  The
reference binary is the one built from its sources under oracle/_ref."""

import json
import os
import subprocess

import pytest

import feyntrop_amd
from feyntrop_amd import workloads as W

REF = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref", "feyntrop_ref")
pytestmark = pytest.mark.skipif(not os.path.exists(REF), reason="reference binary not built (oracle/Makefile ref)")


def _special():
    box = W._four_point(4, (0, 1, 2, 3), 5.0, -2.0, 0.0)
    return {
        "bubble_massless_timelike": {"graph": [[[0, 1], 1], [[0, 1], 1]], "dimension": 4,
                                     "scalarproducts": [[1.0, -1.0], [-1.0, 1.0]], "masses_sqr": [0, 0],
                                     "num_eps_terms": 2, "lambda": 0.5},
        "bubble_below_threshold": {"graph": [[[0, 1], 1], [[0, 1], 1]], "dimension": 4,
                                   "scalarproducts": [[1.0, -1.0], [-1.0, 1.0]], "masses_sqr": [1, 1],
                                   "num_eps_terms": 2, "lambda": 0.5},
        "box_massless": {"graph": [[[0, 1], 1], [[1, 2], 1], [[2, 3], 1], [[3, 0], 1]], "dimension": 6,
                         "scalarproducts": box, "masses_sqr": [0, 0, 0, 0], "num_eps_terms": 1, "lambda": 0.5},
        "triangle_lightlike": {"graph": [[[0, 1], 1], [[1, 2], 1], [[2, 0], 1]], "dimension": 6,
                               "scalarproducts": [[0, 0.5, -0.5], [0.5, 0, -0.5], [-0.5, -0.5, 1]],
                               "masses_sqr": [0, 0, 0], "num_eps_terms": 1, "lambda": 0.5},
    }


def _cases():
    out = {k: v() for k, v in W.all_problems().items()}
    for i, p in enumerate(W.scan_points(64, seed=3)):
        out[f"scan{i}"] = p
    out.update(_special())
    return out


CASES = _cases()


def _pre_start(lines):
    lines = [l for l in lines if not l.startswith("feyntrop_amd: hip")]
    for i, l in enumerate(lines):
        if l.startswith("Started integrating"):
            return lines[:i], True
    return lines, False


@pytest.mark.parametrize("name", list(CASES))
def test_diagnostics_match_reference(name):
    p = dict(CASES[name])
    p["N"], p["seed"] = 1000, 0
    text = json.dumps(p)
    ref = subprocess.run([REF], input=text, capture_output=True, text=True, timeout=120,
                         env={"OMP_NUM_THREADS": "1", "PATH": os.environ.get("PATH", "")})
    env = dict(os.environ)
    env["FEYNTROP_AMD_DEVICE"] = "999"
    ours = subprocess.run([feyntrop_amd.BIN_PATH], input=text, capture_output=True, text=True, timeout=120, env=env)
    r_lines, r_started = _pre_start(ref.stderr.splitlines())
    o_lines, o_started = _pre_start(ours.stderr.splitlines())
    if ref.returncode != 0 and not r_started:
        whats = [l.split("what():", 1)[1].strip() for l in ref.stderr.splitlines() if "what():" in l]
        assert whats, ref.stderr
        assert ours.returncode != 0
        for w in whats:
            assert w in ours.stderr, (w, ours.stderr)
        return
    assert r_started and o_started, (ref.stderr, ours.stderr)
    assert o_lines == r_lines
