"""Diagnostics of the graphs with 13 or more edges (tests/diag_cases.py) on
the host path: the binary, pointed at a device that does not exist, builds
the J-table on the host and must print the reference binary's lines
(tests/golden/cli.json; feyntrop.cpp:165-231) or contain its error text.
The same cases run on the device path in test_gpu_diagnostics.py."""

import json
import os
import subprocess

import pytest

import feyntrop_amd
from diag_cases import large_cases

CASES = large_cases()
HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.parametrize("name", list(CASES))
def test_host_diagnostics_match_reference(name):
    with open(os.path.join(HERE, "golden", "cli.json")) as f:
        gold = json.load(f)["diagnostics_large"][name]
    prob = dict(CASES[name])
    prob["N"], prob["seed"] = 1000, 0
    env = dict(os.environ)
    env["FEYNTROP_AMD_DEVICE"] = "999"
    out = subprocess.run([feyntrop_amd.BIN_PATH], input=json.dumps(prob), capture_output=True, text=True,
                         timeout=300, env=env)
    if "error" in gold:
        assert out.returncode != 0
        for w in gold["error"]:
            assert w in out.stderr, (w, out.stderr)
        return
    lines = [l for l in out.stderr.splitlines() if not l.startswith("feyntrop_amd: hip")]
    start = [i for i, l in enumerate(lines) if l.startswith("Started integrating")]
    assert start, out.stderr
    assert lines[:start[0]] == gold["lines"]
