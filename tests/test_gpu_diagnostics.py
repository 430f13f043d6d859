"""Diagnostics parity on the device path (reference: feyntrop.cpp:165-231 and
the error paths of tropical_sampling.hpp:284-458).  On a GPU host the binary
builds the J-tables of graphs with 13 or more edges on the device, whose
pass also sets the regime flags (generic, pseudo-Euclidean, GP property)
that these lines report.  For every case of tests/diag_cases.py the lines
the binary writes before "Started integrating" must equal the reference
binary's (tests/golden/cli.json), or, where the reference rejects the
graph, our message must contain its what() text.  Run with the J-table
placement chosen automatically and forced onto the device."""

import json
import os
import subprocess

import pytest

from diag_cases import large_cases

pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

CASES = large_cases()
HERE = os.path.dirname(os.path.abspath(__file__))


def _gold():
    with open(os.path.join(HERE, "golden", "cli.json")) as f:
        return json.load(f)["diagnostics_large"]


@pytest.mark.parametrize("jmode", ["auto", "device"])
@pytest.mark.parametrize("name", list(CASES))
def test_device_diagnostics_match_reference(name, jmode):
    import feyntrop_amd
    print(f"[diag-device] {name} jtable {jmode}", flush=True)
    prob = dict(CASES[name])
    assert len(prob["graph"]) >= 13
    prob["N"], prob["seed"] = 1000, 0
    env = dict(os.environ)
    env.pop("FEYNTROP_AMD_DEVICE", None)
    if jmode == "device":
        env["FEYNTROP_AMD_JTABLE"] = "device"
    else:
        env.pop("FEYNTROP_AMD_JTABLE", None)
    out = subprocess.run([feyntrop_amd.BIN_PATH], input=json.dumps(prob), capture_output=True, text=True,
                         timeout=300, env=env)
    gold = _gold()[name]
    if "error" in gold:
        assert out.returncode != 0
        for w in gold["error"]:
            assert w in out.stderr, (w, out.stderr)
        return
    assert out.returncode == 0, out.stderr
    lines = [l for l in out.stderr.splitlines() if not l.startswith("feyntrop_amd: hip")]
    start = [i for i, l in enumerate(lines) if l.startswith("Started integrating")]
    assert start, out.stderr
    assert lines[:start[0]] == gold["lines"]
    json.loads(out.stdout)
