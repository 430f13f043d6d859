"""Host side of the product, on the CPU (no GPU needed):

  - the C-ABI library loads and exports every function include/feyntrop_amd.h
    declares;
  - host-only preparation (ftr_prepare without a device context) reproduces
    the oracle's J-table bitwise and the reference's regime flags;
  - the feyntrop CLI validates its input like the reference and exits nonzero
    with the reference's messages (the device is not needed for that);
  - kernel specialisation (hipRTC for gfx950) compiles without a GPU.
"""

import json
import os
import re
import subprocess

import numpy as np
import pytest

import feyntrop_amd
from feyntrop_amd import workloads as W

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "feyntrop_amd.h")


def test_library_exports_every_declared_symbol():
    lib = feyntrop_amd.lib()
    text = open(HEADER).read()
    names = sorted(set(re.findall(r"\b(ftr_\w+)\s*\(", text)))
    assert len(names) >= 15
    for n in names:
        assert hasattr(lib, n), n


def _host_prepare(prob):
    import ctypes as C
    inp = feyntrop_amd._Input(prob)
    out = C.c_void_p()
    rc = feyntrop_amd.lib().ftr_prepare(None, C.byref(inp.s), C.byref(out))
    if rc:
        raise feyntrop_amd.FeyntropError(rc, feyntrop_amd.lib().ftr_last_error(None).decode())
    return feyntrop_amd.Problem(None, out, prob)


@pytest.mark.parametrize("name", list(W.all_problems()))
def test_jtable_is_its_row_sums(oracle, name):
    """The J recursion of tropical_sampling.hpp:258-276 and :411, bit for
    bit: J(g) is the running sum, from 0.0 in ascending edge order, of
    J(g\e)/r(g\e) over the edges of g, and so equals the last running sum
    of the sampler's row for g (the absent edges' zeros leave the sum
    unchanged).  Checked on the oracle's tables, which equal the
    reference's."""
    Jr, r, _, _ = oracle.problem(W.all_problems()[name]()).jtable()
    E = int(round(np.log2(len(Jr))))
    # (r of the full graph may be 0: it is never the state after a cut)
    w = [float(Jr[g]) / float(r[g]) if r[g] != 0 else float("nan") for g in range(len(Jr))]
    for g in range(1, len(Jr)):
        T = 0.0
        for e in range(E):
            T += w[g ^ (1 << e)] if (g >> e) & 1 else 0.0
        assert T == float(Jr[g]), (g, T, Jr[g])


@pytest.mark.parametrize("name", list(W.all_problems()))
def test_host_jtable_matches_oracle(oracle, name):
    prob = W.all_problems()[name]()
    hp = _host_prepare(prob)
    op = oracle.problem(prob)
    for a, b in zip(hp.jtable(), op.jtable()):
        assert np.array_equal(a, b)
    i = hp.info
    assert (i.E, i.V, i.L) == (op.E, op.V, op.L_loops)
    assert i.W == op.W and i.IGtr == op.IGtr
    flags = [i.euclidean, i.pseudo_euclidean, i.generic, i.gp_property]
    assert flags == [int(x) for x in op.info[5:9]]
    assert i.lambda_eff == op.info[9]
    assert (i.path == 2) == bool(op.info[11])


def _cli(inp):
    text = inp if isinstance(inp, str) else json.dumps(inp)
    env = dict(os.environ)
    env["HIP_VISIBLE_DEVICES"] = env.get("HIP_VISIBLE_DEVICES", "")
    return subprocess.run([feyntrop_amd.BIN_PATH], input=text, capture_output=True, text=True, timeout=120)


BAD = [
    ("missing field", lambda p: p.pop("seed"), "input json must contain the data fields"),
    ("N not a number", lambda p: p.update(N="many"), "input N must be a number"),
    ("tadpole", lambda p: p["graph"].__setitem__(0, [[1, 1], 1]), "tadpole"),
    ("weight", lambda p: p["graph"].__setitem__(0, [[0, 1], 0]), "edge weights must be larger than 0"),
    ("negative vertex", lambda p: p["graph"].__setitem__(0, [[-1, 1], 1]), "vertex numbers must start from 0"),
    ("graph form", lambda p: p["graph"].__setitem__(0, [[0, 1, 2], 1]), "input graph must be a list"),
    ("P shape", lambda p: p.update(scalarproducts=[[0, 0], [0, 0]]), "scalarproducts must be a (V x V)-matrix"),
    ("masses", lambda p: p.update(masses_sqr=[1, 2]), "masses_sqr must be an array of size E"),
    ("asymmetric", lambda p: p["scalarproducts"][0].__setitem__(1, 5), "not symmetric"),
    ("row sums", lambda p: (p["scalarproducts"][0].__setitem__(0, -2), p["scalarproducts"][1].__setitem__(1, -2)),
     "momentum conservation"),
]


@pytest.mark.parametrize("label,mut,msg", BAD, ids=[b[0] for b in BAD])
def test_cli_rejects_bad_input(label, mut, msg):
    p = W.wheel_k4(N=1000)
    mut(p)
    r = _cli(p)
    assert r.returncode != 0
    assert r.stdout == ""
    assert msg in r.stderr


def test_cli_rejects_unknown_proposal():
    p = W.wheel_k4(N=1000)
    p["proposal"] = "magic"
    r = _cli(p)
    assert r.returncode != 0 and r.stdout == ""
    assert 'input proposal must be "plain" or "learned"' in r.stderr


def test_cli_batch_rejects_bad_item():
    """Batch mode names the failing item and keeps the item's own message."""
    good = W.wheel_k4(N=1000)
    bad = W.wheel_k4(N=1000)
    bad["N"] = "many"
    r = _cli({"batch": [good, bad]})
    assert r.returncode != 0 and r.stdout == ""
    assert "batch item 1: input N must be a number" in r.stderr
    r = _cli({"batch": []})
    assert r.returncode != 0 and "non-empty list" in r.stderr


def test_cli_rejects_malformed_json():
    r = _cli("{\"N\": 10, ")
    assert r.returncode != 0 and r.stdout == ""


def test_cli_disconnected_and_subdivergent():
    p = W.wheel_k4(N=1000)
    p["graph"] = [[[0, 1], 1], [[0, 1], 1], [[2, 3], 1], [[2, 3], 1]]
    p["masses_sqr"] = [1, 1, 1, 1]
    r = _cli(p)
    assert r.returncode != 0 and "is not connected" in r.stderr
    # a one-loop bubble inside a two-loop graph at D = 4 is log divergent
    p = W.wheel_k4(N=1000)
    p["graph"] = [[[0, 1], 1], [[0, 1], 1], [[1, 2], 1], [[2, 0], 1]]
    p["scalarproducts"] = [[-1, 1, 0], [1, -1, 0], [0, 0, 0]]
    p["masses_sqr"] = [1, 1, 1, 1]
    r = _cli(p)
    assert r.returncode != 0 and "subdivergence" in r.stderr


def test_cli_singleton_vertex():
    p = W.wheel_k4(N=1000)
    p["graph"] = [[[0, 2], 1], [[0, 2], 1], [[2, 0], 1]]
    r = _cli(p)
    assert r.returncode != 0 and "singleton vertex 1" in r.stderr


def test_precompile_variants_without_gpu(tmp_path):
    """A chosen kernel variant compiles on a host without a GPU: the lane-pair
    kernel of a deformed graph (generic and specialised) and the folded-row
    placement (table mode 4) of a real and a deformed graph write code
    objects; a variant that cannot exist is FTR_ERR_LIMIT (-5): two lanes per
    sample on the real path, a lane pair with the table in LDS, or a table
    mode that does not exist."""
    lib = feyntrop_amd.lib()
    lib.ftr_set_cache_dir(str(tmp_path).encode())
    try:
        prob = W.fixture_1L_3pt_mink()
        prob["masses_sqr"] = [0.3, 0.7, 1.9]
        feyntrop_amd.precompile_variant(prob, 3, 256, lps=2)
        feyntrop_amd.precompile_variant(prob, 3, 128, lps=2, specialized=True)
        wheel = W.wheel_k4()
        wheel["masses_sqr"] = [0.1, 1, 2, 3, 4, 5]
        feyntrop_amd.precompile_variant(wheel, 4, 256)
        feyntrop_amd.precompile_variant(prob, 4, 256, lps=2)
        cos = [f for f in os.listdir(tmp_path) if f.endswith(".co")]
        assert len(cos) == 4
        for tm, blk, lps, q in ((3, 256, 2, W.wheel_k4()), (0, 256, 2, prob), (3, 1024, 2, prob),
                                (5, 256, 1, prob)):
            with pytest.raises(feyntrop_amd.FeyntropError) as e:
                feyntrop_amd.precompile_variant(q, tm, blk, lps=lps)
            assert e.value.code == -5
    finally:
        lib.ftr_set_cache_dir(None)


def test_precompile_without_gpu(tmp_path):
    """hipRTC specialises a new graph for gfx950 on a host without a GPU."""
    lib = feyntrop_amd.lib()
    lib.ftr_set_cache_dir(str(tmp_path).encode())
    try:
        prob = W.fixture_1L_3pt_mink()
        prob["masses_sqr"] = [0.3, 0.7, 1.1]
        feyntrop_amd.precompile(prob)
        cos = [f for f in os.listdir(tmp_path) if f.endswith(".co")]
        assert cos and all(os.path.getsize(tmp_path / f) > 1000 for f in cos)
        with open(tmp_path / cos[0], "rb") as f:
            assert f.read(4) == b"\x7fELF"
    finally:
        lib.ftr_set_cache_dir(None)
