"""Two lanes per sample (deformed path): a lane pair evaluates one sample,
the rows of the deformation Jacobian's LDL^T split between its lanes.  No
reference counterpart (the reference evaluates one sample per OpenMP
thread, feynman_integral.hpp:278-362); its results are pinned against the
one-lane kernel and the oracle.

Bars (written in each test):
  per sample   the pair's drop flags equal the one-lane kernel's and its
               values agree to rounding: the same operations in the same
               order per entry (the pivots and w_k cross lanes by DPP
               moves), the compiler free to contract a different product
               of a complex multiply into an FMA in each kernel (a few
               units in the last place; median <= 1e-14, all <= 1e-10)
  estimates    integrate(N, seed) draws the oracle's one-thread samples:
               counts and drops exact, means within 1e-6 sigma of the
               one-lane kernel's and 1e-4 sigma of the oracle's
  geometry     one lane and two lanes per sample give the same counts and
               sums up to summation order (1e-11)
  autotuner    the recorded choice reloads with the problem; the real paths
               refuse two lanes per sample
"""

import numpy as np
import pytest

from feyntrop_amd import FeyntropError
from feyntrop_amd import workloads as W

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

# deformed-path problems: E = 3 to 13 (the 13-edge ladder's one-lane kernel spills)
DEFORMED = {
    "1L_3pt_mink": lambda: W.FIXTURES["1L_3pt_mink"](N=10**5),
    "2L_4pt_ladder": lambda: W.FIXTURES["2L_4pt_ladder"](N=10**5),
    "2L_5pt_box_pent": lambda: W.FIXTURES["2L_5pt_box_pent"](N=10**5),
    "double_box": lambda: W.double_box(N=10**5),
    "triple_ladder": lambda: W.triple_ladder(N=10**5),
    "ladder4_mink": lambda: W.LARGE["ladder4_mink"](N=50_000, seed=3),
}


def _replay(engine, p, X, px):
    n = X.shape[0]
    Xd = torch.from_numpy(np.ascontiguousarray(X.reshape(-1))).cuda()
    pd = torch.from_numpy(np.ascontiguousarray(px.reshape(-1))).cuda()
    out = torch.zeros(n * p.K * 2, dtype=torch.float64, device="cuda")
    drop = torch.zeros(n, dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    engine.replay(p, n, Xd.data_ptr(), pd.data_ptr(), out.data_ptr(), drop.data_ptr())
    return out.cpu().numpy().reshape(n, p.K, 2), drop.cpu().numpy()


def _one(engine, prob):
    p = engine.prepare(prob)
    engine.set_lanes_per_sample(p, 1)
    assert engine.lanes_per_sample(p) == 1
    return p


def _pair(engine, prob):
    p = engine.prepare(prob)
    assert p.info.path == 2
    engine.set_lanes_per_sample(p, 2)
    assert engine.lanes_per_sample(p) == 2
    assert engine.variant(p)[0] == 3
    return p


def _crel(a, b):
    za = a[..., 0] + 1j * a[..., 1]
    zb = b[..., 0] + 1j * b[..., 1]
    return np.abs(za - zb) / np.maximum(np.abs(zb), 1e-300)


@pytest.mark.parametrize("name", list(DEFORMED))
def test_pair_replay_equals_one_lane(engine, oracle, name):
    """20k oracle points (odd count: the last block's last pair is alone):
    the pair kernel's drop flags are the one-lane kernel's and its values
    agree to rounding (median <= 1e-14, every sample <= 1e-10); against
    the strict oracle the pair is as close as the one-lane kernel (median
    and 99th percentile within 1.5x)."""
    print(f"[lanes-replay] {name}", flush=True)
    prob = DEFORMED[name]()
    p1 = _one(engine, prob)
    p2 = _pair(engine, prob)
    op = oracle.problem(prob)
    X, px, _ = op.sample(oracle.seed(17), 20001)
    ref, rdrop, _ = op.replay(X, px)
    a, da = _replay(engine, p1, X, px)
    b, db = _replay(engine, p2, X, px)
    assert np.array_equal(da, db)
    ok = da == 0
    r = _crel(b[ok], a[ok])
    e1, e2 = _crel(a[ok], ref[ok]), _crel(b[ok], ref[ok])
    print(f"   pair vs one lane: median {np.median(r):.2e} max {r.max():.2e} frac>0 {np.mean(r > 0):.3f} | "
          f"vs strict oracle: one lane median {np.median(e1):.2e} p99 {np.quantile(e1, 0.99):.2e}, "
          f"pair median {np.median(e2):.2e} p99 {np.quantile(e2, 0.99):.2e}", flush=True)
    assert np.median(r) <= 1e-14
    assert r.max() <= 1e-10
    for q in (0.5, 0.99):
        assert np.quantile(e2, q) <= 1.5 * np.quantile(e1, q) + 1e-15, q


@pytest.mark.parametrize("name", list(DEFORMED))
def test_pair_estimate_equals_oracle_one_thread(engine, oracle, name):
    """integrate(N, seed) with two lanes per sample: the oracle's one-thread
    samples -- counts and drops exact; real and imaginary means within
    1e-6 sigma of the one-lane kernel's and within 1e-4 sigma of the
    oracle's; second moments within 1e-6.  (Sums over 1e5 samples are led
    by a few large, ill-conditioned samples, on which rounding-level
    differences of the factorisations grow: 2L_4pt ladder, lambda = 1,
    1.3e-6 sigma between pair and oracle; triple ladder 1.4e-5 sigma, where
    single deformed samples of the reference's own build differ from the
    strict oracle by up to 3e-3 relative.)"""
    print(f"[lanes-estimate] {name}", flush=True)
    prob = DEFORMED[name]()
    p = _pair(engine, prob)
    p1 = _one(engine, prob)
    ref, dropped = oracle.problem(prob).estimate(prob["N"], prob["seed"], 1, 1)
    m = engine.integrate(p, prob["N"], prob["seed"])
    m1 = engine.integrate(p1, prob["N"], prob["seed"])
    n, s1, s2 = m.arrays()
    _, s11, _ = m1.arrays()
    assert m.dropped == dropped == m1.dropped
    for k in range(p.K):
        for q in (0, 1):
            rn, rs1, rs2 = ref[k, q]
            assert n[k, q] == rn, (k, q)
            sig = np.sqrt(max(rs2 / rn - (rs1 / rn) ** 2, 0) / rn)
            print(f"   eps^{k} part {q}: pair - oracle {(s1[k, q] - rs1) / rn / sig:+.2e} sigma, "
                  f"one lane - oracle {(s11[k, q] - rs1) / rn / sig:+.2e} sigma", flush=True)
            assert abs(s1[k, q] - s11[k, q]) / rn <= 1e-6 * sig + 1e-15 * abs(rs1 / rn), (k, q)
            assert abs(s1[k, q] - rs1) / rn <= 1e-4 * sig + 1e-15 * abs(rs1 / rn), (k, q)
            assert abs(s2[k, q] - rs2) <= 1e-6 * abs(rs2) + 1e-300, (k, q)


@pytest.mark.parametrize("name,blocks", [("double_box", (128, 256, 512)), ("ladder4_mink", (128, 256))])
def test_pair_block_sizes_and_one_lane_agree(engine, name, blocks):
    """Every feasible block size of the pair kernel, the kernel specialised
    to the kinematic point, and the one-lane kernel: the same samples
    (counts exact, sums within 1e-11 of each other: summation order)."""
    print(f"[lanes-blocks] {name}", flush=True)
    prob = DEFORMED[name]()
    N = 300_001
    p1 = _one(engine, prob)
    base = engine.integrate(p1, N, 5)
    n0, s10, s20 = base.arrays()
    p2 = _pair(engine, prob)
    runs = []
    for blk in blocks:
        engine.set_variant(p2, 3, blk)
        assert engine.variant(p2) == (3, blk) and engine.lanes_per_sample(p2) == 2
        runs.append(engine.integrate(p2, N, 5))
    engine.set_specialized(p2, True)
    runs.append(engine.integrate(p2, N, 5))
    for m in runs:
        n, s1, s2 = m.arrays()
        assert m.dropped == base.dropped
        assert np.array_equal(n, n0)
        assert np.allclose(s1, s10, rtol=1e-11, atol=1e-11 * np.abs(s10).max())
        assert np.allclose(s2, s20, rtol=1e-11, atol=0)


def test_pair_noalloc_equals_allocating_call(engine):
    """ftr_integrate_noalloc runs the pair kernel too: same moments as
    ftr_integrate (the block partials summed in the same order)."""
    print("[lanes-noalloc] double_box", flush=True)
    p = _pair(engine, DEFORMED["double_box"]())
    for N in (1_000_003, 4099):
        a = engine.integrate(p, N, 2)
        b = engine.integrate_noalloc(p, N, 2)
        assert a.dropped == b.dropped
        for x, y in zip(a.arrays(), b.arrays()):
            assert np.array_equal(x, y)


def test_lanes_per_sample_api(engine):
    """The real paths refuse two lanes per sample (FTR_ERR_LIMIT); lps must
    be 1 or 2; returning to one lane restores the default variant."""
    print("[lanes-api]", flush=True)
    p = engine.prepare(W.wheel_k4(N=1000))
    with pytest.raises(FeyntropError) as e:
        engine.set_lanes_per_sample(p, 2)
    assert e.value.code == -5
    q = engine.prepare(W.double_box(N=1000))
    assert engine.lanes_per_sample(q) == 1      # E = 7: one lane per sample by default
    default = engine.variant(q)
    with pytest.raises(FeyntropError) as e:
        engine.set_lanes_per_sample(q, 3)
    assert e.value.code == -1
    engine.set_lanes_per_sample(q, 2)
    with pytest.raises(FeyntropError):
        engine.set_variant(q, 0, 256)   # the pair kernel reads rows from HBM only
    engine.set_lanes_per_sample(q, 1)
    assert engine.lanes_per_sample(q) == 1 and engine.variant(q) == default
    # a deformed graph whose one-lane kernel spills (13 edges) defaults to a lane pair
    r = engine.prepare(DEFORMED["ladder4_mink"]())
    assert engine.lanes_per_sample(r) == 2 and engine.variant(r)[0] == 3


def test_autotune_chooses_lanes_per_sample(engine, tmp_path):
    """The autotuner times one- and two-lane variants of a deformed graph,
    keeps the fastest and records it: a fresh preparation of the same graph
    (in a cache directory holding the record: links to the built code
    objects, so nothing is compiled here) comes back with the same table
    placement, block size and lanes per sample, and its estimate equals the
    one-lane kernel's sample for sample."""
    print("[lanes-autotune] ladder4_mink", flush=True)
    import os
    import feyntrop_amd
    kc = os.path.join(os.path.dirname(os.path.abspath(feyntrop_amd.__file__)), "kcache")
    for f in os.listdir(kc):
        if f.endswith(".co"):
            os.symlink(os.path.join(kc, f), tmp_path / f)
    prob = DEFORMED["ladder4_mink"]()
    p = _one(engine, prob)
    ref = engine.integrate(p, 200_000, 1)
    try:
        feyntrop_amd.lib().ftr_set_cache_dir(str(tmp_path).encode())
        q = engine.prepare(prob)
        tm, blk, ms = engine.autotune(q, n_probe=400_000)
        engine.set_specialized(q, False)
        lps = engine.lanes_per_sample(q)
        print(f"   chose table mode {tm}, block {blk}, {lps} lane(s) per sample, {ms:.3f} ms", flush=True)
        r = engine.prepare(prob)
        assert engine.variant(r) == (tm, blk) and engine.lanes_per_sample(r) == lps
        m = engine.integrate(r, 200_000, 1)
    finally:
        feyntrop_amd.lib().ftr_set_cache_dir(None)
    assert m.dropped == ref.dropped
    n, s1, s2 = m.arrays()
    n0, s10, s20 = ref.arrays()
    assert np.array_equal(n, n0)
    assert np.allclose(s1, s10, rtol=1e-11, atol=1e-11 * np.abs(s10).max())


def test_cli_runs_pair_kernel_on_spilling_graph(oracle):
    """The drop-in binary on the 13-edge deformed ladder: its default
    variant is a lane pair per sample (no tuning record needed), and its
    printed estimate is the oracle's one-thread estimate of the same
    samples: values within 1e-6 of their standard error, errors within
    1e-6 relative (the second moments agree to 1e-7; the binary prints
    full doubles)."""
    print("[lanes-cli] ladder4_mink", flush=True)
    import json
    import subprocess
    import feyntrop_amd
    prob = DEFORMED["ladder4_mink"]()
    out = subprocess.run([feyntrop_amd.BIN_PATH], input=json.dumps(prob), capture_output=True, text=True,
                         timeout=300)
    assert out.returncode == 0, out.stderr
    d = json.loads(out.stdout)
    op = oracle.problem(prob)
    ref, dropped = op.estimate(prob["N"], prob["seed"], 1, 1)
    res = op.finalize(ref)
    for q in (0, 1):
        v, dv = d["integral"][0][q]
        assert abs(v - res[0, q, 0]) <= 1e-9 * abs(res[0, q, 0]) + 1e-6 * res[0, q, 1], (q, v, res[0, q])
        assert abs(dv - res[0, q, 1]) <= 1e-6 * res[0, q, 1], (q, dv, res[0, q])

