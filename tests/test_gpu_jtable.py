"""The device J-table (kernels/ftr_jtable.hpp) against the host pass
(host_jtable, itself bitwise equal to the oracle and to the compiled
reference's tables in tests/test_host.py, tests/test_oracle.py):
  * Jr, r, mm and L bitwise;
  * IGtr and the regime flags identical;
  * the subdivergence error text identical;
over named problems, batched-scan points, special kinematics and planar
ladders up to 19 edges (2^19 subsets)."""

import numpy as np
import pytest

import feyntrop_amd
from feyntrop_amd import workloads as W

pytestmark = pytest.mark.gpu


def _special():
    out = {}
    out["box_massless"] = {"graph": [[[0, 1], 1], [[1, 2], 1], [[2, 3], 1], [[3, 0], 1]], "dimension": 6,
                           "scalarproducts": W._four_point(4, (0, 1, 2, 3), 5.0, -2.0, 0.0),
                           "masses_sqr": [0, 0, 0, 0], "num_eps_terms": 1, "lambda": 0.5}
    out["triangle_lightlike"] = {"graph": [[[0, 1], 1], [[1, 2], 1], [[2, 0], 1]], "dimension": 6,
                                 "scalarproducts": [[0, 0.5, -0.5], [0.5, 0, -0.5], [-0.5, -0.5, 1]],
                                 "masses_sqr": [0, 0, 0], "num_eps_terms": 1, "lambda": 0.5}
    for k in (3, 4, 5, 6):
        out[f"ladder{k}"] = W.ladder_4pt(k)
    return out


CASES = {k: v() for k, v in W.all_problems().items()}
CASES.update({f"scan{i}": p for i, p in enumerate(W.scan_points(24, seed=11))})
CASES.update(_special())


def _prep(eng, prob, mode):
    eng.set_jtable_mode(mode)
    try:
        return eng.prepare(prob), None
    except feyntrop_amd.FeyntropError as e:
        return None, str(e)
    finally:
        eng.set_jtable_mode(0)


@pytest.mark.parametrize("name", list(CASES))
def test_device_jtable_equals_host(engine, name):
    prob = dict(CASES[name])
    prob.setdefault("N", 1000)
    prob.setdefault("seed", 0)
    print(f"[jtable] {name} E={len(prob['graph'])}", flush=True)
    ph, eh = _prep(engine, prob, 1)
    pd, ed = _prep(engine, prob, 2)
    assert eh == ed
    if ph is None:
        return
    for a, b in zip(ph.jtable(), pd.jtable()):
        assert a.dtype == b.dtype
        assert np.array_equal(a.view(np.uint8), b.view(np.uint8))   # bitwise, NaN/inf included
    ih, idv = ph.info, pd.info
    for f in ("IGtr", "W", "L", "euclidean", "pseudo_euclidean", "generic", "gp_property", "lambda_eff", "path"):
        assert getattr(ih, f) == getattr(idv, f), f


def test_device_jtable_subdivergence_message(engine):
    """A graph with a UV subdivergence (a bubble inside in D = 4): the error
    names the same first subgraph either way."""
    prob = {"graph": [[[0, 1], 1], [[0, 1], 1], [[1, 2], 1], [[2, 0], 1]], "dimension": 4,
            "scalarproducts": [[-1, 0.5, 0.5], [0.5, -1, 0.5], [0.5, 0.5, -1]], "masses_sqr": [1, 1, 1, 1],
            "num_eps_terms": 1, "lambda": 0.0, "N": 1000, "seed": 0}
    _, eh = _prep(engine, prob, 1)
    _, ed = _prep(engine, prob, 2)
    assert eh is not None and "subdivergence" in eh
    assert eh == ed


def test_large_graph_estimate_equals_oracle(engine, oracle):
    """16-edge ladder (J-table built on the device, tables read from L2/HBM
    by the sampler): integrate(N, seed) draws the oracle's one-thread samples
    (sample count exact, means within 1e-6 sigma, second moments 1e-9)."""
    prob = W.ladder_4pt(5, N=200_000, seed=3)
    engine.set_jtable_mode(2)
    try:
        p = engine.prepare(prob)
    finally:
        engine.set_jtable_mode(0)
    op = oracle.problem(prob)
    ref, dropped = op.estimate(prob["N"], prob["seed"], 1, 1)
    m = engine.integrate(p, prob["N"], prob["seed"])
    n, s1, s2 = m.arrays()
    assert m.dropped == dropped
    rn, rs1, rs2 = ref[0, 0]
    assert n[0, 0] == rn
    sig = np.sqrt(max(rs2 / rn - (rs1 / rn) ** 2, 0) / rn)
    assert abs(s1[0, 0] - rs1) / rn <= 1e-6 * sig + 1e-15 * abs(rs1 / rn)
    assert abs(s2[0, 0] - rs2) <= 1e-9 * abs(rs2)


def test_large_deformed_graph_estimate_equals_oracle(engine, oracle):
    """13-edge, 10-vertex Minkowski ladder (deformed complex path, row-table
    placement): integrate(N, seed) draws the oracle's one-thread samples
    (sample count and drops exact, real and imaginary means within 1e-6
    sigma; second moments within 1e-7: they are dominated by a few large
    samples, whose unpivoted complex factorisations differ from the pivoted
    CPU ones in the last ~8 bits)."""
    prob = W.LARGE["ladder4_mink"](N=50_000, seed=3)
    p = engine.prepare(prob)
    assert p.info.path == 2
    op = oracle.problem(prob)
    ref, dropped = op.estimate(prob["N"], prob["seed"], 1, 1)
    m = engine.integrate(p, prob["N"], prob["seed"])
    n, s1, s2 = m.arrays()
    assert m.dropped == dropped
    for q in (0, 1):
        rn, rs1, rs2 = ref[0, q]
        assert n[0, q] == rn
        sig = np.sqrt(max(rs2 / rn - (rs1 / rn) ** 2, 0) / rn)
        assert abs(s1[0, q] - rs1) / rn <= 1e-6 * sig + 1e-15 * abs(rs1 / rn), q
        assert abs(s2[0, q] - rs2) <= 1e-7 * abs(rs2), q
