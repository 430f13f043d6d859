"""Folded rows (table mode 4, the default placement from 16 edges) on a
graph whose tables outgrow the caches: the 6-rung planar ladder (E = 19,
2^19 states).

Bars: against mode 3 (the same cut comparisons and radial arithmetic in
another kernel) the cut order is identical, X and psi/xi within 1e-15
relative, and an integration's counts are exact and its sums within 1e-12
relative; against the oracle (the reference's sampler restated on the CPU)
the cut order is identical and X within 1e-12 relative (the last-bit
rounding of the radial pow), at a stream offset reached by jump-ahead."""

import numpy as np
import pytest

from feyntrop_amd import workloads as W

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu


def _sample(engine, p, first, n, seed):
    X = torch.zeros(n * p.E, dtype=torch.float64, device="cuda")
    px = torch.zeros(2 * n, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    engine.sample(p, first, n, seed, X.data_ptr(), px.data_ptr())
    return X.cpu().numpy().reshape(n, p.E), px.cpu().numpy().reshape(n, 2)


def test_folded_rows_equal_row_tables_and_oracle(engine, oracle):
    prob = W.ladder_4pt(6, 200_000, 0)
    print("[rows] ladder6 E=19", flush=True)
    p = engine.prepare(prob)
    assert p.E == 19
    assert engine.variant(p)[0] == 4
    first, n = 333, 4000
    X4, px4 = _sample(engine, p, first, n, 5)
    m4 = engine.integrate(p, prob["N"], 5)
    engine.set_variant(p, 3, 256)
    assert engine.variant(p) == (3, 256)
    X3, px3 = _sample(engine, p, first, n, 5)
    m3 = engine.integrate(p, prob["N"], 5)
    order = np.argsort(-X3, axis=1, kind="stable")
    assert np.array_equal(np.argsort(-X4, axis=1, kind="stable"), order)
    assert np.max(np.abs(X4 - X3) / np.abs(X3)) <= 1e-15
    assert np.max(np.abs(px4 - px3) / np.abs(px3)) <= 1e-15
    assert m4.dropped == m3.dropped == 0
    (n4, s14, s24), (n3, s13, s23) = m4.arrays(), m3.arrays()
    assert np.array_equal(n4, n3)
    assert np.all(np.abs(s14 - s13) <= 1e-12 * np.abs(s13))
    assert np.all(np.abs(s24 - s23) <= 1e-12 * np.abs(s23))
    print("   oracle J-table (about 20 s)", flush=True)
    op = oracle.problem(prob)
    st = oracle.seed(5)
    oracle.advance(st, first * (2 * p.E - 1))
    Xo, _, _ = op.sample(st, n)
    assert np.array_equal(np.argsort(-X4, axis=1, kind="stable"), np.argsort(-Xo, axis=1, kind="stable"))
    assert np.max(np.abs(X4 - Xo) / np.abs(Xo)) <= 1e-12
