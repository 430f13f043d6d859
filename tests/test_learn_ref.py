"""CPU checks of the numpy restatement of the learned proposal (learn_ref.py),
which the GPU tests use as their checker:
  * the network gradient against central finite differences (FP64);
  * with zero tilt the learned sampler is the tropical sampler: it replays
    the oracle's samples from the same uniforms;
  * the path weight p_trop/q computed two ways (the sampler's running ratio
    and the densities from first principles) agree;
  * E_q[p_trop/q] = 1 (the weights of a random tilt average to one).
"""

import numpy as np
import pytest

import learn_ref as LR
from feyntrop_amd import workloads as W


def _rand_par(rng, scale=0.3):
    return rng.normal(0.0, scale, LR.NPARAMS)


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(1)
    par = _rand_par(rng)
    X = rng.normal(0, 1, (32, LR.NIN))
    dY = rng.normal(0, 1, (32, LR.NOUT))
    Y, cache = LR.forward(par, X)
    g = LR.backward(par, cache, dY)
    idx = rng.choice(LR.NPARAMS, 40, replace=False)
    # one coordinate of every block as well (biases included)
    o = 0
    for _, s in LR.SHAPES:
        idx = np.append(idx, o + int(np.prod(s)) - 1)
        o += int(np.prod(s))
    h = 1e-6
    for i in idx:
        pp, pm = par.copy(), par.copy()
        pp[i] += h
        pm[i] -= h
        fd = (np.sum(dY * LR.forward(pp, X)[0]) - np.sum(dY * LR.forward(pm, X)[0])) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_bf16_rounding():
    x = np.array([1.0, 1.0 + 2**-8, 1.0 + 3 * 2**-9, -2.5, 3.0e38, 1e-40], dtype=np.float32)
    b = LR.bf16(x)
    assert b[0] == 1.0
    assert b[1] == 1.0                          # 1 + 2^-8 is a tie: even is 1.0
    assert b[2] == 1.0 + 2**-7                  # above the tie
    assert b[3] == -2.5


def _zero_tilt_tables(op, E):
    Jr, r, _, _ = op.jtable()
    w = Jr / r
    Y = np.zeros((max(32, 1 << E), 32), np.float32)
    return LR.tables(Y, E, w, Jr, r), w, Jr, r


@pytest.mark.parametrize("name", ["double_box", "triple_ladder", "wheel_k4"])
def test_zero_tilt_replays_tropical_sampler(oracle, name):
    prob = W.all_problems()[name]()
    op = oracle.problem(prob)
    E = op.E
    t, w, Jr, r = _zero_tilt_tables(op, E)
    n = 2000
    U = LR.uniforms(oracle, 5, 0, n, E)
    X, wt, order = LR.replay(t, E, U)
    Xo, _, _ = op.sample(oracle.seed(5), n)
    np.testing.assert_allclose(X, Xo, rtol=1e-12, atol=0)
    np.testing.assert_allclose(wt, 1.0, rtol=1e-13)
    np.testing.assert_allclose(LR.weight_from_path(X, order, t, w, Jr, r), 1.0, rtol=1e-10)


@pytest.mark.parametrize("name", ["double_box", "2L_3pt_tutorial"])
def test_weight_two_ways_and_unit_mean(oracle, name):
    prob = W.all_problems()[name]()
    op = oracle.problem(prob)
    E = op.E
    Jr, r, _, _ = op.jtable()
    w = Jr / r
    rng = np.random.default_rng(3)
    Y = rng.normal(0, 0.7, (max(32, 1 << E), 32)).astype(np.float32)
    t = LR.tables(Y, E, w, Jr, r)
    n = 200000
    U = LR.uniforms(oracle, 11, 0, n, E)
    X, wt, order = LR.replay(t, E, U)
    wp = LR.weight_from_path(X[:5000], order[:5000], t, w, Jr, r)
    np.testing.assert_allclose(wt[:5000], wp, rtol=1e-9)
    mean, err = wt.mean(), wt.std() / np.sqrt(n)
    assert abs(mean - 1.0) <= 4 * err, (mean, err)
    assert err < 0.05
