"""The high-level Python interface (feyntrop_amd/frontend.py) against the
reference's own inputs and outputs:

  kinematics   the seven test-suite problems of tests/test_suite.py:47-117,
               given at the high level (edges with mass expressions,
               replacement rules, phase-space points), must produce the
               low-level inputs whose integrals match the reference's
               tests/*.txt (the fixtures in feyntrop_amd/workloads.py,
               checked on the GPU within 3 sigma); the tutorial notebook's
               printed P_uv and masses (tutorial_2L_3pt.ipynb cell 10)
  prefactor    the notebook's "gamma(2*eps + 3)" (cell 14)
  eps series   the notebook's printed expansion (cell 16) reproduced from
               the tropical coefficients it implies; Gamma-function Taylor
               coefficients from mpmath (numerical derivatives, independent
               of sympy's series); the bubble's 1/eps pole
"""

import json
import os
import subprocess
import sys

import numpy as np
import pytest

sympy = pytest.importorskip("sympy")

from feyntrop_amd import frontend as FE  # noqa: E402
from feyntrop_amd import workloads as W  # noqa: E402

sp = FE.sp

# tests/test_suite.py:47-117, as data
SUITE = {
    "1L_2pt": ([((0, 1), 1, 'mm0'), ((0, 1), 2, 'mm1')], [(sp[0, 0], 'pp')],
               [('mm0', 0.2), ('mm1', 1.2), ('pp', 3)], 4, 5, 1),
    "1L_3pt_eucl": ([((0, 1), 1, 'mm0'), ((1, 2), 1, 'mm1'), ((2, 0), 1, 'mm2')],
                    [(sp[0, 0], 'pp0'), (sp[1, 1], 'pp1'), (sp[0, 1], '(s01-pp0-pp1)/2')],
                    [('mm0', 0.2), ('mm1', 1.2), ('mm2', 2.2), ('pp0', -4), ('pp1', -5), ('s01', -13)], 4, 5, 1),
    "1L_3pt_mink": ([((0, 1), 1, 'mm0'), ((1, 2), 1, 'mm1'), ((2, 0), 1, 'mm2')],
                    [(sp[0, 0], 'pp0'), (sp[1, 1], 'pp1'), (sp[0, 1], '(s01-pp0-pp1)/2')],
                    [('mm0', 0.01), ('mm1', 0.5), ('mm2', 1), ('pp0', 1), ('pp1', 4), ('s01', 3)], 4, 5, 1),
    "2L_2pt_ladder": ([((0, 2), 1, 'mm0'), ((0, 3), 1, 'mm1'), ((1, 2), 1, 'mm2'), ((1, 3), 1, 'mm3'),
                       ((2, 3), 1, 'mm4')], [(sp[0, 0], 'pp')],
                      [('mm0', 0.1), ('mm1', 0.2), ('mm2', 0.3), ('mm3', 0.4), ('mm4', 0.5), ('pp', 3)], 4, 4, 0.5),
    "2L_4pt_ladder": ([((0, 1), 1, '0'), ((0, 4), 1, '0'), ((1, 5), 1, '0'), ((5, 4), 1, '0'), ((5, 2), 1, '0'),
                       ((4, 3), 1, '0'), ((3, 2), 1, '0')],
                      [(sp[0, 0], 'pp0'), (sp[1, 1], 'pp1'), (sp[2, 2], 'pp2'), (sp[0, 1], '(s01-pp0-pp1)/2'),
                       (sp[0, 2], '(s02-pp0-pp2)/2'), (sp[1, 2], '(s12-pp1-pp2)/2')],
                      [('pp0', 1.2), ('pp1', 1.3), ('pp2', 1.4), ('s01', 2.1), ('s02', 2.2), ('s12', 2.3)], 4, 5, 1),
    "2L_3pt_tutorial": ([((0, 1), 1, 'mm'), ((1, 3), 1, 'mm'), ((3, 2), 1, 'mm'), ((2, 0), 1, 'mm'),
                         ((0, 3), 1, 'mm')],
                        [(sp[0, 0], '0'), (sp[1, 1], '0'), (sp[0, 1], 's01/2')], [('mm', 0.2), ('s01', 1)],
                        2, 5, 7.6),
    "2L_5pt_box_pent": ([((0, 1), 1, '0'), ((1, 2), 1, 'mm'), ((2, 6), 1, '0'), ((6, 3), 1, 'mm'),
                         ((3, 4), 1, '0'), ((4, 5), 1, 'mm'), ((5, 0), 1, '0'), ((5, 6), 1, 'mm')],
                        [(sp[0, 0], '0'), (sp[1, 1], 'mm'), (sp[2, 2], 'mm'), (sp[3, 3], 'mm'),
                         (sp[0, 1], '(s01-mm)/2'), (sp[0, 2], '(s02-mm)/2'), (sp[0, 3], '(s03-mm)/2'),
                         (sp[1, 2], '(s12-2*mm)/2'), (sp[1, 3], '(s13-2*mm)/2'), (sp[2, 3], '(s23-2*mm)/2')],
                        [('mm', 1 / 2), ('s01', 2.2), ('s02', 2.3), ('s02', 2.3), ('s03', 2.4), ('s12', 2.5),
                         ('s13', 2.6), ('s23', 2.7)], 6, 5, 0.28),
}


@pytest.mark.parametrize("name", list(SUITE))
def test_kinematics_of_reference_suite(name):
    edges, rules, point, D0, K, lam = SUITE[name]
    P, m2 = FE.prepare_kinematic_data(edges, rules, point)
    fx = W.FIXTURES[name]()
    np.testing.assert_allclose(np.array(P), np.array(fx["scalarproducts"], dtype=float), rtol=0, atol=1e-15)
    np.testing.assert_allclose(m2, fx["masses_sqr"], rtol=0, atol=1e-15)
    assert [list(e[0]) for e in edges] == [g[0] for g in fx["graph"]]
    assert (D0, K, lam) == (fx["dimension"], fx["num_eps_terms"], fx["lambda"])


def test_tutorial_notebook_kinematics_and_prefactor():
    edges = [((0, 1), 1, 'mm'), ((1, 3), 1, 'mm'), ((3, 2), 1, 'mm'), ((2, 0), 1, 'mm'), ((0, 3), 1, 'mm')]
    rules = [(sp[0, 0], '0'), (sp[1, 1], '0'), (sp[0, 1], 'pp2/2')]
    P, m2 = FE.prepare_kinematic_data(edges, rules, [('mm', 0.2), ('pp2', 1)])
    assert P == [[0, 0.5, -0.5, 0], [0.5, 0, -0.5, 0], [-0.5, -0.5, 1.0, 0], [0, 0, 0, 0]]
    assert m2 == [0.2] * 5
    assert str(FE.prefactor(edges, 2, 5)) == "gamma(2*eps + 3)"


def _coeffs(expr, K, lo=0):
    expr = expr.removeO()
    return [complex(expr.coeff(FE.eps, k)) for k in range(lo, K)]


def test_eps_expansion_reproduces_notebook():
    """Cell 16 printed prefactor x (tropical series) to 10 digits; the
    tropical coefficients it implies, fed back in, give it again."""
    edges = [((0, 1), 1, 'mm'), ((1, 3), 1, 'mm'), ((3, 2), 1, 'mm'), ((2, 0), 1, 'mm'), ((0, 3), 1, 'mm')]
    printed = [complex(-93.1295061, 174.3662845), complex(-720.6568699, 544.4786666),
               complex(-2115.10607, 497.0359226), complex(-3571.773163, -676.3796294),
               complex(-3872.607825, -2725.212451)]
    # prefactor Taylor coefficients gamma(3 + 2 eps) = sum c_j eps^j
    c = _coeffs(sympy.series(sympy.gamma(3 + 2 * FE.eps), FE.eps, 0, 5).evalf(30), 5)
    I = []
    for k in range(5):   # invert the triangular product
        I.append((printed[k] - sum(c[j] * I[k - j] for j in range(1, k + 1))) / c[0])
    res = [[[v.real, 0.0], [v.imag, 0.0]] for v in I]
    # the implied eps^0 coefficient is the notebook's printed tropical result
    assert abs(I[0].real - (-46.56)) < 0.01 and abs(I[0].imag - 87.18) < 0.01
    got = _coeffs(FE.eps_expansion(res, edges, 2), 5)
    for a, b in zip(got, printed):
        assert abs(a - b) <= 1e-9 * abs(b) + 1e-7, (a, b)


def test_eps_expansion_gamma_taylor_mpmath():
    """Prefactor series against mpmath's numerical derivatives of Gamma:
    omega = 3 + 2 eps (tutorial), coefficients 2^k Gamma^(k)(3)/k!."""
    mpmath = pytest.importorskip("mpmath")
    edges = [((0, 1), 1, 'a'), ((1, 3), 1, 'a'), ((3, 2), 1, 'a'), ((2, 0), 1, 'a'), ((0, 3), 1, 'a')]
    res = [[[1.0, 0.0], [0.0, 0.0]]] + [[[0.0, 0.0], [0.0, 0.0]]] * 4
    got = _coeffs(FE.eps_expansion(res, edges, 2), 5)
    for k in range(5):
        ref = float(2 ** k * mpmath.diff(mpmath.gamma, 3, k) / mpmath.factorial(k))
        assert abs(got[k].real - ref) <= 1e-8 * abs(ref), (k, got[k], ref)


def test_eps_expansion_bubble_pole():
    """Massive bubble in D0 = 4: omega = eps, Gamma(eps) = 1/eps - gamma_E +
    (gamma_E^2/2 + pi^2/12) eps + ..."""
    edges = [((0, 1), 1, 'm'), ((0, 1), 1, 'm')]
    res = [[[1.0, 0.0], [0.0, 0.0]], [[0.0, 0.0], [0.0, 0.0]]]
    ex = FE.eps_expansion(res, edges, 4).removeO()
    ge = 0.57721566490153286
    assert abs(complex(ex.coeff(FE.eps, -1)) - 1.0) < 1e-9
    assert abs(complex(ex.coeff(FE.eps, 0)) + ge) < 1e-9
    assert abs(complex(ex.coeff(FE.eps, 1)) - (ge ** 2 / 2 + np.pi ** 2 / 12)) < 1e-8


def test_print_res_rounds_to_error(capsys):
    FE.print_res([[[-46.5647, 0.1312], [87.1831, 0.1201]], [[-274.3912, 0.0055], [1.0, 0.0]]])
    out = capsys.readouterr().out.strip().splitlines()
    assert out[0].startswith("-- eps^0: [") and "-46.56" in out[0] and "0.13" in out[0] and "87.18" in out[0]
    assert "-274.3912" in out[1] and "0.0055" in out[1]


@pytest.mark.parametrize("value,error,expect", [
    (12.3456, 3.4, ("12.35", "3.40")),          # error >= 1: |trunc(log10)| + 2 decimals
    (1234.56789, 25.0, ("1234.568", "25.000")),
    (0.123456, 0.001, ("0.12346", "0.00100")),   # exact power of ten
    (-46.5647, 0.1312, ("-46.56", "0.13")),
    (5.0, 0.0, ("5.0", "0.0")),
])
def test_round_pair_is_the_reference_rule(value, error, expect):
    """py_feyntrop.py:51-63: decimals = |int(log10(err))| + 2, both rounded."""
    assert FE._round_pair(value, error) == expect


def test_missing_rule_is_an_error():
    edges = [((0, 1), 1, 'm'), ((1, 2), 1, 'm'), ((2, 0), 1, 'm')]
    with pytest.raises(ValueError):
        FE.get_P_uv(edges, [('m', 1), ('a', 1)], [(sp[0, 0], 'a'), (sp[1, 1], 'a')])   # sp[0,1] missing
    with pytest.raises(ValueError):
        FE.get_m_sqr(edges, [('q', 1)])                                                  # m unassigned


@pytest.mark.gpu
@pytest.mark.parametrize("name,proposal", [("1L_2pt", "plain"), ("2L_3pt_tutorial", "plain"),
                                           ("2L_3pt_tutorial", "learned")])
def test_frontend_tropical_integration_matches_reference(golden_fixtures, name, proposal):
    """tropical_integration on the reference's own test-suite inputs: every
    eps coefficient within 3 combined standard errors of tests/*.txt."""
    edges, rules, point, D0, K, lam = SUITE[name]
    print(f"[frontend] {name} {proposal}", flush=True)
    res, Itr = FE.tropical_integration(10**7, D0, lam, K, edges, rules, point, proposal=proposal,
                                       verbose=False)
    gold = golden_fixtures[name]
    for k in range(K):
        for q in range(2):
            a, da = res[k][q]
            b, db = gold[k][q]
            assert abs(a - b) <= 3.0 * np.hypot(da, db) + 1e-12, (k, q, a, da, b, db)
