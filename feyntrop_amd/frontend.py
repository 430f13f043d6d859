"""High-level Python interface on the MI355X engine, in-process.

Mirrors the reference's front-end (py_feyntrop.py) with the same names,
argument meanings and return values:

    from feyntrop_amd.frontend import *
    edges = [((0, 1), 1, 'mm'), ((1, 3), 1, 'mm'), ...]      # (vertices, nu, m^2)
    replacement_rules = [(sp[0, 0], '0'), (sp[0, 1], 'pp2/2'), ...]
    phase_space_point = [('mm', 0.2), ('pp2', 1)]
    trop_res, Itr = tropical_integration(N, D0, Lambda, eps_order, edges,
                                         replacement_rules, phase_space_point)
    print(eps_expansion(trop_res, edges, D0))

Differences, all additions:
- no ./feyntrop subprocess: the problem goes through the C ABI on GPU 0, or
  on FEYNTROP_AMD_DEVICE;
- `seed` and `proposal` keywords, where proposal="learned" trains the
  learned proposal first and lets the plain sampler judge it;
- the kinematic expressions are evaluated with sympy on the phase-space
  point's values, not by textual substitution.

Reference: py_feyntrop.py:19-298.
"""

import math
import os
import sys

from sympy import Float, IndexedBase, Integer, gamma, series, symbols, sympify

import feyntrop_amd

sp = IndexedBase("sp")      # sp[i, j]: scalar product of external momenta i and j
eps = symbols("eps")        # dimensional regulator, D = D0 - 2 eps

_engine = None


def _get_engine():
    global _engine
    if _engine is None:
        _engine = feyntrop_amd.Engine(int(os.environ.get("FEYNTROP_AMD_DEVICE", "0")))
    return _engine


def _vertex_count(edges):
    return len({v for e in edges for v in e[0]})


def _point(phase_space_point):
    """{symbol: value} of the phase-space point (later entries win)."""
    return {symbols(str(k)): sympify(v) for k, v in phase_space_point}


def _evaluate(expr, point):
    """A kinematic expression (string, number or sympy) at the point."""
    val = sympify(expr).subs(point)
    if val.free_symbols:
        raise ValueError(f"kinematic expression {expr!r} has unassigned variables {sorted(map(str, val.free_symbols))}")
    return float(val)


def get_m_sqr(edges, phase_space_point):
    """Squared masses: the third entry of every edge at the point
    (py_feyntrop.py:171-183)."""
    point = _point(phase_space_point)
    return [_evaluate(e[2], point) for e in edges]


def get_P_uv(edges, phase_space_point, replacement_rules):
    """V x V matrix of scalar products of the momenta entering the vertices
    (py_feyntrop.py:186-238).  Externals are vertices 0..n; the rules give
    sp[i, j] for i <= j < n, the last external momentum is minus the sum of
    the others (momentum conservation), internal vertices have zero rows."""
    V = _vertex_count(edges)
    P = [[0.0] * V for _ in range(V)]
    if not replacement_rules:
        return P
    point = _point(phase_space_point)
    block = {}
    n = 0
    for lhs, rhs in replacement_rules:
        i, j = (int(x) for x in lhs.indices)
        i, j = min(i, j), max(i, j)
        block[(i, j)] = _evaluate(rhs, point)
        n = max(n, j + 1)
    if n + 1 > V:
        raise ValueError(f"replacement rules name {n + 1} external vertices but the graph has {V}")
    for i in range(n):
        for j in range(n):
            key = (min(i, j), max(i, j))
            if key not in block:
                raise ValueError(f"no replacement rule for sp[{key[0]},{key[1]}]")
            P[i][j] = block[key]
    for j in range(n):
        P[n][j] = -sum(P[i][j] for i in range(n))
    for i in range(n + 1):
        P[i][n] = -sum(P[i][j] for j in range(n))
    return P


def prepare_kinematic_data(edges, replacement_rules, phase_space_point):
    """(P_uv, m_sqr) for the low-level input (py_feyntrop.py:241-247)."""
    return get_P_uv(edges, phase_space_point, replacement_rules), get_m_sqr(edges, phase_space_point)


def prefactor(edges, D0, eps_order=None):
    """Gamma(omega) / prod Gamma(nu_e) with omega = sum nu - L (D0 - 2 eps)/2,
    the factor between the tropical integral and the Feynman integral
    (py_feyntrop.py:100-112)."""
    nu = [sympify(e[1]) for e in edges]
    L = len(edges) - _vertex_count(edges) + 1
    omega = sum(nu) - L * (sympify(D0) - 2 * eps) / 2
    den = Integer(1)
    for x in nu:
        den *= gamma(x)
    return gamma(omega) / den


def eps_expansion(res, edges, D0):
    """The Feynman integral's Laurent series in eps to the order of res:
    prefactor x sum_k res_k eps^k, coefficients to 10 digits
    (py_feyntrop.py:115-132)."""
    K = len(res)
    body = sum(complex(r[0][0], r[1][0]) * eps ** k for k, r in enumerate(res))
    return series(prefactor(edges, D0) * body, eps, 0, K).evalf(10)


def _round_pair(value, error):
    """Value and error as strings, as the reference prints them
    (py_feyntrop.py:51-63): the number of decimals is |trunc(log10 error)| + 2
    (so 3.4 -> '3.40', 25 -> '25.000', 0.001 -> '0.00100'), both numbers
    rounded to it first; an exact zero error prints both as str()."""
    if error == 0:
        return str(value), str(error)
    decimals = abs(int(math.log10(error))) + 2
    return f"{round(value, decimals):.{decimals}f}", f"{round(error, decimals):.{decimals}f}"


def print_res(res):
    """One line per eps order: [re +/- dre] + i * [im +/- dim] (py_feyntrop.py:68-97)."""
    rows = [(_round_pair(*r[0]), _round_pair(*r[1])) for r in res]
    w = [max(len(row[c][d]) for row in rows) for c in range(2) for d in range(2)]
    print("")
    for k, ((re, dre), (im, dim)) in enumerate(rows):
        print(f"-- eps^{k}: [{re:^{w[0]}} +/- {dre:^{w[1]}}]  +  i * [{im:^{w[2]}} +/- {dim:^{w[3]}}]")


def tropical_integration(N, D0, Lambda, eps_order, edges, replacement_rules, phase_space_point, seed=0,
                         proposal="plain", verbose=True):
    """Tropical Monte Carlo estimate of the integral without its prefactor
    (py_feyntrop.py:250-298).  Returns (trop_res, IGtr), trop_res[k] =
    [[re, dre], [im, dim]] for eps^k.  Raises RuntimeError with the engine's
    message when the input is rejected."""
    P_uv, m_sqr = prepare_kinematic_data(edges, replacement_rules, phase_space_point)
    graph = [[[int(e[0][0]), int(e[0][1])], float(e[1])] for e in edges]
    prob = {"graph": graph, "dimension": D0, "scalarproducts": P_uv, "masses_sqr": m_sqr,
            "num_eps_terms": int(eps_order), "lambda": Lambda, "N": int(N), "seed": int(seed)}
    if verbose:
        print("Starting integration using feyntrop_amd with input:")
        print("Graph with edge weights:", [(tuple(g[0]), g[1]) for g in graph])
        print("Dimension:", D0)
        print("Scalarproducts (matrix element (u,v) is the scalar product of ext. momentum flowing into "
              "vertices u and v):\n", P_uv)
        print("Squared masses:", m_sqr)
        print("Epsilon order:", eps_order)
        print("Deformation parameter Lambda:", Lambda)
        print("Sample points:", N)
    eng = _get_engine()
    try:
        p = eng.prepare(prob)
    except feyntrop_amd.FeyntropError as e:
        raise RuntimeError(f"feyntrop_amd rejected the input: {e}") from None
    if proposal == "plain":
        m = eng.integrate(p, prob["N"], prob["seed"])
    elif proposal == "learned":
        lrn = feyntrop_amd.Learner(eng, p, init_seed=1)
        lrn.train(400, 1 << 18, seed=int(seed) + 100, lr=3e-3)
        use, _, _ = lrn.select(n_probe=min(4_000_000, max(int(N) // 10, 100_000)), seed=int(seed) + 7)
        if verbose:
            print("Proposal:", "learned" if use else "plain (the learned proposal was not better)")
        m = lrn.integrate(prob["N"], prob["seed"])
    else:
        raise ValueError("proposal must be 'plain' or 'learned'")
    r = p.finalize(m)
    trop_res = [[[float(r[k, 0, 0]), float(r[k, 0, 1])], [float(r[k, 1, 0]), float(r[k, 1, 1])]]
                for k in range(p.K)]
    if verbose:
        if m.dropped:
            print(f"Warning: {m.dropped} sampled values were not finite and were dropped.", file=sys.stderr)
        print("Prefactor: " + str(prefactor(edges, D0, eps_order)) + ".")
        print_res(trop_res)
    return trop_res, float(p.info.IGtr)


__all__ = ["sp", "eps", "tropical_integration", "prepare_kinematic_data", "get_P_uv", "get_m_sqr",
           "prefactor", "eps_expansion", "print_res"]
