"""Named problem instances (graphs + kinematics) used by tests, smoke and bench.

Each problem is a plain dict with the keys of the feyntrop low-level JSON
input (reference: srcs/feyntrop.cpp:25-96, README.md:86-100):
    graph, dimension, scalarproducts, masses_sqr, num_eps_terms, lambda, N, seed

The seven ``fixture_*`` problems are the reference test suite
(tests/test_suite.py:47-117) restated at the low level: the P matrix follows
the front-end's rule (py_feyntrop.py:186-234: the external block is sp[i,j],
the last external row/column is minus the sums, internal vertices are zero).
Their golden values live in the reference's tests/*.txt and are committed
under tests/golden/.
"""

import random


def _ext_block(V, sp):
    """V x V scalar-product matrix from the (n x n) block sp of the first n
    external momenta; the (n+1)-th external vertex closes momentum
    conservation, all other vertices are internal (zero rows)."""
    n = len(sp)
    P = [[0.0] * V for _ in range(V)]
    for i in range(n):
        for j in range(n):
            P[i][j] = sp[i][j]
    for j in range(n):
        P[n][j] = -sum(P[i][j] for i in range(n))
    for i in range(n + 1):
        P[i][n] = -sum(P[i][j] for j in range(n))
    return P


def _four_point(V, ext, s, t, p2=1.0):
    """Scalar products for 4 external momenta p_i^2 = p2 on vertices ext with
    Mandelstam s, t and u = 4 p2 - s - t (rows sum to zero)."""
    a, b, c, d = ext
    u = 4 * p2 - s - t
    P = [[0.0] * V for _ in range(V)]
    for v in ext:
        P[v][v] = p2
    for (x, y, val) in ((a, b, (s - 2 * p2) / 2), (c, d, (s - 2 * p2) / 2),
                        (a, c, (t - 2 * p2) / 2), (b, d, (t - 2 * p2) / 2),
                        (a, d, (u - 2 * p2) / 2), (b, c, (u - 2 * p2) / 2)):
        P[x][y] = P[y][x] = val
    return P


def _st_from_seed(seed):
    rng = random.Random(seed)
    return rng.uniform(4, 20), rng.uniform(-10, -1)


def _edges(pairs, nu=1):
    return [[[a, b], nu] for (a, b) in pairs]


def wheel_k4(N=10**7, seed=0):
    """K4 wheel with three spokes, Euclidean, masses [0..5]
    (low_level_input.json / README.md:86-100); eps^0 = 6 zeta(3)."""
    return {
        "graph": _edges([(0, 1), (0, 2), (0, 3), (1, 2), (2, 3), (3, 1)]),
        "dimension": 4,
        "scalarproducts": [[-3, 1, 1, 1], [1, -3, 1, 1], [1, 1, -3, 1], [1, 1, 1, -3]],
        "masses_sqr": [0, 1, 2, 3, 4, 5],
        "num_eps_terms": 3, "lambda": 0, "N": N, "seed": seed,
    }


def readme_wheel(N=10**7, seed=0):
    """The shipped low_level_input.json: the K4 wheel with masses [1..6]
    (README.md:54-61 prints its result)."""
    p = wheel_k4(N, seed)
    p["masses_sqr"] = [1, 2, 3, 4, 5, 6]
    return p


DOUBLE_BOX_S, DOUBLE_BOX_T = _st_from_seed(0)


def double_box(N=10**8, seed=0):
    """2-loop 4-point double box, Minkowski, lambda = 0.5, D = 6, m^2 = 1,
    s, t drawn with random.Random(0) in (4,20) x (-10,-1)."""
    return {
        "graph": _edges([(0, 1), (2, 3), (4, 5), (0, 2), (2, 4), (1, 3), (3, 5)]),
        "dimension": 6,
        "scalarproducts": _four_point(6, (0, 1, 4, 5), DOUBLE_BOX_S, DOUBLE_BOX_T),
        "masses_sqr": [1.0] * 7,
        "num_eps_terms": 2, "lambda": 0.5, "N": N, "seed": seed,
    }


# Random cubic 2-point 5-loop topology (stub pairing, seed 0), D = 4,
# p^2 = -1 on vertices 0-1; masses drawn U[0,2] with random.Random(4).
CUBIC5_EDGES = [(1, 5), (4, 6), (0, 2), (3, 7), (6, 7), (8, 9), (1, 2), (2, 8),
                (4, 7), (3, 9), (5, 6), (0, 3), (5, 8), (4, 9)]


def cubic5(N=10**9, seed=0):
    rng = random.Random(4)
    m2 = [rng.uniform(0, 2) for _ in CUBIC5_EDGES]
    P = [[0.0] * 10 for _ in range(10)]
    P[0][0] = P[1][1] = -1.0
    P[0][1] = P[1][0] = 1.0
    return {
        "graph": _edges(CUBIC5_EDGES),
        "dimension": 4,
        "scalarproducts": P,
        "masses_sqr": m2,
        "num_eps_terms": 1, "lambda": 0, "N": N, "seed": seed,
    }


def triple_ladder(N=10**8, seed=0, lam=0.4):
    """3-loop 4-point ladder, Minkowski, lambda = 0.4, D = 6, m^2 = 1,
    s, t as for the double box (random.Random(0))."""
    return {
        "graph": _edges([(0, 1), (2, 3), (4, 5), (6, 7), (0, 2), (2, 4), (4, 6),
                         (1, 3), (3, 5), (5, 7)]),
        "dimension": 6,
        "scalarproducts": _four_point(8, (0, 1, 6, 7), DOUBLE_BOX_S, DOUBLE_BOX_T),
        "masses_sqr": [1.0] * 10,
        "num_eps_terms": 1, "lambda": lam, "N": N, "seed": seed,
    }


def fixture_1L_2pt(N=10**7):
    return {"graph": [[[0, 1], 1], [[0, 1], 2]], "dimension": 4,
            "scalarproducts": [[3.0, -3.0], [-3.0, 3.0]], "masses_sqr": [0.2, 1.2],
            "num_eps_terms": 5, "lambda": 1, "N": N, "seed": 0}


def fixture_1L_3pt_eucl(N=10**7):
    pp0, pp1, s01 = -4, -5, -13
    sp = [[pp0, (s01 - pp0 - pp1) / 2], [(s01 - pp0 - pp1) / 2, pp1]]
    return {"graph": _edges([(0, 1), (1, 2), (2, 0)]), "dimension": 4,
            "scalarproducts": _ext_block(3, sp), "masses_sqr": [0.2, 1.2, 2.2],
            "num_eps_terms": 5, "lambda": 1, "N": N, "seed": 0}


def fixture_1L_3pt_mink(N=10**7):
    pp0, pp1, s01 = 1, 4, 3
    sp = [[pp0, (s01 - pp0 - pp1) / 2], [(s01 - pp0 - pp1) / 2, pp1]]
    return {"graph": _edges([(0, 1), (1, 2), (2, 0)]), "dimension": 4,
            "scalarproducts": _ext_block(3, sp), "masses_sqr": [0.01, 0.5, 1],
            "num_eps_terms": 5, "lambda": 1, "N": N, "seed": 0}


def fixture_2L_2pt_ladder(N=10**7):
    return {"graph": _edges([(0, 2), (0, 3), (1, 2), (1, 3), (2, 3)]), "dimension": 4,
            "scalarproducts": _ext_block(4, [[3]]), "masses_sqr": [0.1, 0.2, 0.3, 0.4, 0.5],
            "num_eps_terms": 4, "lambda": 0.5, "N": N, "seed": 0}


def fixture_2L_4pt_ladder(N=10**7):
    pp0, pp1, pp2, s01, s02, s12 = 1.2, 1.3, 1.4, 2.1, 2.2, 2.3
    sp01, sp02, sp12 = (s01 - pp0 - pp1) / 2, (s02 - pp0 - pp2) / 2, (s12 - pp1 - pp2) / 2
    sp = [[pp0, sp01, sp02], [sp01, pp1, sp12], [sp02, sp12, pp2]]
    return {"graph": _edges([(0, 1), (0, 4), (1, 5), (5, 4), (5, 2), (4, 3), (3, 2)]),
            "dimension": 4, "scalarproducts": _ext_block(6, sp), "masses_sqr": [0] * 7,
            "num_eps_terms": 5, "lambda": 1, "N": N, "seed": 0}


def fixture_2L_3pt_tutorial(N=10**7):
    sp = [[0, 1 / 2], [1 / 2, 0]]
    return {"graph": _edges([(0, 1), (1, 3), (3, 2), (2, 0), (0, 3)]), "dimension": 2,
            "scalarproducts": _ext_block(4, sp), "masses_sqr": [0.2] * 5,
            "num_eps_terms": 5, "lambda": 7.6, "N": N, "seed": 0}


def fixture_2L_5pt_box_pent(N=10**7):
    mm = 1 / 2
    s01, s02, s03, s12, s13, s23 = 2.2, 2.3, 2.4, 2.5, 2.6, 2.7
    sp = [[0, (s01 - mm) / 2, (s02 - mm) / 2, (s03 - mm) / 2],
          [(s01 - mm) / 2, mm, (s12 - 2 * mm) / 2, (s13 - 2 * mm) / 2],
          [(s02 - mm) / 2, (s12 - 2 * mm) / 2, mm, (s23 - 2 * mm) / 2],
          [(s03 - mm) / 2, (s13 - 2 * mm) / 2, (s23 - 2 * mm) / 2, mm]]
    return {"graph": _edges([(0, 1), (1, 2), (2, 6), (6, 3), (3, 4), (4, 5), (5, 0), (5, 6)]),
            "dimension": 6, "scalarproducts": _ext_block(7, sp),
            "masses_sqr": [0, mm, 0, mm, 0, mm, 0, mm],
            "num_eps_terms": 5, "lambda": 0.28, "N": N, "seed": 0}


FIXTURES = {
    "1L_2pt": fixture_1L_2pt,
    "1L_3pt_eucl": fixture_1L_3pt_eucl,
    "1L_3pt_mink": fixture_1L_3pt_mink,
    "2L_2pt_ladder": fixture_2L_2pt_ladder,
    "2L_4pt_ladder": fixture_2L_4pt_ladder,
    "2L_3pt_tutorial": fixture_2L_3pt_tutorial,
    "2L_5pt_box_pent": fixture_2L_5pt_box_pent,
}

def ladder_4pt(rungs=6, N=10**7, seed=0, D=4, mass=1.0, minkowski=False, lam=0.5):
    """Planar 4-point ladder with `rungs` loops (V = 2 (rungs + 1), E =
    3 rungs + 1), all masses equal: a large J-table (2^E entries).
    Euclidean kinematics s = -3, t = -2 on the corner vertices (real path),
    or with minkowski=True the double box's s, t with p_i^2 = 1 and the
    deformation lambda (complex path)."""
    top = list(range(rungs + 1))
    bot = [rungs + 1 + i for i in range(rungs + 1)]
    pairs = [(top[i], bot[i]) for i in range(rungs + 1)]
    pairs += [(top[i], top[i + 1]) for i in range(rungs)] + [(bot[i], bot[i + 1]) for i in range(rungs)]
    V = 2 * (rungs + 1)
    if minkowski:
        s_, t_ = _st_from_seed(0)
        P = _four_point(V, (top[0], bot[0], top[-1], bot[-1]), s_, t_, 1.0)
    else:
        P = _four_point(V, (top[0], bot[0], top[-1], bot[-1]), -3.0, -2.0, -1.0)
    return {"graph": _edges(pairs), "dimension": D, "scalarproducts": P, "masses_sqr": [mass] * len(pairs),
            "num_eps_terms": 1, "lambda": lam if minkowski else 0.0, "N": N, "seed": seed}


CONFIGS = {
    "wheel_k4": wheel_k4,
    "readme_wheel": readme_wheel,
    "double_box": double_box,
    "cubic5": cubic5,
    "triple_ladder": triple_ladder,
}


# large J-tables: device-built 2^E tables, sampler tables read from L2/HBM
LARGE = {
    "ladder5": lambda N=10**7, seed=0: ladder_4pt(5, N, seed),
    "ladder6": lambda N=10**7, seed=0: ladder_4pt(6, N, seed),
    "ladder7": lambda N=10**8, seed=0: ladder_4pt(7, N, seed),
    "ladder4_mink": lambda N=10**7, seed=0: ladder_4pt(4, N, seed, D=6, minkowski=True),
}


def all_problems():
    d = dict(CONFIGS)
    d.update(FIXTURES)
    d.update(DROPS)
    return d


def drop_double_box(N=10**5, seed=0, D=5.5, K=2):
    """The double box's Minkowski kinematics with lambda = 0: no deformation,
    so the real path runs on indefinite kinematics (the reference does the
    same).  The second Symanzik pxi changes sign over the domain: with a
    non-integer W (D = 5.5: W = 1.5) pxi^-W is NaN where pxi < 0 and those
    samples are dropped and counted (feynman_integral.hpp:254-261); with an
    integer W (D = 6) the eps^0 values stay finite but log(pxi) is NaN, so
    the eps^k (k >= 1) sums are NaN and the output prints them as null
    (feyntrop.cpp:264-278)."""
    p = double_box(N, seed)
    p["lambda"] = 0
    p["dimension"] = D
    p["num_eps_terms"] = K
    return p


DROPS = {
    "drop_double_box_D5.5": lambda N=10**5, seed=0: drop_double_box(N, seed, 5.5, 2),
    "drop_double_box_D6": lambda N=10**5, seed=0: drop_double_box(N, seed, 6.0, 2),
}


# Batched scan: 1- to 3-loop topologies with 3-10 edges and 3-8 vertices,
# (edges, external vertices).
SCAN_TOPOLOGIES = {
    "triangle": ([(0, 1), (1, 2), (2, 0)], (0, 1, 2)),
    # 3-loop Mercedes: the K4 wheel (hub 0, rim 1-2-3) with a vertex on
    # each rim edge, 4 external legs
    "mercedes": ([(0, 1), (0, 2), (0, 3), (1, 4), (4, 2), (2, 5), (5, 3), (3, 6), (6, 1)], (1, 4, 5, 6)),
    "triple_ladder": ([(0, 1), (2, 3), (4, 5), (6, 7), (0, 2), (2, 4), (4, 6), (1, 3), (3, 5), (5, 7)],
                      (0, 1, 6, 7)),
    "box": ([(0, 1), (1, 2), (2, 3), (3, 0)], (0, 1, 2, 3)),
    "pentagon": ([(0, 1), (1, 2), (2, 3), (3, 4), (4, 0)], (0, 1, 2, 3, 4)),
    "kite": ([(0, 2), (0, 3), (1, 2), (1, 3), (2, 3)], (0, 1)),
    "twoloop_3pt": ([(0, 1), (1, 3), (3, 2), (2, 0), (0, 3)], (0, 1, 2)),
    "double_box": ([(0, 1), (2, 3), (4, 5), (0, 2), (2, 4), (1, 3), (3, 5)], (0, 1, 4, 5)),
    "twoloop_4pt": ([(0, 1), (0, 4), (1, 5), (5, 4), (5, 2), (4, 3), (3, 2)], (0, 1, 2, 3)),
    "penta_box": ([(0, 1), (1, 2), (2, 6), (6, 3), (3, 4), (4, 5), (5, 0), (5, 6)], (0, 1, 2, 3, 4)),
    "wheel": ([(0, 1), (0, 2), (0, 3), (1, 2), (2, 3), (3, 1)], (0, 1, 2, 3)),
    "threeloop_2pt": ([(0, 2), (0, 3), (2, 3), (2, 4), (3, 5), (4, 5), (4, 1), (5, 1)], (0, 1)),
}


def _scan_P(V, ext, minkowski, rng):
    """Scalar products of random external momenta (4-dimensional), momentum
    conservation on the last external vertex; Euclidean: P = -Gram (negative
    semi-definite); Minkowski: metric (+,-,-,-)."""
    n = len(ext)
    mom = []
    for _ in range(n - 1):
        mom.append([rng.uniform(-1.5, 1.5) for _ in range(4)])
    mom.append([-sum(m[k] for m in mom) for k in range(4)])
    sig = (1.0, -1.0, -1.0, -1.0) if minkowski else (-1.0, -1.0, -1.0, -1.0)
    P = [[0.0] * V for _ in range(V)]
    for a in range(n):
        for b in range(n):
            P[ext[a]][ext[b]] = sum(sig[k] * mom[a][k] * mom[b][k] for k in range(4))
    # exact zero row sums: the last row/column closes conservation
    last = ext[-1]
    for a in range(n - 1):
        P[ext[a]][last] = P[last][ext[a]] = -sum(P[ext[a]][ext[b]] for b in range(n - 1))
    P[last][last] = -sum(P[last][ext[b]] for b in range(n - 1))
    return P


def scan_points(n=4096, seed=0, accept=None):
    """The batched scan: n kinematic points over SCAN_TOPOLOGIES, D in {4, 6},
    m^2 ~ U[0,4], Euclidean (lambda = 0) or Minkowski (lambda in {0.3, 0.7}),
    N log-uniform in [1e5, 1e7], K = 1.  Candidates are drawn from the seeded
    generator until n of them pass `accept` (a callable prob -> bool: the
    preprocessing's acceptance, e.g. feyntrop_amd.host_accepts)."""
    rng = random.Random(seed)
    names = sorted(SCAN_TOPOLOGIES)
    out = []
    while len(out) < n:
        name = names[rng.randrange(len(names))]
        edges, ext = SCAN_TOPOLOGIES[name]
        V = max(max(e) for e in edges) + 1
        mink = rng.random() < 0.5
        D = rng.choice((4, 6))
        prob = {
            "graph": _edges(edges),
            "dimension": D,
            "scalarproducts": _scan_P(V, ext, mink, rng),
            "masses_sqr": [rng.uniform(0, 4) for _ in edges],
            "num_eps_terms": 1,
            "lambda": rng.choice((0.3, 0.7)) if mink else 0,
            "N": int(10 ** rng.uniform(5, 7)),
            "seed": rng.randrange(2 ** 32),
            "_topology": name,
        }
        if accept is None or accept(prob):
            out.append(prob)
    return out


def scan_problems(path_of=None):
    """One representative per (topology, kernel path) of the scan, so that
    their kernels can be specialised ahead of time.  path_of(prob) returns the
    path (0, 1, 2) or None when the preprocessing rejects the point."""
    reps = {}
    for prob in scan_points(3072, seed=12345):
        path = path_of(prob) if path_of else (2 if prob["lambda"] else 0)
        if path is None:
            continue
        key = (prob["_topology"], path)
        if key not in reps:
            reps[key] = prob
    return {f"scan_{k[0]}_path{k[1]}": v for k, v in reps.items()}
