"""Problems with 13 or more edges for diagnostics parity on the device path:
on a GPU host the product builds their J-tables on the device (ftr_prepare
uses the device from E = 13 on), so the regime, GP-property, exceptional-
kinematics and error lines come from the device pass there.  The expected
lines are the reference binary's (feyntrop.cpp:165-231), recorded once by
tests/golden/make_golden.py into tests/golden/cli.json."""

from feyntrop_amd import workloads as W


def _ladder(rungs, s, t, p2, masses, D=4, lam=0.5, K=1):
    p = W.ladder_4pt(rungs, N=1000, seed=0, D=D)
    V = 2 * (rungs + 1)
    top, bot = list(range(rungs + 1)), [rungs + 1 + i for i in range(rungs + 1)]
    p["scalarproducts"] = W._four_point(V, (top[0], bot[0], top[-1], bot[-1]), s, t, p2)
    p["masses_sqr"] = list(masses) if isinstance(masses, (list, tuple)) else [masses] * len(p["graph"])
    p["lambda"] = lam
    p["num_eps_terms"] = K
    return p


def large_cases():
    s, t = W.DOUBLE_BOX_S, W.DOUBLE_BOX_T
    out = {
        "cubic5": W.cubic5(N=1000),
        # 16 edges; spacelike legs with u > 0: pseudo-Euclidean
        "ladder5_16_edges": W.LARGE["ladder5"](N=1000),
        "ladder4_mink": W.LARGE["ladder4_mink"](N=1000),
        # spacelike legs and s, t < 0 but u > 0: pseudo-Euclidean, so the
        # requested deformation is not applied
        "ladder4_pseudo_eucl": _ladder(4, -3.0, -2.0, -1.0, 1.0, D=6, lam=0.5),
        # massless propagators with on-shell massless legs: exceptional
        # kinematics, GP property not fulfilled
        "ladder4_massless": _ladder(4, s, t, 0.0, 0.0, D=6, lam=0.5),
        # half the propagators massless, massive legs
        "ladder4_half_massless": _ladder(4, s, t, 1.0, [0.0, 1.0] * 6 + [0.0], D=6, lam=0.5, K=2),
        # every box is UV divergent at D = 8: the subdivergence error
        "ladder4_subdivergent": _ladder(4, -3.0, -2.0, -1.0, 1.0, D=8, lam=0.0),
    }
    return out
