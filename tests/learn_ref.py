"""numpy restatement of the learned proposal (test infrastructure only): the
network of feyntrop_amd/csrc/kernels/ftr_mlp.hpp in FP64 (optionally with
the kernels' bf16 roundings), its tables, and a replay of the learned
sampler (ftr_device.hpp learn_body) from the xoshiro256+ uniforms.

The learned proposal has no reference counterpart; it reweights the tropical
sampler of feynman_integral.hpp:31-77, whose cut probabilities
J(g\\e)/(w(g\\e) J(g)) and radial densities w y^(w-1) the weight checks use.
"""

import numpy as np

H, NIN, NOUT, MAX_E = 128, 32, 32, 16
DMAX, SMAX = 8.0, 0.6   # a = r e^sigma < 2r: p/q has finite variance
SHAPES = [("W1", (H, NIN)), ("b1", (H,)), ("W2", (H, H)), ("b2", (H,)),
          ("W3", (H, H)), ("b3", (H,)), ("W4", (NOUT, H)), ("b4", (NOUT,))]
NPARAMS = sum(int(np.prod(s)) for _, s in SHAPES)


def unpack(par):
    out, o = {}, 0
    for name, s in SHAPES:
        n = int(np.prod(s))
        out[name] = np.asarray(par[o:o + n], dtype=np.float64).reshape(s)
        o += n
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], dtype=np.float64).reshape(-1) for name, _ in SHAPES])


def bf16(x):
    """float32 -> bf16 (round to nearest even), returned as float64."""
    f = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = f.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return (u.astype(np.uint32)).view(np.float32).astype(np.float64)


def kinematics(P):
    """Scalar products among the vertices with external momentum (rows of P
    with a nonzero entry): upper triangle, row-major, at most 13, scaled by
    max(1, their largest magnitude)."""
    P = np.asarray(P, dtype=np.float64)
    ext = [i for i in range(P.shape[0]) if np.any(P[i] != 0.0)]
    kin = [P[a, b] for ia, a in enumerate(ext) for b in ext[ia:]][:NIN - MAX_E - 3]
    ks = max([1.0] + [abs(v) for v in kin])
    return np.array(kin) / ks


def features(E, r, P=None):
    """Network input per state g: edge bits, popcount/E, 1, r(g)/E, kinematics."""
    n = 1 << E
    M = max(32, n)
    X = np.zeros((M, NIN))
    g = np.arange(n)
    for e in range(E):
        X[:n, e] = (g >> e) & 1
    X[:n, MAX_E] = np.array([bin(int(v)).count("1") for v in g], dtype=np.float64) / E
    X[:n, MAX_E + 1] = 1.0
    X[:n, MAX_E + 2] = (np.asarray(r, dtype=np.float64) / E).astype(np.float32)
    if P is not None:
        kin = kinematics(P)
        X[:n, MAX_E + 3:MAX_E + 3 + len(kin)] = kin
    return X.astype(np.float32).astype(np.float64)


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(par, X, emulate=False):
    """Y and the cache for backward.  emulate: round GEMM operands to bf16 as
    the kernels do (weights, inputs, activations after SiLU)."""
    p = unpack(par)
    rb = bf16 if emulate else (lambda a: a)
    A = rb(X)
    cache = {"A0": A}
    for i in (1, 2, 3):
        Z = A @ rb(p["W%d" % i]).T + p["b%d" % i]
        cache["Z%d" % i] = Z
        A = rb(silu(Z))
        cache["A%d" % i] = A
    Y = A @ rb(p["W4"]).T + p["b4"]
    return Y, cache


def backward(par, cache, dY, emulate=False):
    """Parameter gradient of sum(dY * Y)."""
    p = unpack(par)
    rb = bf16 if emulate else (lambda a: a)
    g = {}
    d = dY
    g["W4"] = rb(d).T @ cache["A3"]
    g["b4"] = d.sum(0)
    dA = rb(d) @ rb(p["W4"])
    for i in (3, 2, 1):
        dZ = dA * dsilu(cache["Z%d" % i])
        g["W%d" % i] = rb(dZ).T @ cache["A%d" % (i - 1)]
        g["b%d" % i] = dZ.sum(0)
        if i > 1:
            dA = rb(dZ) @ rb(p["W%d" % i])
    return pack(g)


def tables(Y, E, w, Jr, r):
    """Learned tables (FP64) from network outputs Y, as mlp_build_tables."""
    n = 1 << E
    lw = np.zeros((n, E))
    g = np.arange(n)
    for e in range(E):
        has = ((g >> e) & 1) == 1
        lw[has, e] = w[g[has] ^ (1 << e)] * np.exp(DMAX * np.tanh(Y[:n][has, e].astype(np.float64) / DMAX))
    lZ = np.zeros(n)
    for e in range(E):          # ascending e, as the kernel sums
        lZ = lZ + lw[:, e]
    lratio = np.zeros((n, E))
    for e in range(E):
        has = ((g >> e) & 1) == 1
        lratio[has, e] = (lZ[has] * w[g[has] ^ (1 << e)]) / (Jr[has] * lw[has, e])
    # radial exponent after cutting e from g: a = r(g\e) e^sigma(g,e), sigma
    # from column E+e; zero where no radius follows (e not in g, or g\e empty)
    lia, lrad, ldiff = np.zeros((n, E)), np.zeros((n, E)), np.zeros((n, E))
    for e in range(E):
        has = (((g >> e) & 1) == 1) & (g != (1 << e))
        c = g[has] ^ (1 << e)
        sg = SMAX * np.tanh(Y[:n][has, E + e].astype(np.float64) / SMAX)
        a = r[c] * np.exp(sg)
        lia[has, e] = 1.0 / a
        lrad[has, e] = -sg
        ldiff[has, e] = r[c] - a
    return dict(lw=lw, lZ=lZ, lratio=lratio, lia=lia, lrad=lrad, ldiff=ldiff)


def replay(t, E, U):
    """Learned sampler on the uniforms U (n x (2E-1), stream order: cut draw
    of step j at 2j, radial draw at 2j+1).  Returns X, weight, cut order."""
    n = U.shape[0]
    full = (1 << E) - 1
    g = np.full(n, full, dtype=np.int64)
    X = np.zeros((n, E))
    lX = np.zeros(n)
    XX = np.ones(n)
    ratio = np.ones(n)
    lrad = np.zeros(n)
    order = np.zeros((n, E), dtype=np.int64)
    rows = np.arange(n)
    for j in range(E):
        R = U[:, 2 * j] * t["lZ"][g]
        T = np.zeros(n)
        sel = np.full(n, -1, dtype=np.int64)
        last = np.zeros(n, dtype=np.int64)
        for e in range(E):
            has = ((g >> e) & 1) == 1
            T = np.where(has, T + t["lw"][g, e], T)
            hit = has & (T > R) & (sel < 0)
            sel[hit] = e
            last[has] = e
        sel = np.where(sel < 0, last, sel)
        order[:, j] = sel
        ratio = ratio * t["lratio"][g, sel]
        c = g ^ (1 << sel)
        X[rows, sel] = XX
        if j < E - 1:
            lY = t["lia"][g, sel] * np.log(1.0 - U[:, 2 * j + 1])
            lX = lX + lY
            XX = np.exp(lX)
            lrad = lrad + t["lrad"][g, sel] + t["ldiff"][g, sel] * lY
        g = c
    return X, ratio * np.exp(lrad), order


def weight_from_path(X, order, t, w, Jr, r):
    """p_trop(path)/q(path) from first principles: cut probabilities
    w(g\\e)/J(g) against lw/lZ, radial densities r y^(r-1) (r of the state
    the previous cut reached) against a y^(a-1) with a = 1/lia of that
    previous cut; y from consecutive X along the path."""
    n, E = X.shape
    logw = np.zeros(n)
    g = np.full(n, (1 << E) - 1, dtype=np.int64)
    rows = np.arange(n)
    prevX = np.ones(n)
    pg, pe = g, order[:, 0]
    for j in range(E):
        e = order[:, j]
        c = g ^ (1 << e)
        logw += np.log(w[c] / Jr[g]) - np.log(t["lw"][g, e] / t["lZ"][g])
        if j > 0:
            y = X[rows, e] / prevX
            ly = np.log(y)
            a = 1.0 / t["lia"][pg, pe]
            rr = r[g]
            logw += (np.log(rr) + (rr - 1) * ly) - (np.log(a) + (a - 1) * ly)
        prevX = X[rows, e]
        pg, pe = g, e
        g = c
    return np.exp(logw)


def uniforms(oracle, seed, first, n, E):
    """The xoshiro256+ uniforms of samples [first, first+n) (2E-1 each)."""
    st = oracle.seed(seed)
    per = 2 * E - 1
    if first:
        oracle.advance(st, first * per)
    u = oracle.next(st, n * per)
    return ((u >> np.uint64(11)).astype(np.float64) * 2.0**-53).reshape(n, per)
