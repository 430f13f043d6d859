"""Derivation of the polynomial in log_pos (ftr_device.hpp): a weighted Remez
exchange for R(z) = 2 atanh(sqrt z)/sqrt z - 2 = z P(z), P of degree d, on
z in [0, zmax], zmax = ((sqrt 2 - 1)/(sqrt 2 + 1))^2, minimising max |z (P - P_true)|.
Prints the coefficients (lowest order first, as hex floats) and the maximal
error for d = 5, 6, 7; log_pos uses d = 6.  Needs mpmath; not part of any
build or test."""
import mpmath as mp
mp.mp.dps = 60
zmax = ((mp.sqrt(2) - 1) / (mp.sqrt(2) + 1)) ** 2
def Ptrue(z):
    if z == 0: return mp.mpf(2) / 3
    s = mp.sqrt(z)
    return (2 * mp.atanh(s) / s - 2) / z
def fit(d, iters=12):
    n = d + 2
    # initial reference: Chebyshev extrema
    xs = [zmax * (1 - mp.cos(mp.pi * k / (n - 1))) / 2 for k in range(n)]
    for it in range(iters):
        # solve z_i (sum c_j z_i^j - P(z_i)) = (-1)^i E  (weighted Remez on weight z)
        A = mp.matrix(n, n); b = mp.matrix(n, 1)
        for i, x in enumerate(xs):
            w = x if x > 0 else mp.mpf('1e-30')
            for j in range(d + 1): A[i, j] = w * x ** j
            A[i, d + 1] = (-1) ** i
            b[i] = w * Ptrue(x)
        sol = mp.lu_solve(A, b)
        c = [sol[j] for j in range(d + 1)]
        err = lambda x: x * (sum(c[j] * x ** j for j in range(d + 1)) - Ptrue(x))
        # new extrema on a fine grid
        grid = [zmax * k / 4000 for k in range(4001)]
        ev = [err(x) for x in grid]
        ext = []
        for k in range(1, len(grid) - 1):
            if (ev[k] - ev[k-1]) * (ev[k+1] - ev[k]) <= 0: ext.append(grid[k])
        cand = [grid[0]] + ext + [grid[-1]]
        if len(cand) >= n:
            # keep n points with alternating sign and largest |err|
            cand.sort()
            xs = cand[:n] if len(cand) == n else sorted(sorted(cand, key=lambda x: -abs(err(x)))[:n])
        mx = max(abs(e) for e in ev)
    return c, mx
for d in (5, 6, 7):
    c, mx = fit(d)
    print(d, 'max |z dP| =', mp.nstr(mx, 5), ' log2 =', mp.nstr(mp.log(mx, 2), 5))
    for x in c: print('   ', float(x).hex(), mp.nstr(x, 20))
