"""The kernels' FP64 elementary functions (feyntrop_amd/csrc/kernels/ftr_math.hpp:
log_pos, log_any, sincos_fast with its Payne-Hanek branch, atan2_fast,
rcp_nr), compiled as plain C++ with g++ and run against mpmath at 50 digits:
  * error bounds in ulp of the exact value (log 1 ulp; sin, cos, atan2 2 ulp
    where the value is not tiny, absolute 2^-52 where it is);
  * the large-argument reduction over the whole exponent range, including
    arguments within 1e-6 of multiples of pi/2;
  * special values as the C library's functions: NaN, infinities, zeros.
On the device v_rcp_f64 seeds the reciprocal instead of the exact quotient;
the two Newton steps make the two agree to an ulp (the GPU tests compare
whole integrands with the oracle)."""

import math
import os
import random
import subprocess

import numpy as np
import pytest

mp = pytest.importorskip("mpmath")
mp.mp.dps = 50

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
KERNELS = os.path.join(ROOT, "feyntrop_amd", "csrc", "kernels")


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("math") / "math_driver")
    subprocess.run(["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-I", KERNELS, "-o", exe,
                    os.path.join(ROOT, "tests", "math_driver.cpp")], check=True)

    def run(fn, *args):
        lines = "".join(f"{fn} " + " ".join(float(v).hex() for v in a) + "\n" for a in zip(*args))
        out = subprocess.run([exe], input=lines, capture_output=True, text=True, check=True).stdout.split("\n")
        return [tuple(float.fromhex(t) for t in l.split()) for l in out if l]
    return run


def _ulp(v):
    v = abs(float(v))
    return math.ulp(v) if v > 0 else math.ulp(0.0)


def _err_ulp(got, exact):
    return abs(mp.mpf(got) - exact) / _ulp(exact)


def test_log_pos_within_one_ulp(driver):
    rng = random.Random(1)
    ys = [1.0, 2.0 ** -53, 1 - 2.0 ** -53, 0.5, 0.7071067811865476, 0.7071067811865475, 2.0, 1e300, 1e-300]
    ys += [(2 ** 53 - rng.getrandbits(53)) * 2.0 ** -53 for _ in range(4000)]
    ys += [rng.uniform(0.70, 0.72) for _ in range(1000)] + [math.exp(rng.uniform(-700, 700)) for _ in range(1000)]
    got = driver("log_pos", ys)
    worst = max(_err_ulp(g[0], mp.log(mp.mpf(y))) if y != 1.0 else abs(g[0]) for g, y in zip(got, ys))
    print(f"log_pos max {float(worst):.3f} ulp")
    assert worst <= 1.0


def test_log_any_special_values(driver):
    xs = [0.0, -0.0, -1.0, float("inf"), float("nan"), 5e-324, 1e-310, 3.0]
    got = [g[0] for g in driver("log_any", xs)]
    assert got[0] == -math.inf and got[1] == -math.inf
    assert math.isnan(got[2]) and math.isnan(got[4])
    assert got[3] == math.inf
    for x, g in zip(xs[5:], got[5:]):
        assert _err_ulp(g, mp.log(mp.mpf(x))) <= 1.0


def _check_sincos(driver, xs, bound):
    got = driver("sincos", xs)
    worst = 0.0
    for (s, c), x in zip(got, xs):
        S, C = mp.sin(mp.mpf(x)), mp.cos(mp.mpf(x))
        for g, v in ((s, S), (c, C)):
            if abs(v) > 1e-3:
                e = _err_ulp(g, v)
            else:
                e = abs(mp.mpf(g) - v) / 2.0 ** -53 / 2   # absolute, in units of 2^-52
            worst = max(worst, float(e))
    return worst


def test_sincos_small_arguments(driver):
    rng = random.Random(2)
    xs = [0.0, 1e-300, 1e-10, 0.5, math.pi / 4, 1.0, 2.0, 3.0, -3.0, 100.0, 1e5, 2.0 ** 20, -(2.0 ** 20)]
    xs += [rng.uniform(-4, 4) for _ in range(3000)] + [rng.uniform(-1e3, 1e3) for _ in range(1000)]
    xs += [float(k * mp.pi / 2) + rng.uniform(-1e-6, 1e-6) for k in range(-300, 300)]
    worst = _check_sincos(driver, xs, 2.0)
    print(f"sincos |x| <= 2^20 max {worst:.3f} ulp")
    assert worst <= 2.0


def test_sincos_large_arguments(driver):
    """The Payne-Hanek branch over the whole exponent range."""
    rng = random.Random(3)
    xs = [2.0 ** 20 * 1.0000001, 1e6, 1e10, 2.0 ** 52, 2.0 ** 53, 2.0 ** 55, 1e22, 1e100, 1e300,
          1.7976931348623157e308, -1e22, -2.0 ** 60]
    xs += [math.ldexp(rng.uniform(1, 2), rng.randint(21, 1023)) * rng.choice((1, -1)) for _ in range(3000)]
    xs += [float(mp.mpf(k) * mp.pi / 2) for k in [2 ** 21 + 3, 10 ** 7 + 1, 10 ** 12 + 7, 10 ** 15 + 3]]
    worst = _check_sincos(driver, xs, 2.0)
    print(f"sincos |x| > 2^20 max {worst:.3f} ulp")
    assert worst <= 2.0


def test_sincos_special_values(driver):
    got = driver("sincos", [float("inf"), -float("inf"), float("nan"), -0.0])
    for s, c in got[:3]:
        assert math.isnan(s) and math.isnan(c)
    s, c = got[3]
    assert s == 0.0 and math.copysign(1.0, s) == -1.0 and c == 1.0


def test_atan2_within_two_ulp(driver):
    rng = random.Random(4)
    pts = [(0.0, 1.0), (1.0, 0.0), (1.0, 1.0), (-1.0, 1.0), (1.0, -1.0), (-1.0, -1.0), (1e-300, 1.0), (1.0, 1e-300)]
    for _ in range(4000):
        r, th = math.exp(rng.uniform(-30, 30)), rng.uniform(-math.pi, math.pi)
        pts.append((r * math.sin(th), r * math.cos(th)))
    ys, xs = zip(*pts)
    got = driver("atan2", ys, xs)
    worst = max(float(_err_ulp(g[0], mp.atan2(mp.mpf(y), mp.mpf(x)))) for g, y, x in zip(got, ys, xs))
    print(f"atan2 max {worst:.3f} ulp")
    assert worst <= 2.0


def test_atan2_signed_zeros(driver):
    ys, xs = [0.0, -0.0, 0.0, -0.0], [0.0, 0.0, -0.0, -0.0]
    got = [g[0] for g in driver("atan2", ys, xs)]
    assert got == [math.atan2(y, x) for y, x in zip(ys, xs)]
    assert [math.copysign(1, g) for g in got] == [math.copysign(1, math.atan2(y, x)) for y, x in zip(ys, xs)]


def test_rcp_within_one_ulp(driver):
    rng = random.Random(5)
    ds = [math.ldexp(rng.uniform(1, 2), rng.randint(-1000, 1000)) for _ in range(3000)] + [1.0, 3.0, 1e-300, 1e300]
    got = driver("rcp", ds)
    worst = max(float(_err_ulp(g[0], 1 / mp.mpf(d))) for g, d in zip(got, ds))
    assert worst <= 1.0
