"""ftr_integrate_noalloc (no reference counterpart: ftr_integrate for callers
that cannot allocate device memory) against ftr_integrate, through the C ABI.

The no-allocation entry point keeps every scratch buffer (block start
states, jump polynomials, block partials, totals) in one fixed device array
per GPU and sums the block partials on the device in block order, the order
of ftr_integrate's host sum.  Same kernels and launch geometry, so the
moments must be equal bit for bit.  Checked on two calls in a row, the
second with a smaller N (the array keeps the first call's contents), on the
deformed and the real path (the 19-edge ladder with its J-table as folded
rows too), and after a call of the allocating entry point; a launch too
large for the array is rejected with FTR_ERR_LIMIT."""

import ctypes as C

import numpy as np
import pytest

from feyntrop_amd import workloads as W

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu


def _call(lib, fn, eng, p, N, seed, lanes=0):
    import feyntrop_amd
    m = feyntrop_amd.ftr_moments()
    opts = feyntrop_amd.ftr_opts(0, -1, int(lanes))
    rc = getattr(lib, fn)(eng.h, p.h, int(N), int(seed), C.byref(opts), C.byref(m))
    return rc, m


def _same(a, b):
    na, s1a, s2a = a.arrays()
    nb, s1b, s2b = b.arrays()
    return (np.array_equal(na, nb) and a.dropped == b.dropped and np.array_equal(s1a, s1b)
            and np.array_equal(s2a, s2b) and a.lanes == b.lanes and a.samples_per_lane == b.samples_per_lane)


PROBLEMS = dict(W.all_problems())
PROBLEMS["ladder6"] = lambda: W.ladder_4pt(6)


@pytest.mark.parametrize("name,N1,N2", [("double_box", 3_000_000, 200_000), ("wheel_k4", 5_000_000, 77_777),
                                        ("drop_double_box_D5.5", 400_000, 65_000), ("ladder6", 2_000_000, 90_001)])
def test_noalloc_equals_integrate(engine, name, N1, N2):
    import feyntrop_amd
    lib = feyntrop_amd.lib()
    print(f"[noalloc] {name} N {N1} then {N2}", flush=True)
    p = engine.prepare(PROBLEMS[name]())
    for N, seed in ((N1, 5), (N2, 6)):
        rc, m0 = _call(lib, "ftr_integrate_noalloc", engine, p, N, seed)
        assert rc == 0, lib.ftr_last_error(engine.h)
        rc, m1 = _call(lib, "ftr_integrate", engine, p, N, seed)
        assert rc == 0, lib.ftr_last_error(engine.h)
        assert _same(m0, m1), (name, N)
        assert m0.n[0][0] + m0.dropped == N
    # once more after the allocating call: the shared array is rewritten
    rc, m2 = _call(lib, "ftr_integrate_noalloc", engine, p, N2, 6)
    assert rc == 0 and _same(m2, m1)


def test_noalloc_rejects_launch_beyond_its_array(engine):
    import feyntrop_amd
    lib = feyntrop_amd.lib()
    p = engine.prepare(W.double_box())
    rc, _ = _call(lib, "ftr_integrate_noalloc", engine, p, 10**8, 1, lanes=64 * 10**6)
    assert rc == -5
    msg = lib.ftr_last_error(engine.h).decode()
    assert "no-allocation" in msg and "scratch" in msg, msg
    # the entry point still works afterwards
    rc, m = _call(lib, "ftr_integrate_noalloc", engine, p, 100_000, 1)
    assert rc == 0 and m.n[0][0] + m.dropped == 100_000
