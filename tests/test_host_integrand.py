"""The kernels' own integrand code on the CPU (test infrastructure, no GPU):
ftr_device.hpp's ftr::integrand compiled with g++ for a problem's generated
graph spec, under the undefined-behaviour and address sanitizers
(-fno-sanitize-recover: any finding aborts the run), evaluated on oracle
sample points and compared with the strict oracle to the bars of the GPU
replay test.  The 3-loop ladder (10 edges, deformed path) is the graph whose
kernel once came out of the GPU compiler with wrong values when its tail
branched on finiteness; a clean sanitized run rules out undefined behaviour
in the kernel's own code as the cause."""

import glob
import os
import re
import subprocess

import numpy as np
import pytest

import feyntrop_amd
from feyntrop_amd import workloads as W

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
KERNELS = os.path.join(ROOT, "feyntrop_amd", "csrc", "kernels")


def _spec(prob, tmp):
    """The generated graph spec (structs FG and FGP) of the problem's kernel
    specialised to its kinematic point, and its (K, MODE)."""
    d = tmp / "src"
    d.mkdir()
    env_old = os.environ.get("FEYNTROP_AMD_DUMP_SRC")
    os.environ["FEYNTROP_AMD_DUMP_SRC"] = str(d)
    try:
        feyntrop_amd.precompile_specialized(prob)
    finally:
        if env_old is None:
            del os.environ["FEYNTROP_AMD_DUMP_SRC"]
        else:
            os.environ["FEYNTROP_AMD_DUMP_SRC"] = env_old
    src = [open(f).read() for f in glob.glob(str(d / "*.hip"))]
    src = [s for s in src if "struct FGP" in s][0]
    start = src.index("\nstruct FGP {")
    spec = src[start:src.index("FTR_INSTANTIATE(ftrk", start)]
    K, MODE = map(int, re.search(r"FTR_INSTANTIATE\(ftrk, FG, (\d+), (\d+),", src).groups())
    return spec, K, MODE


@pytest.mark.parametrize("name", ["triple_ladder", "double_box", "cubic5"])
def test_integrand_host_build_sanitized(oracle, oracle_ld, replay_stats, tmp_path, name):
    prob = W.all_problems()[name]()
    spec, K, MODE = _spec(prob, tmp_path)
    hdr = tmp_path / "spec.hpp"
    hdr.write_text(spec)
    exe = str(tmp_path / "integrand")
    subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-mfma", "-ffp-contract=fast",
                    "-fsanitize=undefined,address", "-fno-sanitize-recover=all",
                    "-I", KERNELS, "-I", HERE, f"-DFTR_SPEC=\"{hdr}\"", f"-DFTR_K={K}", f"-DFTR_MODE={MODE}",
                    "-o", exe, os.path.join(HERE, "integrand_driver.cpp")], check=True)
    op, ol = oracle.problem(prob), oracle_ld.problem(prob)
    n = 20000
    X, px, _ = op.sample(oracle.seed(11), n)
    ref, rdrop, _ = op.replay(X, px)
    tru, _, _ = ol.replay(X, px)
    lines = [str(n)] + [" ".join(float(v).hex() for v in np.concatenate([X[i], px[i]])) for i in range(n)]
    r = subprocess.run([exe], input="\n".join(lines) + "\n", capture_output=True, text=True, timeout=600,
                       env=dict(os.environ, ASAN_OPTIONS="detect_leaks=0"))
    assert r.returncode == 0, r.stderr[-3000:]
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr, r.stderr[-3000:]
    vals = np.array([[float.fromhex(t) for t in l.split()[1:]] for l in r.stdout.split("\n") if l]).reshape(n, K, 2)
    drop = np.array([int(l.split()[0]) for l in r.stdout.split("\n") if l])
    assert np.array_equal(drop.astype(bool), rdrop.astype(bool))
    keep = ~rdrop.astype(bool)
    za = vals[keep, 0, 0] + 1j * vals[keep, 0, 1]
    zr = ref[keep, 0, 0] + 1j * ref[keep, 0, 1]
    zt = tru[keep, 0, 0] + 1j * tru[keep, 0, 1]
    d = np.abs(za - zr) / np.abs(zr)
    e_host, e_ref = np.abs(za - zt) / np.abs(zt), np.abs(zr - zt) / np.abs(zt)
    rs = replay_stats[name]
    print(f"[host integrand] {name}: median {np.median(d):.2e} frac>1e-12 {np.mean(d > 1e-12):.5f} "
          f"(reference build {rs['frac_gt_1e-12']:.5f}) max {d.max():.2e}")
    assert np.median(d) <= 3.0 * max(rs["median"], 5e-15)
    assert np.mean(d > 1e-12) <= 2.0 * rs["frac_gt_1e-12"] + 1e-3
    assert np.mean(e_host > 1e-12) <= np.mean(e_ref > 1e-12) + 2e-4
    for q in (0.5, 0.9, 0.99):
        assert np.quantile(e_host, q) <= 2.0 * np.quantile(e_ref, q) + 2e-15, q
