"""Benchmark: plain tropical sampling throughput on one MI355X.

Workload (N=1): the 2-loop 4-point double box in the deformed Minkowski regime
(7 edges, 6 vertices, D = 6, m^2 = 1, s, t from seed 0, lambda = 0.5, two
eps orders), N = 1e8 samples per step, FP64 complex-LU path.  One step is one
ftr_integrate call: 1e8 samples drawn, evaluated and reduced on the GPU.

Reported: samples/s (value), time to 1e-3 relative error of the eps^0
coefficient extrapolated from the measured error, the FP64 roofline of the
sampling kernel (algorithmic flops per sample x samples per launch / mean
HIP-event duration of the launch), and the reference CPU path (compiled from
its sources, OpenMP on the host cores) on a bounded sample of the same
workload as the baseline.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload double_box]
"""

import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64_TFLOPS = 78.6   # MI355X FP64 (vector and matrix)

# HBM traffic of one ftrk_int launch (N = 1e8) from two rocprofv3 --pmc runs of
# this script (FETCH_SIZE, then WRITE_SIZE; KiB per dispatch, timed launches),
# each scaled by what the same counter read for a 256 MiB device-to-device
# copy in the same run (FETCH 131088.1 KiB, WRITE 262144 KiB for 262144 KiB
# moved).  Per kernel kind: generic (parameters read from memory) and
# specialised to the kinematic point; both with the row-table placement.
# Summary and commands in profiles/ftrk_int_double_box.md.
_FETCH_SCALE = 262144 / 131088.125
PMC_TRAFFIC_BYTES = {
    ("double_box", "generic"): (452.0 * _FETCH_SCALE + 24.0) * 1024,
    ("double_box", "specialised"): (361.5 * _FETCH_SCALE + 24.0) * 1024,
}


def flops_per_sample(info):
    """Algorithmic FP64 flops of one sample (+ - x /, FMA = 2; transcendentals
    excluded), fixed by (E, V, K, path):
      sampler   E(E+1) + 4E
      real      n^3/3 (LDL^T) + 2n^3 (n-RHS solve) + n^2 + V^2 + 12E + 3K
      complex   n^3/3 + 2n^3 (inverse) + 4n^3 (LPGPLT) + 5.5E^2 (derivatives)
                + 8/3 n^3 (complex LU) + 8n^3 (complex n-RHS solve)
                + 8/3 E^3 (complex Jacobian LU) + 2E^2 + 40E + 8K
    with n = V - 1."""
    E, V, K = info.E, info.V, info.K
    n = V - 1
    f = E * (E + 1) + 4 * E
    if info.path == 2:
        f += (n ** 3 / 3 + 2 * n ** 3 + 4 * n ** 3 + 5.5 * E ** 2 + 8 / 3 * n ** 3 + 8 * n ** 3
              + 8 / 3 * E ** 3 + 2 * E ** 2 + 40 * E + 8 * K)
    else:
        f += n ** 3 / 3 + 2 * n ** 3 + n ** 2 + V ** 2 + 12 * E + 3 * K
    return f


def transcendentals_per_sample(info, prob):
    """Transcendental evaluations of one sample on the real paths, which the
    flop count excludes: a log and an exp per radial draw (E - 1), the
    projective rescale's exp, the logs of psi and pxi for the eps expansion
    (K > 1), and a log and an exp per non-integer power (-D/2, -W)."""
    E, K = info.E, info.K
    t = 2 * (E - 1) + 1 + (2 if K > 1 else 0)
    for x in (-prob["dimension"] / 2, -info.W):
        if x != round(x):
            t += 2
    return t


def cpu_baseline(prob, n_samples, threads):
    """The reference binary (oracle/_ref/feyntrop_ref, built from the reference
    sources with its own flags) on a bounded sample of the workload."""
    ref = os.path.join(ROOT, "oracle", "_ref", "feyntrop_ref")
    kind = "reference"
    if not os.path.exists(ref):
        return None
    q = dict(prob)
    q["N"] = int(n_samples)
    env = dict(os.environ)
    env["OMP_NUM_THREADS"] = str(threads)
    t0 = time.time()
    out = subprocess.run([ref], input=json.dumps(q), capture_output=True, text=True, env=env, timeout=600)
    wall = time.time() - t0
    if out.returncode != 0:
        return None
    d = json.loads(out.stdout)
    secs = d["seconds sampling"]
    return {"value": q["N"] / secs, "unit": "samples/s", "cores": threads, "kind": kind,
            "sample": f"same workload, N={q['N']:.0e}, seed 0, OMP_NUM_THREADS={threads} "
                      f"('seconds sampling' {secs:.2f} s, wall {wall:.2f} s)"}


def run_scan(eng, n_points, seed=0):
    """Batched scan: n_points kinematic points (1- to 3-loop topologies,
    Euclidean and Minkowski, N log-uniform in [1e5, 1e7]) prepared together
    (ftr_prepare_batch: host phases in parallel) and integrated in one
    hipGraph of multi-problem kernels.  value = samples / integration wall
    time; end_to_end_value also counts the preparation."""
    import feyntrop_amd
    from feyntrop_amd import workloads as W
    cand = W.scan_points(n_points * 2, seed)          # candidates; the rejected ones are skipped
    pts, probs = [], []
    t0 = time.perf_counter()
    n_cand = 0
    for c0 in range(0, len(cand), 1024):               # prepare in chunks until n_points are accepted
        chunk = cand[c0:c0 + 1024]
        n_cand += len(chunk)
        for pt, q in zip(chunk, eng.prepare_batch(chunk)):
            if not isinstance(q, Exception) and len(probs) < n_points:
                pts.append(pt)
                probs.append(q)
        if len(probs) >= n_points:
            break
    t_prep = time.perf_counter() - t0
    Ns = [p["N"] for p in pts]
    seeds = [p["seed"] for p in pts]
    eng.integrate_batch(probs[:8], [10000] * 8, seeds[:8])       # warm-up
    t0 = time.perf_counter()
    outs = eng.integrate_batch(probs, Ns, seeds)
    wall = time.perf_counter() - t0
    tot = float(sum(Ns))
    paths = {}
    for q in probs:
        paths[q.info.path] = paths.get(q.info.path, 0) + 1
    return {"workload": f"batched scan: {len(pts)} points, {len(set(p['_topology'] for p in pts))} topologies",
            "samples": tot, "value": tot / wall, "unit": "samples/s", "graph_ms": outs[0].kernel_ms,
            "wall_s": wall, "prep_s": t_prep, "prepared_candidates": n_cand,
            "end_to_end_value": tot / (wall + t_prep), "paths": paths,
            "dropped": int(sum(m.dropped for m in outs))}


# HBM bytes per sample of the 22-edge ladder's kernel (folded rows, table
# mode 4; N = 1e8 per launch): rocprofv3 --pmc FETCH_SIZE of this script's
# --workload ladder7 run, 1.856e8 KiB per launch, scaled by the same counter
# for a 256 MiB device copy in that run (131085.4 KiB counted); the sampler's
# row reads are the bound there.  Summary in profiles/ftrk_int_ladder7.md.
LADDER7_HBM_BYTES_PER_SAMPLE = 1.856e8 * (262144 / 131085.375) * 1024 / 1e8
PEAK_HBM_BYTES_S = 8.0e12   # MI355X HBM3E


def run_large(eng, name="ladder7", steps=2):
    """A 22-edge planar ladder: the 2^22-entry J-table built on the device
    (timed against the host pass), then plain sampling with the tables read
    from HBM (folded rows, table mode 4)."""
    from feyntrop_amd import workloads as W
    prob = W.LARGE[name]()
    out = {}
    for mode, tag in ((2, "device"), (1, "host")):
        eng.set_jtable_mode(mode)
        t0 = time.perf_counter()
        p = eng.prepare(prob)
        out[f"prepare_s_{tag}_jtable"] = time.perf_counter() - t0
    eng.set_jtable_mode(0)
    N = int(prob["N"])
    eng.integrate(p, 1_000_000, 1)
    t0 = time.perf_counter()
    ms = [eng.integrate(p, N, 100 + i).kernel_ms for i in range(steps)]
    wall = time.perf_counter() - t0
    fps = flops_per_sample(p.info)
    tm, blk = eng.variant(p)
    out.update({"workload": f"{name}: E={p.E} V={p.V} K={p.K}, 2^{p.E} J-table entries",
                "value": N * steps / wall, "unit": "samples/s", "kernel_ms_mean": float(np.mean(ms)),
                "variant": {"table_mode": tm, "block": blk},
                "roofline_frac": fps * N / (np.mean(ms) * 1e-3) / 1e12 / PEAK_FP64_TFLOPS})
    if name == "ladder7" and tm == 4:
        bps = LADDER7_HBM_BYTES_PER_SAMPLE
        out.update({"bound": "hbm", "hbm_bytes_per_sample_pmc": bps,
                    "hbm_frac": bps * N / (np.mean(ms) * 1e-3) / PEAK_HBM_BYTES_S})
    return out


def run_tuned(eng, name, steps=2):
    """A named workload with the autotuner's kernel variant."""
    from feyntrop_amd import workloads as W
    prob = W.all_problems()[name]()
    p = eng.prepare(prob)
    tm, blk, _ = eng.autotune(p, n_probe=min(int(prob["N"]), 20_000_000))
    eng.integrate(p, prob["N"], 1)
    t0 = time.perf_counter()
    ms = []
    for i in range(steps):
        ms.append(eng.integrate(p, prob["N"], 100 + i).kernel_ms)
    wall = time.perf_counter() - t0
    fps = flops_per_sample(p.info)
    out = {"workload": f"{name}: E={p.E} V={p.V} K={p.K}", "N_per_step": prob["N"],
           "value": prob["N"] * steps / wall, "unit": "samples/s", "kernel_ms_mean": float(np.mean(ms)),
           "variant": {"table_mode": tm, "block": blk, "lanes_per_sample": eng.lanes_per_sample(p)},
           "roofline_frac": fps * prob["N"] / (np.mean(ms) * 1e-3) / 1e12 / PEAK_FP64_TFLOPS}
    if p.info.path != 2:
        # the work the flop count leaves out (real paths: logs and exps)
        t = transcendentals_per_sample(p.info, prob)
        out["transcendentals_per_sample"] = t
        out["transcendentals_per_s_kernel"] = t * prob["N"] / (np.mean(ms) * 1e-3)
    return out


def run_lanes(eng, name="ladder4_mink", steps=2):
    """Lanes per sample on the 13-edge deformed ladder (V = 10), whose
    one-lane kernel spills: one lane and a lane pair per sample, generic and
    specialised to the kinematic point, then the autotuner's choice for a
    fresh preparation of the graph."""
    from feyntrop_amd import workloads as W
    prob = W.LARGE[name]()
    N = int(prob["N"])
    p = eng.prepare(prob)
    fps = flops_per_sample(p.info)
    out = {"workload": f"{name}: E={p.E} V={p.V} K={p.K}, deformed path, N={N:.0e} per call"}
    for lps in (1, 2):
        eng.set_lanes_per_sample(p, lps)
        for spec in (False, True):
            eng.set_specialized(p, spec)
            eng.integrate(p, N // 10, 1)
            ms = float(np.mean([eng.integrate(p, N, 100 + i).kernel_ms for i in range(steps)]))
            tm, blk = eng.variant(p)
            out[f"lps{lps}_{'specialised' if spec else 'generic'}"] = {
                "kernel_ms": ms, "table_mode": tm, "block": blk,
                "roofline_frac": fps * N / (ms * 1e-3) / 1e12 / PEAK_FP64_TFLOPS}
        eng.set_specialized(p, False)
    q = eng.prepare(prob)
    tm, blk, ms = eng.autotune(q, n_probe=2_000_000)
    out["autotune"] = {"table_mode": tm, "block": blk, "lanes_per_sample": eng.lanes_per_sample(q),
                       "specialised": eng.specialized(q), "probe_kernel_ms": ms}
    return out


def _rel_err(p, m):
    r = p.finalize(m)
    I0 = complex(r[0, 0, 0], r[0, 1, 0])
    return float(np.hypot(r[0, 0, 1], r[0, 1, 1]) / abs(I0)), I0, r


def _combined(p, ms):
    """Moments of several runs added into one ftr_moments (independent
    streams: the runs use different seeds)."""
    import feyntrop_amd
    t = feyntrop_amd.ftr_moments()
    t.K = ms[0].K
    for m in ms:
        for k in range(m.K):
            for q in range(2):
                t.n[k][q] += m.n[k][q]
                t.s1[k][q] += m.s1[k][q]
                t.s2[k][q] += m.s2[k][q]
        t.dropped += m.dropped
    if p.info.path != 2:   # the real path's two zero updates, once
        for k in range(t.K):
            t.n[k][1] = 2
    return t


def _run_to(p, fn, target=1e-3, cap_s=90.0, seed0=1000):
    """Calls fn(seed) (one integration of N samples each, new seed each call)
    until the combined eps^0 relative error is <= target; returns (wall s,
    rel. err reached, samples, calls)."""
    ms, t0 = [], time.perf_counter()
    while True:
        ms.append(fn(seed0 + len(ms)))
        rel, _, _ = _rel_err(p, _combined(p, ms))
        wall = time.perf_counter() - t0
        if rel <= target or wall > cap_s:
            return wall, rel, sum(int(m.n[0][0]) + m.dropped for m in ms), len(ms)


def run_learned(eng, name="triple_ladder", steps=400, batch=1 << 18, lr=3e-3, seeds=3):
    """Learned proposal against the plain sampler on the 3-loop ladder: the
    network is trained on steps x batch (about 1e8) proposal samples, plain
    sampling judges it (select), then both integrate N = 1e8 samples with
    `seeds` seeds each.  time-to-1e-3 is reported two ways:
      direct:    wall clock of repeated N-sample calls (new seed each) until
                 the combined rel. error of eps^0 is <= 1e-3 (the learned
                 proposal's time includes its training);
      median:    over the seeds, (rel. err at N / 1e-3)^2 x wall time at N
                 (+ training for the learned proposal).
    Equal wall-clock: the plain error at the learned proposal's median time
    to 1e-3."""
    import feyntrop_amd
    from feyntrop_amd import workloads as W
    prob = W.all_problems()[name]()
    N = int(prob["N"])
    p = eng.prepare(prob)
    # the plain sampler at its fastest: specialised to the kinematic point
    # (as the learned kernels are) with the autotuner's table placement and
    # block size
    eng.set_specialized(p, True)
    tm, blk, _ = eng.autotune(p, n_probe=min(N, 20_000_000))
    eng.integrate(p, 1_000_000, 1)
    plain = []
    for s in range(seeds):
        t0 = time.perf_counter()
        m = eng.integrate(p, N, 11 + 17 * s)
        plain.append((time.perf_counter() - t0, m))
    L = feyntrop_amd.Learner(eng, p, init_seed=1)
    t0 = time.perf_counter()
    st = L.train(steps, batch, seed=100, lr=lr)
    t_train = time.perf_counter() - t0
    use, eff_p, eff_l = L.select(n_probe=4_000_000, seed=7)
    L.integrate(1_000_000, 2)
    learned = []
    for s in range(seeds):
        t0 = time.perf_counter()
        m = L.integrate(N, 12 + 17 * s)
        learned.append((time.perf_counter() - t0, m))
    rel_p = [_rel_err(p, m)[0] for _, m in plain]
    rel_l = [_rel_err(p, m)[0] for _, m in learned]
    t13_p = [(r / 1e-3) ** 2 * t for r, (t, _) in zip(rel_p, plain)]
    t13_l = [(r / 1e-3) ** 2 * t + t_train for r, (t, _) in zip(rel_l, learned)]
    ratios = [a / b for a, b in zip(t13_p, t13_l)]
    # direct: run until 1e-3 (the learned clock starts with its training)
    wall_p, rel_pd, n_pd, c_pd = _run_to(p, lambda sd: eng.integrate(p, N, sd), seed0=5000)
    wall_l, rel_ld, n_ld, c_ld = _run_to(p, lambda sd: L.integrate(N, sd), seed0=6000)
    wall_l += t_train
    Ip = _rel_err(p, _combined(p, [m for _, m in plain]))
    Il = _rel_err(p, _combined(p, [m for _, m in learned]))
    rp, rl = Ip[2], Il[2]
    z = [abs(rl[k, q, 0] - rp[k, q, 0]) / max(np.hypot(rl[k, q, 1], rp[k, q, 1]), 1e-300)
         for k in range(p.K) for q in range(2)]
    fps = flops_per_sample(p.info)
    kp = float(np.mean([m.kernel_ms for _, m in plain]))
    kl = float(np.mean([m.kernel_ms for _, m in learned]))
    med_p, med_l = float(np.median(t13_p)), float(np.median(t13_l))
    return {"workload": f"{name}: E={p.E} V={p.V} K={p.K}, N={N:.0e} per call, lambda={prob['lambda']}",
            "network": "MLP 32-128-128-128-32 SiLU, BF16 MFMA (f32 accumulate), Adam",
            "train_samples": steps * batch, "train_s": t_train,
            "train_rel2_first_last": [float(st[0, 0]), float(st[-1, 0])],
            "selected": "learned" if use else "plain",
            "seeds": seeds,
            "plain": {"kernel": f"specialised, table mode {tm}, block {blk}",
                      "samples_per_s": N / float(np.mean([t for t, _ in plain])), "rel_err": rel_p,
                      "time_to_1e-3_s": t13_p, "kernel_ms": kp,
                      "roofline_frac": fps * N / (kp * 1e-3) / 1e12 / PEAK_FP64_TFLOPS},
            "learned": {"samples_per_s": N / float(np.mean([t for t, _ in learned])), "rel_err": rel_l,
                        "time_to_1e-3_s": t13_l, "kernel_ms": kl,
                        "roofline_frac": fps * N / (kl * 1e-3) / 1e12 / PEAK_FP64_TFLOPS},
            "variance_gain_median": float(np.median([(a / b) ** 2 for a, b in zip(rel_p, rel_l)])),
            "speedup_time_to_1e-3_median": float(np.median(ratios)),
            "speedup_time_to_1e-3_per_seed": ratios,
            "direct_time_to_1e-3": {"plain_s": wall_p, "plain_rel_err": rel_pd, "plain_samples": n_pd,
                                    "learned_s_incl_training": wall_l, "learned_rel_err": rel_ld,
                                    "learned_samples": n_ld, "speedup": wall_p / wall_l},
            "rel_err_at_equal_wallclock": {"wall_s": med_l, "plain": 1e-3 * np.sqrt(med_p / med_l),
                                           "learned": 1e-3},
            "max_sigma_plain_vs_learned": float(max(z)),
            "eps0": {"plain": [Ip[1].real, Ip[1].imag], "learned": [Il[1].real, Il[1].imag]}}


def calibration_copy(nbytes):
    """One device-to-device hipMemcpy of nbytes (the runtime's copy kernel
    reads nbytes and writes nbytes): the reference for scaling the HBM
    traffic counters of the sampling kernel."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(nbytes)) == 0
    assert hip.hipDeviceSynchronize() == 0
    assert hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3) == 0     # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    print(f"calibration copy: {nbytes >> 20} MiB read + {nbytes >> 20} MiB written", file=sys.stderr)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="double_box")
    ap.add_argument("--N", type=float, default=0)
    ap.add_argument("--cpu-samples", type=float, default=0)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true")
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="after the run, one device-to-device copy of this many MiB (the "
                         "reference kernel for scaling HBM traffic counters)")
    ap.add_argument("--scan-points", type=int, default=4096)
    ap.add_argument("--lanes-per-sample", type=int, default=0, choices=(0, 1, 2),
                    help="force one lane or a lane pair per sample (deformed path) for the main workload; "
                         "0: the engine's choice")
    ap.add_argument("--learned-only", action="store_true",
                    help="profiling run: train the learned proposal on the 3-loop ladder and integrate "
                         "N = 1e8 twice with it (one JSON line of kernel times, not the benchmark line)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this engine runs on one GPU; --gpus {args.gpus} is not supported "
              "(sharding the samples across GPUs is not built)", file=sys.stderr)
        sys.exit(2)

    import feyntrop_amd
    from feyntrop_amd import workloads as W

    if args.learned_only:
        eng = feyntrop_amd.Engine(0)
        prob = W.triple_ladder()
        p = eng.prepare(prob)
        L = feyntrop_amd.Learner(eng, p, init_seed=1)
        st = L.train(400, 1 << 18, seed=100, lr=3e-3)
        use, _, _ = L.select(n_probe=4_000_000, seed=7)
        ms = [L.integrate(int(prob["N"]), 12 + i).kernel_ms for i in range(2)]
        print(json.dumps({"workload": "learned proposal, triple_ladder", "selected": "learned" if use else "plain",
                          "train_rel2_last": float(st[-1, 0]), "kernel_ms": ms}), flush=True)
        return

    prob = (W.all_problems().get(args.workload) or W.LARGE[args.workload])()
    if args.N:
        prob["N"] = int(args.N)
    N = int(prob["N"])
    eng = feyntrop_amd.Engine(0)
    p = eng.prepare(prob)
    if args.lanes_per_sample:
        eng.set_lanes_per_sample(p, args.lanes_per_sample)
    # kernel choice (preparation, untimed): the generic kernel (the one the
    # CLI runs) or the one specialised to this kinematic point, whichever is
    # faster on a probe; the other one is timed after the main line
    probe = {}
    for spec in (False, True):
        eng.set_specialized(p, spec)
        eng.integrate(p, min(N, 10**7), 7)
        probe[spec] = eng.integrate(p, min(N, 10**7), 8).kernel_ms
    eng.set_specialized(p, probe[True] < probe[False])
    for i in range(args.warmup):
        eng.integrate(p, N, i + 1)
    kms, res, seeds = [], [], []
    t0 = time.perf_counter()
    for i in range(args.steps):
        m = eng.integrate(p, N, 1000 + i)
        kms.append(m.kernel_ms)
        res.append(m)
    wall = time.perf_counter() - t0
    ms_step = wall / args.steps * 1e3
    value = N * args.steps / wall

    # accuracy of eps^0 from the combined moments of all timed steps
    rr = [p.finalize(m) for m in res]
    I0 = np.mean([r[0, 0, 0] for r in rr]) + 1j * np.mean([r[0, 1, 0] for r in rr])
    d0 = np.mean([np.hypot(r[0, 0, 1], r[0, 1, 1]) for r in rr])
    rel_at_N = d0 / abs(I0)
    t_1e3 = (rel_at_N / 1e-3) ** 2 * N / value

    fps = flops_per_sample(p.info)
    mean_kms = float(np.mean(kms))
    achieved = fps * N / (mean_kms * 1e-3) / 1e12
    line = {
        "metric": "samples/s and time-to-1e-3 rel. err, plain vs learned proposal, 1 MI355X",
        "value": value, "unit": "samples/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic (kinematics from seed 0)",
        "config": {"workload": f"{args.workload}: E={p.E} V={p.V} K={p.K} path="
                               f"{['real', 'real-psi', 'deformed-complex'][p.info.path]}",
                   "N_per_step": N, "lambda": prob["lambda"], "dimension": prob["dimension"],
                   "kernel": "specialised to the kinematic point" if eng.specialized(p) else "generic",
                   "lanes_per_sample": eng.lanes_per_sample(p)},
        "time_to_1e-3_s": t_1e3,
        "rel_err_eps0_at_N": rel_at_N,
        "eps0": [I0.real, I0.imag],
        "roofline": {"bound": "fp64-valu", "achieved": achieved, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "bound_note": "FP64 vector ALU issue (per-sample matrices of order <= 9: no MFMA or HBM "
                                   "work); peak = MI355X FP64 78.6 TFLOP/s",
                     "frac": achieved / PEAK_FP64_TFLOPS,
                     "traffic": PMC_TRAFFIC_BYTES.get((args.workload, "specialised" if eng.specialized(p) else
                                                       "generic")) if N == 10**8 else None,
                     "kernel": "ftrk_int", "flops_per_sample": fps, "kernel_ms_mean": mean_kms},
        "learned": None,
    }
    # the other kernel kind at full N (the CLI runs the generic one)
    other = not eng.specialized(p)
    eng.set_specialized(p, other)
    eng.integrate(p, N, 2)
    oms = [eng.integrate(p, N, 2000 + i).kernel_ms for i in range(3)]
    eng.set_specialized(p, not other)
    okms = float(np.mean(oms))
    line["roofline"]["other_kernel"] = {
        "kernel": "specialised to the kinematic point" if other else "generic (the CLI's)",
        "kernel_ms_mean": okms, "samples_per_s_kernel": N / (okms * 1e-3),
        "frac": fps * N / (okms * 1e-3) / 1e12 / PEAK_FP64_TFLOPS}
    if not args.no_learned:
        try:
            line["learned"] = run_learned(eng)
        except Exception as e:  # reported, not fatal for the main line
            line["learned"] = {"error": str(e)[:300]}
    if not args.no_extras:
        extras = {}
        for key, fn in (("scan", lambda: run_scan(eng, args.scan_points)),
                        ("wheel_k4", lambda: run_tuned(eng, "wheel_k4")),
                        ("cubic5", lambda: run_tuned(eng, "cubic5")),
                        ("triple_ladder", lambda: run_tuned(eng, "triple_ladder")),
                        ("large_graph", lambda: run_large(eng)),
                        ("lanes_per_sample", lambda: run_lanes(eng))):
            try:
                extras[key] = fn()
            except Exception as e:  # an extra must not take the main line down
                extras[key] = {"error": str(e)[:300]}
        line["extra_workloads"] = extras
    if not args.no_cpu:
        # every core this process may run on, capped by OMP_NUM_THREADS when
        # the environment sets it (the host's share of cores for this GPU)
        aff = len(os.sched_getaffinity(0))
        omp = os.environ.get("OMP_NUM_THREADS")
        threads = args.cpu_threads or (min(aff, int(omp)) if omp and omp.isdigit() else aff)
        ns = args.cpu_samples or 2.5e6 * threads
        line["cpu_baseline"] = cpu_baseline(prob, ns, threads)
        if line["cpu_baseline"]:
            line["cpu_baseline"].update({"affinity_cores": aff, "nproc": os.cpu_count(),
                                         "omp_num_threads_env": omp})
        # the same reference binary on one core (bounded: 2e6 samples)
        line["cpu_baseline_1core"] = cpu_baseline(prob, 2e6, 1)
    print(json.dumps(line), flush=True)
    if args.calib_copy_mib:
        calibration_copy(args.calib_copy_mib << 20)


if __name__ == "__main__":
    main()
