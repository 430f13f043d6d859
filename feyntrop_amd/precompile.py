"""Ahead-of-time kernel specialisation: compiles (hipRTC, no GPU needed) the
kernels of every named workload into feyntrop_amd/kcache so that GPU runs of
these problems load cached code objects instead of compiling.

    python -m feyntrop_amd.precompile
"""

import concurrent.futures as cf
import os
import sys

from feyntrop_amd import host_path, precompile, precompile_learned, precompile_specialized, workloads


def problems():
    probs = dict(workloads.all_problems())
    probs["ladder7"] = workloads.LARGE["ladder7"]   # the large-graph bench line
    probs["ladder5"] = workloads.LARGE["ladder5"]   # learned proposal on row tables (tests)
    probs["ladder4_mink"] = workloads.LARGE["ladder4_mink"]   # large deformed graph (tests)
    probs["ladder6"] = lambda: workloads.ladder_4pt(6)   # folded rows against row tables (tests)
    for k, v in workloads.scan_problems(host_path).items():
        probs[k] = v
    return probs


def batches():
    """Problem lists whose batched integration the benchmark and the tests
    run: the 4096-point scan and its warm-up, the mixed scan checked against
    the oracle, and the batch CLI cases."""
    from feyntrop_amd import host_accepts
    scan = [c for c in workloads.scan_points(8192, seed=0) if host_accepts(c)][:4096]
    out = [scan, scan[:8]]
    for seed in (21, 33):
        cand = workloads.scan_points(1500, seed=seed)
        seen, pts = set(), []
        for c in cand:
            path = host_path(c)
            if path is not None and (c["_topology"], path) not in seen:
                seen.add((c["_topology"], path))
                pts.append(c)
        out.append(pts)
    for seed, n in ((9, 12), (7, 24), (11, 12)):
        pts = workloads.scan_points(n, seed=seed, accept=host_accepts)
        out.append(pts)
        out.append(pts[1:])
    return out


def lanes_variants():
    """(problem, table mode, block, lanes per sample, specialised) of the
    two-lanes-per-sample kernels the tests and the benchmark run, the folded
    row placement of the placement tests, and every candidate variant of the
    13-edge deformed ladder (the autotuner test)."""
    fx = workloads.FIXTURES
    small = {"1L_3pt_mink": fx["1L_3pt_mink"](), "2L_4pt_ladder": fx["2L_4pt_ladder"](),
             "2L_5pt_box_pent": fx["2L_5pt_box_pent"](), "triple_ladder": workloads.triple_ladder()}
    out = [(prob, 3, 256, 2, False) for prob in small.values()]
    out.append((workloads.triple_ladder(), 3, 256, 2, True))
    db = workloads.double_box()
    out += [(db, 3, blk, 2, False) for blk in (128, 256, 512)] + [(db, 3, 512, 2, True), (db, 3, 256, 2, True)]
    out += [(prob, 4, 256, 1, False) for prob in (workloads.cubic5(), workloads.triple_ladder(), db,
                                                  workloads.wheel_k4())]
    out.append((workloads.ladder_4pt(6), 3, 256, 1, False))
    l4 = workloads.LARGE["ladder4_mink"]()
    out += [(l4, tm, blk, 1, False) for tm in range(4) for blk in (256, 512, 1024)] + [(l4, 3, 256, 1, True)]
    out += [(l4, 3, blk, 2, sp) for blk in (128, 256, 512) for sp in (False, True)]
    return out


def _variant(item):
    from feyntrop_amd import FeyntropError, precompile_variant
    prob, tm, blk, lps, sp = item
    try:
        precompile_variant(prob, tm, blk, lps, sp)
    except FeyntropError as e:
        if e.code != -5:   # infeasible variants are skipped, as the autotuner skips them
            return str(e)
    return None


def _one(item):
    name, fn, learned = item
    try:
        prob = fn() if callable(fn) else fn
        precompile(prob)
        if learned and len(prob["graph"]) <= 16:
            precompile_learned(prob)
        if name in workloads.CONFIGS:   # bench workloads: the kinematics-specialised kernel too
            precompile_specialized(prob)
        return name, None
    except Exception as e:  # report, keep going
        return name, str(e)


def main():
    named = set(workloads.all_problems()) | {"ladder5"}
    # learned-proposal kernels for the named workloads only
    items = [(k, v, k in named) for k, v in problems().items()]
    workers = min(8, os.cpu_count() or 1)
    bad = 0
    with cf.ThreadPoolExecutor(max_workers=workers) as ex:
        for name, err in ex.map(_one, items):
            if err:
                bad += 1
                print(f"precompile {name}: {err}", file=sys.stderr)
    print(f"precompiled {len(items) - bad}/{len(items)} problems")
    with cf.ThreadPoolExecutor(max_workers=workers) as ex:
        for err in ex.map(_variant, lanes_variants()):
            if err:
                bad += 1
                print(f"precompile variant: {err}", file=sys.stderr)
    print("precompiled the lanes-per-sample and autotuner variants")
    from feyntrop_amd import precompile_batch
    for b in batches():
        try:
            precompile_batch(b)
        except Exception as e:  # report, keep going
            bad += 1
            print(f"precompile batch of {len(b)}: {e}", file=sys.stderr)
    print("precompiled the bucket kernels of the benchmark's and the tests' batches")
    return 1 if bad else 0


if __name__ == "__main__":
    sys.exit(main())
