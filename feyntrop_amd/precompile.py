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
    for k, v in workloads.scan_problems(host_path).items():
        probs[k] = v
    return probs


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
    return 1 if bad else 0


if __name__ == "__main__":
    sys.exit(main())
