"""Autotuner of the launch shapes of the particle stage (phx_tuning, include/phx.h).

For one workload (dimension, interpolation order, particles per cell, precision family) it runs
the level on the GPU and searches the tunable shapes by coordinate descent: each parameter in
turn takes every value of its candidate list while the others keep their best values so far,
each candidate timed over whole predictor-corrector steps (`phx_advance`, the sorts included at
their interval) of a fresh level from the workload's initial state, two passes. The winners go into tuning_table.json beside this file, keyed by
(dim, order, precision, particles per cell); phx.Level applies the matching entry before its
first initialisation, and C++ callers pass the same values to phx_set_tuning.

Tunables (their meaning is in include/phx.h): the segment shape of the cell-segmented storage,
the most segment rows per workgroup of the segmented push, the pitch padding of the LDS tile,
the sort interval, the moment copies of the linear-storage push's tile (the 3D staged path), and
for the mixed-precision family the register window of the deposit.

Usage: python -m phare_amd.autotune --workload harris2d [--precision fp64|fp32] [--write]
"""
import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

from phare_amd import phx  # noqa: E402
from phare_amd import workloads as W  # noqa: E402

TABLE = os.path.join(HERE, "tuning_table.json")

# candidate values per dimension (the 3D staged path pushes segmented storage with one workgroup
# per segment: its segment shape and sort interval)
SPACE = {
    1: {"rows_per_item": [8, 16, 32, 64, 128], "sort_interval": [10, 20, 30, 60]},
    2: {"segment_shape": [[2, 32, 1], [4, 16, 1], [1, 64, 1], [8, 8, 1]],
        "rows_per_item": [16, 24, 32, 48, 64], "tile_row_pad": [1, 8, 16, 32],
        "sort_interval": [15, 20, 30, 40, 60]},
    3: {"segment_shape": [[4, 4, 4], [2, 4, 8]], "tile_row_pad": [1, 4],
        "sort_interval": [5, 10, 15, 30]},
}


def workload(name):
    if name == "harris2d":
        return W.harris2d()
    if name == "uniform1d":
        return W.uniform1d()
    if name.startswith("uniform3d"):
        order, n = 2, 128
        for tok in name.split("-")[1:]:
            if tok.startswith("o"):
                order = int(tok[1:])
            if tok.startswith("n"):
                n = int(tok[1:])
        return W.uniform3d(ncells=(n, n, n), order=order)
    raise SystemExit(f"unknown workload {name}")


def time_steps(lvl, wl, warmup, steps):
    lvl.advance(wl.dt, warmup)
    lvl.sync()
    t0 = time.perf_counter()
    lvl.advance(wl.dt, steps)
    lvl.sync()
    return 1e3 * (time.perf_counter() - t0) / steps


def trial(wl, prec, tuning, warmup, steps):
    """ms per PPC step of the workload under one tuning, from its initial state (each candidate
    runs the same physics)"""
    lvl = phx.Level(wl.dim, wl.order, wl.ncells, wl.dx, masses=wl.masses, Te=wl.Te,
                    capacity=wl.n_particles, precision=prec, apply_tuning_table=False)
    if tuning is not None:
        lvl.set_tuning(**tuning)
    wl.setup(lvl, phx.F_B)
    got = lvl.get_tuning()
    ms = time_steps(lvl, wl, warmup, steps)
    lvl.close()
    return got, ms


def tune(wl, precision, warmup, steps, passes=2, log=print):
    prec = phx.PRECISION_FP32 if precision == "fp32" else phx.PRECISION_FP64
    best, best_ms = trial(wl, prec, None, warmup, steps)
    log(f"[autotune] start {best} -> {best_ms:.4f} ms/step")
    trials = [dict(tuning=dict(best), ms_step=best_ms)]
    space = dict(SPACE[wl.dim])
    if precision == "fp32" and wl.dim < 3:
        space["deposit_window"] = [0, 1]
    for _ in range(passes):
        for key, values in space.items():
            for v in values:
                if best[key] == v:
                    continue
                cand = dict(best)
                cand[key] = v
                _, ms = trial(wl, prec, cand, warmup, steps)
                trials.append(dict(tuning=dict(cand), ms_step=ms))
                log(f"[autotune] {key}={v}: {ms:.4f} ms/step")
                if ms < best_ms:
                    best, best_ms = cand, ms
    return best, best_ms, trials


def write_entry(wl, precision, best, best_ms, path=TABLE):
    table = []
    if os.path.exists(path):
        with open(path) as f:
            table = json.load(f)
    key = dict(dim=wl.dim, order=wl.order, precision=precision, ppc=wl.ppc)
    table = [e for e in table if any(e.get(k) != v for k, v in key.items())]
    table.append(dict(key, workload=f"{wl.name} {'x'.join(map(str, wl.ncells))}",
                      tuning=best, ms_step=best_ms))
    table.sort(key=lambda e: (e["dim"], e["order"], e["precision"], e["ppc"]))
    with open(path, "w") as f:
        json.dump(table, f, indent=1)
        f.write("\n")


def lookup(dim, order, precision, ppc, path=TABLE):
    """the table entry of (dim, order, precision) whose particles per cell are nearest to ppc
    (within a factor of 2), or None"""
    if not os.path.exists(path):
        return None
    with open(path) as f:
        table = json.load(f)
    near = [e for e in table if e["dim"] == dim and e["order"] == order
            and e["precision"] == precision and ppc > 0 and 0.5 <= e["ppc"] / ppc <= 2.0]
    if not near:
        return None
    return min(near, key=lambda e: abs(e["ppc"] - ppc))["tuning"]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="harris2d")
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32"])
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--steps", type=int, default=60)
    ap.add_argument("--passes", type=int, default=2)
    ap.add_argument("--write", action="store_true", help="store the winner in tuning_table.json")
    ap.add_argument("--json", default=None, help="write every trial to this file")
    args = ap.parse_args()
    wl = workload(args.workload)
    best, best_ms, trials = tune(wl, args.precision, args.warmup, args.steps, args.passes,
                                 log=lambda m: print(m, file=sys.stderr, flush=True))
    out = dict(workload=args.workload, precision=args.precision, best=best, ms_step=best_ms,
               trials=trials)
    if args.json:
        with open(args.json, "w") as f:
            json.dump(out, f, indent=1)
    if args.write:
        write_entry(wl, args.precision, best, best_ms)
    print(json.dumps(dict(workload=args.workload, precision=args.precision, best=best,
                          ms_step=best_ms, trials=len(trials))), flush=True)


if __name__ == "__main__":
    main()
