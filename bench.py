#!/usr/bin/env python3
"""Benchmark of the hybrid-PIC ion update + field solve on one MI355X.

A step is one full predictor-corrector step (SolverPPC::advanceLevel: two updatePopulations
passes with push + gather + deposit, and the Faraday/Ampere/Ohm field solve), replayed from a
captured hipGraph (2D) or launched stage by stage (3D), over one periodic level resident in HBM.
Default workload: the 3D uniform Maxwellian, 128^3 cells, 64 particles per cell (134M particles),
interpolation order 2, plain FP64 (the largest single-patch configuration); orders 1 and 3 and the
2D Harris sheet (256x128 cells, 100 ppc, both precision families) are reported under extra keys.

Prints one JSON line: particle-steps/s, the roofline of the dominant kernel (the fused
push+gather+deposit kernel, algorithmic bytes per launch / its mean duration over the timed
steps), the same for the mixed-precision family, and the CPU baseline (the FP64 oracle
restatement of the reference method, 1 thread, timed on a bounded sample of the same workload).

The push's duration is taken inside the timed steps: in 2D each launch stamps the device wall clock
at its first workgroup's start and its last workgroup's end (a HIP event recorded inside a captured
step would keep only the last replay's time); in 3D HIP events bracket each push launch. rocprofv3's
kernel trace of the same command agrees (profiles/).
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from phare_amd import phx  # noqa: E402
from phare_amd import workloads as W  # noqa: E402

HBM_PEAK_GBS = 8000.0

# HBM bytes per push dispatch (mode all) from rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE runs of
# this script, FETCH_SIZE scaled x2 and WRITE_SIZE x1 by a 1 GiB device copy's readings; keyed by
# workload and precision family (k_push_seg / k_push_seg32 in 2D, k_push_segwg / k_push32_segwg on
# the 3D staged path)
MEASURED_TRAFFIC = {("harris2d", "fp64"): 395.0e6, ("harris2d", "fp32"): 224.3e6,
                    ("uniform3d", "fp64"): 19.41e9, ("uniform3d", "fp32"): 11.03e9}


def _hip():
    for name in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
        try:
            return ctypes.CDLL(name)
        except OSError:
            continue
    raise RuntimeError("HIP runtime not found")


def device_sync():
    rc = _hip().hipDeviceSynchronize()
    if rc != 0:
        raise RuntimeError(f"hipDeviceSynchronize failed: {rc}")


def calibrate_copy(nbytes):
    """one device-to-device copy of nbytes (HIP's blit kernel), so that the FETCH_SIZE /
    WRITE_SIZE counters of a kernel with known traffic can scale the push kernel's readings"""
    hip = _hip()
    src, dst = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(src), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(ctypes.byref(dst), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMemset(src, 1, ctypes.c_size_t(nbytes)) == 0
    for _ in range(3):
        assert hip.hipMemcpy(dst, src, ctypes.c_size_t(nbytes), 3) == 0  # device to device
    device_sync()
    hip.hipFree(src)
    hip.hipFree(dst)
    print(json.dumps({"calibration": "device-to-device copy", "bytes": nbytes, "copies": 3}))


def workload(name):
    if name == "harris2d":
        return W.harris2d()
    if name == "uniform1d":
        return W.uniform1d()
    if name.startswith("uniform3d"):
        # uniform3d[-o<order>][-n<cells per side>]
        order, n = 2, 128
        for tok in name.split("-")[1:]:
            if tok.startswith("o"):
                order = int(tok[1:])
            if tok.startswith("n"):
                n = int(tok[1:])
        return W.uniform3d(ncells=(n, n, n), order=order)
    raise SystemExit(f"unknown workload {name}")


def record_bytes(wl, precision):
    """(R, M): bytes of a particle record read, and of its mutable part written back. FP64 SoA:
    iCell i32 x D, delta f64 x D, v f64 x 3, weight and charge f64; mixed precision: iCell i32 x D,
    delta f32 x D, v f32 x 3, weight f32 (the charge is the population's)"""
    D = wl.dim
    if precision == "fp32":
        return 8 * D + 16, 8 * D + 12
    return 12 * D + 40, 12 * D + 24


def push_bytes(wl, mode, precision="fp64"):
    """algorithmic HBM bytes of one fused push+deposit launch: the record read, in mode all the
    write-back of its mutable part, plus 6 E/B component reads and 5 moment writes per cell"""
    R, M = record_bytes(wl, precision)
    per_particle = R + (M if mode == phx.MODE_ALL else 0)
    return wl.n_particles * per_particle + wl.n_cells * 11 * 8


def step_bytes(wl, precision="fp64"):
    """algorithmic bytes of one PPC step: N_p (2R+M) + N_c * 8 * (189 + 20 npop)"""
    R, M = record_bytes(wl, precision)
    return wl.n_particles * (2 * R + M) + wl.n_cells * 8 * (189 + 20 * len(wl.masses))


def run_family(wl, precision, steps, warmup, drift_steps=0):
    """steps/s of one precision family on the workload, and its push kernel's time measured
    inside the timed steps"""
    prec = phx.PRECISION_FP32 if precision == "fp32" else phx.PRECISION_FP64
    lvl = phx.Level(wl.dim, wl.order, wl.ncells, wl.dx, masses=wl.masses, Te=wl.Te,
                    capacity=wl.n_particles, precision=prec)
    t_setup = time.perf_counter()
    wl.setup(lvl, phx.F_B)
    diag0 = lvl.diagnostics()
    print(f"[bench] {precision}: level loaded and initialised in {time.perf_counter() - t_setup:.1f} s",
          file=sys.stderr, flush=True)
    # the pushes of the captured steps time themselves (first workgroup start to last workgroup
    # end, device wall clock); enabled before the warm-up so that the timed steps replay the
    # step captured with it, and reset right before the timed steps
    lvl.set_push_timing(True)
    lvl.advance(wl.dt, warmup)
    lvl.sync()
    device_sync()
    print(f"[bench] {precision}: {warmup} warm-up steps done", file=sys.stderr, flush=True)
    lvl.set_push_timing(True)
    t0 = time.perf_counter()
    lvl.advance(wl.dt, steps)
    lvl.sync()
    device_sync()
    elapsed = time.perf_counter() - t0
    print(f"[bench] {precision}: {steps} timed steps in {elapsed:.2f} s", file=sys.stderr, flush=True)
    ms_all, n_all = lvl.push_timing(phx.MODE_ALL)
    ms_dom, n_dom = lvl.push_timing(phx.MODE_DOMAIN_ONLY)
    assert n_all == steps and n_dom == steps, (n_all, n_dom, steps)
    diag1 = lvl.diagnostics()
    # untimed: continue to the workload's full run length for its energy drift
    done = warmup + steps
    if drift_steps > done:
        lvl.advance(wl.dt, drift_steps - done)
        done = drift_steps
    diagN = lvl.diagnostics()
    lvl.close()
    e0 = diag0["kinetic"] + diag0["magnetic"]
    e1 = diag1["kinetic"] + diag1["magnetic"]
    eN = diagN["kinetic"] + diagN["magnetic"]
    return dict(value=wl.n_particles * steps / elapsed, ms_step=1e3 * elapsed / steps,
                ms_all=ms_all, ms_dom=ms_dom,
                energy_drift=(e1 - e0) / e0,
                charge_drift=(diag1["charge"] - diag0["charge"]) / diag0["charge"],
                steps_done=warmup + steps,
                energy_drift_run=(eN - e0) / e0, run_steps=done,
                charge_drift_run=(diagN["charge"] - diag0["charge"]) / diag0["charge"])


def cpu_baseline(wl, steps, runs=3):
    """the oracle (FP64 restatement of the reference method: AoS particles, CellMap, the same
    loops) on one host thread, on `steps` PPC steps of the workload after one warm-up step; the
    median rate of `runs` runs"""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O

    rates, secs = [], []
    for _ in range(runs):
        sim = O.Sim(wl.dim, wl.order, wl.ncells, wl.dx, masses=wl.masses, Te=wl.Te)
        wl.setup(sim, O.F_B)
        sim.advance(wl.dt, 1)
        t0 = time.perf_counter()
        sim.advance(wl.dt, steps)
        el = time.perf_counter() - t0
        rates.append(wl.n_particles * steps / el)
        secs.append(el)
        del sim
    med = float(np.median(rates))
    return dict(value=med, unit="particle-steps/s", cores=1, kind="port",
                sample=f"{wl.name} {'x'.join(map(str, wl.ncells))} o{wl.order} {wl.ppc} ppc "
                       f"({wl.n_particles} particles): {steps} PPC steps after 1 warm-up step, "
                       f"median of {runs} runs ({', '.join(f'{x:.1f}' for x in secs)} s), "
                       f"1 thread, no MPI",
                seconds=float(np.sum(secs)))


def cpu_sample(wl):
    """a bounded CPU sample of a workload: the same profiles, order and particles per cell on a
    smaller box when the workload is too large for ~10-30 s of one host core"""
    if wl.name == "uniform3d" and wl.n_cells > 32 ** 3:
        return W.uniform3d(ncells=(32, 32, 32), ppc=wl.ppc, order=wl.order, dx=wl.dx[0]), 3
    return wl, 20


def family_line(wl, workload_name, precision, res):
    """the throughput and the push roofline of one run_family result"""
    b = push_bytes(wl, phx.MODE_ALL, precision)
    a = b / (res["ms_all"] * 1e-3) / 1e9
    if wl.dim == 3:
        kernel = ("k_push32_segwg (mixed precision" if precision == "fp32" else "k_push_segwg (FP64") + \
            ", cell-segmented storage, one workgroup per 4x4x4-cell segment: prePush+gather+Boris+" \
            "postPush+deposit, mode all; 3D staged path)"
    elif precision == "fp32":
        kernel = "k_push_seg32 (mixed precision, mode all)"
    else:
        kernel = "k_push_seg (fused prePush+gather+Boris+postPush+deposit on cell-segmented storage, mode all)"
    return {"value": res["value"], "unit": "particle-steps/s", "ms_per_step": res["ms_step"],
            "roofline": {"bound": "hbm", "achieved": a, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": a / HBM_PEAK_GBS,
                         "traffic": MEASURED_TRAFFIC.get((workload_name, precision)),
                         "kernel": kernel, "kernel_ms": res["ms_all"],
                         "kernel_ms_domain_only": res["ms_dom"], "algorithmic_bytes_per_launch": b},
            "step_hbm_fraction": step_bytes(wl, precision) / (res["ms_step"] * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "energy_drift": res["energy_drift"], "charge_drift": res["charge_drift"],
            "run": {"steps": res["run_steps"], "energy_drift": res["energy_drift_run"],
                    "charge_drift": res["charge_drift_run"]}}


def describe(wl, precision):
    return (f"{wl.name} {'x'.join(map(str, wl.ncells))} cells, order {wl.order}, {wl.ppc} ppc, "
            f"{wl.n_particles} particles, dt={wl.dt}, {'plain FP64' if precision == 'fp64' else 'mixed FP32/FP64'}, "
            f"one periodic patch")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="uniform3d",
                    help="uniform3d[-o<order>][-n<cells>] (3D uniform, 128^3 x 64 ppc: the largest "
                         "single-patch configuration), harris2d or uniform1d")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision family")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the extra lines (2D Harris both families; 3D orders 1 and 3)")
    ap.add_argument("--no-drift-run", action="store_true",
                    help="skip the untimed continuation to the workload's run length (energy drift)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="BYTES",
                    help="only run a device-to-device copy of BYTES (counter calibration)")
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
        return
    if args.gpus != 1:
        print("bench.py runs on one GPU: this build has no multi-GPU path (patch sharding over "
              "xGMI is a design, not built); use --gpus 1", file=sys.stderr)
        sys.exit(2)

    wl = workload(args.workload)
    drift_steps = 0 if args.no_drift_run else wl.nsteps
    print(f"[bench] {describe(wl, 'fp64')}: {args.warmup} + {args.steps} steps", file=sys.stderr,
          flush=True)
    f64 = run_family(wl, "fp64", args.steps, args.warmup, drift_steps)
    line = family_line(wl, args.workload, "fp64", f64)
    out = {
        "metric": "particle-steps/s (push+gather+deposit)",
        "value": f64["value"],
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": f64["ms_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": describe(wl, "fp64"), "particles": wl.n_particles,
                   "cells": wl.n_cells, "interp_order": wl.order, "precision": "fp64",
                   "parallelism": "single GPU"},
        "roofline": line["roofline"],
        "step_hbm_fraction": line["step_hbm_fraction"],
        "energy_drift": f64["energy_drift"],
        "charge_drift": f64["charge_drift"],
        "run": line["run"],
        "headline": "the north star's headline (batched FP32 vs plain FP64 on the 3D AMR "
                    "two-population configuration) is not measurable in this build: its "
                    "hierarchy driver has two levels with one finer patch, and the finer level "
                    "runs in FP64 only; mixed_fp32 below is the same comparison on this "
                    "single-patch 3D workload",
    }
    if not args.no_mixed:
        f32 = run_family(wl, "fp32", args.steps, args.warmup, drift_steps)
        m = family_line(wl, args.workload, "fp32", f32)
        m["speedup_vs_fp64"] = f32["value"] / f64["value"]
        m["energy_drift_minus_fp64"] = f32["energy_drift"] - f64["energy_drift"]
        m["run"]["energy_drift_minus_fp64"] = f32["energy_drift_run"] - f64["energy_drift_run"]
        out["mixed_fp32"] = m
    if not args.no_extra and args.workload == "uniform3d":
        # the other interpolation orders of the 3D workload (fewer steps, no drift run), and the
        # 2D Harris workload in both families (its 1000-step drift run included)
        for o in (1, 3):
            w = workload(f"uniform3d-o{o}")
            print(f"[bench] {describe(w, 'fp64')}", file=sys.stderr, flush=True)
            r = run_family(w, "fp64", max(5, args.steps // 2), 2, 0)
            out[f"uniform3d_order{o}"] = family_line(w, f"uniform3d-o{o}", "fp64", r)
        h = workload("harris2d")
        print(f"[bench] {describe(h, 'fp64')} (both families)", file=sys.stderr, flush=True)
        h64 = run_family(h, "fp64", 200, 20, 0 if args.no_drift_run else h.nsteps)
        h32 = run_family(h, "fp32", 200, 20, 0 if args.no_drift_run else h.nsteps)
        hl = family_line(h, "harris2d", "fp64", h64)
        hm = family_line(h, "harris2d", "fp32", h32)
        hm["speedup_vs_fp64"] = h32["value"] / h64["value"]
        hm["run"]["energy_drift_minus_fp64"] = h32["energy_drift_run"] - h64["energy_drift_run"]
        hl["config"] = describe(h, "fp64")
        hl["mixed_fp32"] = hm
        out["harris2d"] = hl
    if not args.no_cpu:
        cw, csteps = cpu_sample(wl)
        print(f"[bench] CPU baseline on {describe(cw, 'fp64')}", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(cw, csteps)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
