"""Benchmark: 64^3 MDE contour steps/s of the SCFT propagator on one MI355X.

Workload (the cubic diblock of the headline metric): 64^3 grid, cubic cell
a = 4.0 Rg = 4.0*sqrt(100/6) b, AB diblock N = 100, fA = 0.35, bA = bB = 1, Ns = 100
(the reference's chain_grid adjusts it to 102, so one density evaluation is 204 MDE
steps: 102 for q and 102 for q+), chi N = 20.  Fields are 0.1*N(0,1) from numpy's
default_rng(seed) plus a BCC {110} cosine bias of amplitude 1.0.  Propagator in FP32,
fields, densities and sums in FP64.

A "step" is one density!(sys) evaluation of every instance in flight (204 MDE steps
each, plus Q and the Simpson densities).  value = MDE steps/s over the whole job.
"""
import argparse
import json
import math
import os
import statistics
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "polyfields.jl_amd"))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12
N = 100
FA = 0.35
CHI = 20.0 / N
GRID = 64
A_CELL = 4.0 * math.sqrt(N / 6.0)


def seeded_fields(shape, seed):
    rng = np.random.default_rng(seed)
    nz, ny, nx = shape
    z, y, x = np.meshgrid(np.arange(nz) / nz, np.arange(ny) / ny, np.arange(nx) / nx,
                          indexing="ij")
    tp = 2 * np.pi
    bias = (np.cos(tp * (x + y)) + np.cos(tp * (x - y)) + np.cos(tp * (y + z)) +
            np.cos(tp * (y - z)) + np.cos(tp * (x + z)) + np.cos(tp * (x - z))) / 6.0
    wa = 0.1 * rng.standard_normal(shape) + 1.0 * bias
    wb = 0.1 * rng.standard_normal(shape) - 1.0 * bias
    return {1: wa, 2: wb}


def build(mod, precision=None, seed=0):
    kw = {} if precision is None else {"precision": precision}
    cell = mod.UnitCell(3, "cubic", A_CELL)
    a, b = mod.Monomer(1), mod.Monomer(2)
    import warnings
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        ch = mod.Diblock(a, b, N, FA, 1.0, 1.0, 100)
    itx = mod.FloryInteraction()
    itx.set_interaction(1, 2, CHI)
    sys_ = mod.FieldSystem((GRID, GRID, GRID), cell, monomers=[a, b], **kw)
    sys_.add_species(ch, 1.0)
    sys_.add_interaction(itx)
    return sys_, ch


def cpu_baseline(seed):
    """The FP64 CPU oracle (numpy/scipy; transforms threaded over the host's cores,
    pointwise work single-threaded as in the reference) on one density evaluation."""
    import oracle as O
    cores = len(os.sched_getaffinity(0))
    O.set_fft_workers(cores)
    so, ch = build(O)
    for m, w in seeded_fields(so.shape, seed).items():
        so.fields[m] = w
    so.density_update()                       # warm-up
    ts = []
    for _ in range(3):
        t0 = time.perf_counter()
        so.density_update()
        ts.append(time.perf_counter() - t0)
    t = statistics.median(ts)
    return {"value": 2 * ch.Ns / t, "unit": "MDE steps/s", "cores": cores, "kind": "port",
            "sample": f"median of 3 FP64 density! evaluations of the 64^3 diblock "
                      f"({2 * ch.Ns} MDE steps each) after 1 warm-up; CPU oracle "
                      f"(scipy.fft workers={cores})"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--precision", default="fp32", choices=["fp32", "fp64"])
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the SCFT backend runs on a single MI355X "
              "(sharding is described in DESIGN.md, not built).", file=sys.stderr)
        sys.exit(2)

    import polyfields_hip as P

    sg, ch = build(P, args.precision, args.seed)
    P.fieldinit(sg, fields=seeded_fields(sg.shape, args.seed))
    nsteps_mde = 2 * ch.Ns
    for _ in range(args.warmup):
        sg.density_update()
    sg.synchronize()

    sg.timing(True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        sg.density_update()
    sg.synchronize()
    dt = time.perf_counter() - t0
    slab_ms, slab_n = sg.timing_read(0)
    pen_ms, pen_n = sg.timing_read(1)
    sg.timing(False)

    ms_per_step = dt * 1e3 / args.steps
    value = nsteps_mde * args.steps / dt
    s = 4 if args.precision == "fp32" else 8
    n = GRID ** 3
    nk = (GRID // 2 + 1) * GRID * GRID
    lanes = 2
    # algorithmic bytes per launch: slab pass reads the half spectrum (2nk reals) and LW (n),
    # writes q (n) and the next half spectrum (2nk); the pencil pass reads and writes 2nk.
    slab_bytes = s * (2 * n + 4 * nk) * lanes
    pen_bytes = s * (4 * nk) * lanes
    slab_avg = slab_ms / max(slab_n, 1) * 1e-3
    pen_avg = pen_ms / max(pen_n, 1) * 1e-3
    if slab_ms >= pen_ms:
        dom = ("xy-slab pass", slab_bytes, slab_avg)
    else:
        dom = ("z-pencil pass", pen_bytes, pen_avg)
    achieved = dom[1] / dom[2] if dom[2] > 0 else 0.0
    b_step = s * (3 * n + 4 * nk)
    line = {
        "metric": "64^3 MDE steps/s (FP32 propagator) + % HBM roofline",
        "value": value,
        "unit": "MDE steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if args.precision == "fp32" else "f64",
        "data": "synthetic (seeded 0.1*N(0,1) + BCC cosine bias)",
        "config": {"workload": "64^3 cubic AB diblock N=100 fA=0.35 chiN=20 Ns=102, "
                               "one density! (204 MDE steps) per step",
                   "instances": 1, "lanes_in_flight": lanes, "precision": args.precision},
        "roofline": {"bound": "hbm", "kernel": dom[0], "achieved": achieved / 1e9,
                     "peak": HBM_PEAK / 1e9, "unit": "GB/s", "frac": achieved / HBM_PEAK,
                     "traffic": None,
                     "bytes_per_launch": dom[1], "avg_launch_us": dom[2] * 1e6},
        "step_roofline_frac": value * b_step / HBM_PEAK,
        "kernel_avg_us": {"slab": slab_avg * 1e6, "pencil": pen_avg * 1e6},
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(args.seed)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
