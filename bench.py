"""Benchmark of the SCFT hot path on one MI355X.

Headline: 64^3 MDE contour steps/s.  Workload: 64^3 grid, cubic cell
a = 4.0 Rg = 4.0*sqrt(100/6) b, AB diblock N = 100, fA = 0.35, bA = bB = 1, Ns = 100
(the reference's chain_grid adjusts it to 102, so one density evaluation is 204 MDE steps:
102 for q and 102 for q+), chi N = 20.  Fields are 0.1*N(0,1) from numpy's
default_rng(seed + instance) plus a BCC {110} cosine bias of amplitude 1.0.  FP32
propagator, FP64 fields and sums.  `--instances` independent systems are in flight (one
batch: every propagator launch covers all of their lanes).  A "step" is one density!
evaluation of every instance; value = MDE steps/s over the whole job.

Secondary (`batched_scft`): SCFT solves/s of the batched sweep workload -- 2-D hexagonal
diblocks with grids drawn from {32^2, 48^2, 64^2, 96^2}, Ns in [64, 128], integer N_A in
[25, 45], chi N in [12, 30], a in [3, 6] Rg, fields 0.1*N(0,1) -- batched in FP32 versus
the plain FP64 path solving the same instances one at a time.  Protocol for both paths:
scft!(AndersonUpdater(nhist=10, iter_euler=10^9, lam_euler=0.05); nsteps=1500, ftol=1e-2)
(pure Euler mixing) then scft!(AndersonUpdater(nhist=20); ftol=1e-6); a solve is an
instance reaching ferr < 1e-6.  (Euler steps of 0.1 oscillate without converging on about
two thirds of these seeds, in the oracle as on the GPU, and Anderson histories of 5 or 10
diverge on most of the rest: the reference's mixing step reaches 1 once the history is
full.)  The batched path runs the Anderson phase in FP32 to ferr < 1e-5, then
switches the propagator to FP64 and finishes to 1e-6 (FP32 residual noise sits near
1e-6).
"""
import argparse
import json
import math
import os
import statistics
import sys
import time
import warnings

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


def build(mod, precision=None, seed=0, chi=CHI):
    kw = {} if precision is None else {"precision": precision}
    cell = mod.UnitCell(3, "cubic", A_CELL)
    a, b = mod.Monomer(1), mod.Monomer(2)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        ch = mod.Diblock(a, b, N, FA, 1.0, 1.0, 100)
    itx = mod.FloryInteraction()
    itx.set_interaction(1, 2, chi)
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
    state = {"Q": ch.Q, "fhelm": so.free_energy(),
             "rho": {m: np.asarray(so.density[m], dtype=float) for m in so.mids}}
    return ({"value": 2 * ch.Ns / t, "unit": "MDE steps/s", "cores": cores, "kind": "port",
             "sample": f"median of 3 FP64 density! evaluations of one 64^3 instance "
                       f"({2 * ch.Ns} MDE steps each) after 1 warm-up; CPU oracle "
                       f"(scipy.fft workers={cores}, pointwise single-threaded)"}, state)


def gpu_vs_oracle(P, seed, state):
    """The GPU's density!, Q and fhelm of the headline's first instance in FP64 and FP32
    against the oracle's evaluation of the same fields (the CPU baseline's run)."""
    out = {}
    for prec in ("fp64", "fp32"):
        s, ch = build(P, prec, seed)
        for m, w in seeded_fields(s.shape, seed).items():
            s.fields[m] = w
        s.density_update()
        q = s.species_scalars(0)[0]
        dr = max(float(np.abs(np.asarray(s.density[m]) - state["rho"][m]).max() /
                       np.abs(state["rho"][m]).max()) for m in s.mids)
        out[prec] = {"Q_rel": abs(q - state["Q"]) / abs(state["Q"]),
                     "rho_max_rel": dr, "fhelm_abs": abs(s.free_energy() - state["fhelm"])}
        del s
    out["workload"] = "instance 0 of the headline batch (64^3, seeded fields), one density!"
    return out


# ---------------------------------------------------------------------------------
def sweep_instances(count, seed=7):
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        n = int(rng.choice([32, 48, 64, 96]))
        Ns = int(rng.integers(64, 129))
        NA = int(rng.integers(25, 46))
        chiN = float(rng.uniform(12, 30))
        a = float(rng.uniform(3, 6)) * math.sqrt(N / 6.0)
        out.append((n, Ns, NA, chiN, a, 1000 + k))
    return out


def sweep_system(P, c, precision=None):
    n, Ns, NA, chiN, a, fs = c
    cell = P.UnitCell(2, "hexagonal", a)
    A, B = P.Monomer(1), P.Monomer(2)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        ch = P.Multiblock([A, B], [NA, N - NA], [1.0, 1.0], Ns)
    itx = P.FloryInteraction()
    itx.set_interaction(1, 2, chiN / N)
    kw = {} if precision is None else {"precision": precision}
    s = P.FieldSystem((n, n, 1), cell, monomers=[A, B], **kw)
    s.add_species(ch, 1.0)
    s.add_interaction(itx)
    rng = np.random.default_rng(fs)
    for m in (1, 2):
        s.fields[m] = 0.1 * rng.standard_normal((1, n, n))
    return s


PH1 = dict(nhist=10, iter_euler=10 ** 9, lam_euler=0.05, nsteps=1500, ftol=1e-2)
PH2 = dict(nhist=20, iter_euler=1, lam_euler=0.01, nsteps=1000, ftol=1e-6)
PH2_FP32 = dict(PH2, ftol=1e-5)


def sweep_step_bytes(n, elem):
    """SURVEY-style algorithmic bytes of one MDE step on an n x n plane: elem (3n + 4nk)."""
    return elem * (3 * n * n + 4 * (n // 2 + 1) * n)


def cpu_sweep_estimate(cfg, work):
    """The FP64 CPU oracle's time for the batch's solved instances, estimated: one
    residuals! evaluation of one instance per grid size is timed (median of 3 after a
    warm-up), scaled per MDE step, and multiplied by the MDE steps each solved instance
    took on the GPU (work: index -> MDE steps)."""
    import oracle as O
    cores = len(os.sched_getaffinity(0))
    O.set_fft_workers(cores)
    per_step = {}
    for i in work:
        n = cfg[i][0]
        if n in per_step:
            continue
        so = sweep_system(O, cfg[i])
        ns = so.species[0].Ns
        so.residuals_update()
        ts = []
        for _ in range(3):
            t0 = time.perf_counter()
            so.residuals_update()
            ts.append(time.perf_counter() - t0)
        per_step[n] = statistics.median(ts) / (2 * ns)
    secs = sum(work[i] * per_step[cfg[i][0]] for i in work)
    return {"solves_per_s": len(work) / secs if secs > 0 else None, "cores": cores,
            "kind": "port", "seconds_estimated": secs, "instances": len(work),
            "sample": "FP64 CPU oracle residuals! (scipy.fft workers=%d, pointwise "
                      "single-threaded) of one instance per grid size %s, per MDE step, "
                      "times the MDE steps each solved instance took in the batched run"
                      % (cores, sorted(per_step))}


def _amplitudes(s):
    """|rho_hat(k)| of the first monomer's density, a translation-invariant fingerprint of a
    converged state."""
    return np.abs(np.fft.rfftn(np.asarray(s.density[s.mids[0]], dtype=float)))


def batched_scft(P, count, plain_sample):
    cfg = sweep_instances(count)
    systems = [sweep_system(P, c, "fp32") for c in cfg]
    b = P.Batch(systems)
    t0 = time.perf_counter()
    r1 = b.scft(**PH1)
    mask = [int(r.converged) for r in r1]
    r15 = b.scft(mask=mask, **PH2_FP32)
    mask = [int(r.converged) for r in r15]
    b.set_precision("fp64")                  # FP64 polish below the FP32 noise floor
    r2 = b.scft(mask=mask, **PH2)
    t_batch = time.perf_counter() - t0
    solved = [i for i, r in enumerate(r2) if r.converged]
    # MDE steps of the batched run: one density! (2 Ns_eff steps) per scft! step; the two
    # first phases in FP32, the polish in FP64
    ns_eff = [sy.species[0].Ns for sy in systems]
    steps32 = [2 * ns_eff[i] * (r1[i].nsteps + r15[i].nsteps) for i in range(count)]
    steps64 = [2 * ns_eff[i] * r2[i].nsteps for i in range(count)]
    agg_bytes = sum(steps32[i] * sweep_step_bytes(cfg[i][0], 4) +
                    steps64[i] * sweep_step_bytes(cfg[i][0], 8) for i in range(count))
    work = {i: steps32[i] + steps64[i] for i in solved}
    fhelm32 = [r.fhelm for r in r2]
    amp32 = {i: _amplitudes(systems[i]) for i in solved}
    del b, systems
    # FP64 reference for every instance: the same protocol batched in FP64 (the batch
    # follows each system's one-at-a-time trajectory to rounding, tested)
    sys64 = [sweep_system(P, c, "fp64") for c in cfg]
    b64 = P.Batch(sys64)
    q1 = b64.scft(**PH1)
    q2 = b64.scft(mask=[int(r.converged) for r in q1], **PH2)
    both = [i for i in solved if q2[i].converged]
    all_errs = [abs(fhelm32[i] - q2[i].fhelm) for i in both]
    # same converged state or another one: the largest change of the translation-
    # invariant density amplitudes |rho_hat(k)| between the two runs
    state = [float(np.max(np.abs(amp32[i] - _amplitudes(sys64[i]))) /
                   np.max(np.abs(amp32[i]))) for i in both]
    same = [e for e, d in zip(all_errs, state) if d < 1e-3]
    n64 = sum(1 for r in q2 if r.converged)
    del b64, sys64
    # plain FP64 path, one instance at a time, same protocol
    t_plain, n_plain_solved, errs = 0.0, 0, []
    for i, c in enumerate(cfg[:plain_sample]):
        s = sweep_system(P, c, "fp64")
        t0 = time.perf_counter()
        ok = False
        try:
            q1 = P.scft(s, P.AndersonUpdater(nhist=10, iter_euler=10 ** 9, lam_euler=0.05),
                        opts=P.SCFTOptions(nsteps=PH1["nsteps"], ftol=PH1["ftol"]))
            if q1.converged:
                q2 = P.scft(s, P.AndersonUpdater(nhist=20),
                            opts=P.SCFTOptions(nsteps=PH2["nsteps"], ftol=PH2["ftol"]))
                ok = q2.converged
        except P.PolyFieldsError:
            ok = False
        t_plain += time.perf_counter() - t0
        if ok:
            n_plain_solved += 1
            if r2[i].converged:
                errs.append(abs(r2[i].fhelm - q2.fhelm))
    batch_rate = len(solved) / t_batch
    plain_rate = n_plain_solved / t_plain if t_plain > 0 else 0.0
    return {
        "instances": count, "batched_fp32_solved": len(solved), "batched_fp32_seconds": t_batch,
        "batched_solves_per_s": batch_rate,
        "plain_fp64_sample": plain_sample, "plain_fp64_solved": n_plain_solved,
        "plain_fp64_seconds": t_plain, "plain_solves_per_s": plain_rate,
        "speedup_vs_plain": batch_rate / plain_rate if plain_rate > 0 else None,
        "fhelm_abs_err_fp32_vs_fp64": {
            "max": max(all_errs) if all_errs else None,
            "median": float(np.median(all_errs)) if all_errs else None, "n": len(all_errs),
            "n_above_1e-6": sum(1 for e in all_errs if e > 1e-6),
            "n_other_state": len(all_errs) - len(same),
            "max_same_state": max(same) if same else None,
            "reference": "the same protocol batched in FP64, every instance both runs solve; "
                         "other state: the density amplitudes |rho_hat(k)| of the two "
                         "converged fields differ by 1e-3 or more of their largest"},
        "fp64_batched_solved": n64,
        "aggregate_step_roofline_frac": agg_bytes / t_batch / HBM_PEAK,
        "mde_steps": {"fp32": int(sum(steps32)), "fp64": int(sum(steps64))},
        "cpu_oracle_estimate": cpu_sweep_estimate(cfg, work),
        "fhelm_abs_err_vs_plain_sample": {"max": max(errs) if errs else None,
                                          "n": len(errs)},
        "protocol": "scft!(Anderson iter_euler=1e9 lam_euler=0.05, ftol 1e-2) then "
                    "scft!(Anderson nhist=20, ftol 1e-6); the batched path runs the second "
                    "phase in FP32 to ftol 1e-5 and polishes in FP64 to 1e-6; solved = "
                    "ferr < 1e-6 in FP64",
    }


# Memory bytes per lane and launch from rocprofv3 --pmc TCC_EA0_RDREQ_DRAM_32B_sum and
# TCC_EA0_WRREQ_WRITE_DRAM_32B_sum (read and write requests from the L2 to the memory
# controller, in 32-byte units; median over the 32-lane dispatches), per xy-slab form: 3 the
# shuffle-exchange kernel (default), 2 the 512-thread LDS kernel.  The same two counters
# read exactly the bytes of a plain 68 MiB device copy (bench.py --workload copy).  Stored
# from the counter runs recorded in profiles/, not measured by this script.
TRAFFIC_PER_LANE = {("xy-slab pass", 3): (2137002 + 2132760) * 32 / 32,
                    ("xy-slab pass", 2): (2138600 + 2129920) * 32 / 32,
                    ("z-pencil pass", None): (1087608 + 1081344) * 32 / 32}


TRAFFIC_SOURCE = ("stored: rocprofv3 --pmc TCC_EA0_RDREQ_DRAM_32B_sum and "
                  "TCC_EA0_WRREQ_WRITE_DRAM_32B_sum runs of this kernel form at this workload "
                  "(profiles/propagator_64cubed_32inst_fp32.md): bytes requested from the "
                  "memory controller, which read a plain copy's bytes exactly; Infinity Cache "
                  "hits are not told apart")


def traffic_per_launch(kernel, args, slab_lanes, pen_lanes, slab_form):
    if args.precision != "fp32" or GRID != 64:
        return None
    if kernel == "z-pencil pass":
        return TRAFFIC_PER_LANE[(kernel, None)] * pen_lanes
    form = 3 if slab_form < 0 else slab_form
    per = TRAFFIC_PER_LANE.get((kernel, form))
    return per * slab_lanes if per is not None else None


def precision_error(P, systems32):
    """FP32 propagator against FP64 on the same fields: the FP32 systems' current
    densities, Q and fhelm against a fresh FP64 evaluation (max over the systems)."""
    rho, rms, q, f = [], [], [], []
    for s32 in systems32:
        s64, ch64 = build(P, "fp64")
        for m in s64.mids:
            s64.fields[m] = np.array(s32.fields[m])
        s64.density_update()
        q32, q64 = s32.species_scalars(0)[0], s64.species_scalars(0)[0]
        q.append(abs(q32 - q64) / abs(q64))
        f.append(abs(s32.free_energy() - s64.free_energy()))
        for m in s64.mids:
            d = np.asarray(s32.density[m]) - np.asarray(s64.density[m])
            rho.append(float(np.abs(d).max()))
            rms.append(float(np.sqrt(np.mean(d * d))))
        del s64
    # np.max keeps a NaN, so a broken state cannot read as agreement
    return {"rho_max_abs": float(np.max(rho)), "rho_rms_abs": float(np.max(rms)),
            "Q_rel": float(np.max(q)), "fhelm_abs": float(np.max(f)),
            "instances": len(systems32)}


def fp64_rate(P, args, nsteps_mde):
    """The headline workload with the FP64 propagator (the plain mode): MDE steps/s of
    the same batch of instances, replayed from its captured graph."""
    systems = []
    for k in range(args.instances):
        s, ch = build(P, "fp64", args.seed + k)
        for m, w in seeded_fields(s.shape, args.seed + k).items():
            s.fields[m] = w
        systems.append(s)
    batch = P.Batch(systems)
    for _ in range(args.warmup):
        batch.density_update()
    # the FP64 schedule for this grid, tuned on this workload like the FP32 one
    schedule = batch.autotune(reps=4) if not args.no_autotune else []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        batch.density_update()
    dt = time.perf_counter() - t0
    value = nsteps_mde * args.instances * args.steps / dt
    n, nk = GRID ** 3, (GRID // 2 + 1) * GRID * GRID
    return {"value": value, "unit": "MDE steps/s", "ms_per_step": dt * 1e3 / args.steps,
            "step_roofline_frac": value * 8 * (3 * n + 4 * nk) / HBM_PEAK,
            "instances": args.instances, "schedule": schedule}


def single_instance(P, args, nsteps_mde):
    s1, _ = build(P, args.precision, args.seed)
    for m, w in seeded_fields(s1.shape, args.seed).items():
        s1.fields[m] = w
    for _ in range(2):
        s1.density_update()
    s1.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        s1.density_update()
    s1.synchronize()
    return nsteps_mde * args.steps / (time.perf_counter() - t0)


# ---------------------------------------------------------------------------------
# Other configurations, timed on request (--workload): the 128^3 blend and the 64^3
# free-energy gradient.  Each prints one JSON line of its own.
def blend_system(P, precision):
    """128^3 orthorhombic cell, ABC triblock (N = 200, 60/80/60, b = 1.0/1.2/0.9, Ns = 200)
    plus A homopolymer (N = 100, phi = 0.15); chi N = 25 / 20 / 35 at N = 200; cell edges
    uniform in [6, 10] Rg from a fixed seed."""
    nref = 200
    rg = math.sqrt(nref / 6.0)
    params = tuple(float(x) * rg for x in np.random.default_rng(5).uniform(6.0, 10.0, 3))
    cell = P.UnitCell(3, "orthorhombic", *params)
    A, B, C = P.Monomer(1), P.Monomer(2), P.Monomer(3)
    tri = P.Multiblock([A, B, C], [60, 80, 60], [1.0, 1.2, 0.9], 200)
    hom = P.Homopolymer(A, 100, 1.0, 100)
    itx = P.FloryInteraction()
    itx.set_interaction(1, 2, 25.0 / nref)
    itx.set_interaction(2, 3, 20.0 / nref)
    itx.set_interaction(1, 3, 35.0 / nref)
    s = P.FieldSystem((128, 128, 128), cell, monomers=[A, B, C], precision=precision)
    s.add_species(tri, 0.85)
    s.add_species(hom, 0.15)
    s.add_interaction(itx)
    rng = np.random.default_rng(11)
    for m in s.mids:
        s.fields[m] = 0.1 * rng.standard_normal(s.shape)
    return s


def time_calls(fn, sync, steps, warmup):
    for _ in range(warmup):
        fn()
    sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    sync()
    return (time.perf_counter() - t0) / steps


def blend_workload(P, args):
    n = 128 ** 3
    nk = 65 * 128 * 128
    per_density = 2 * 200 + 100          # q and q+ of the triblock, q of the homopolymer
    out = {}
    for prec in ("fp32", "fp64"):
        s = blend_system(P, prec)
        dt = time_calls(s.density_update, s.synchronize, args.steps, args.warmup)
        s.timing(True)
        s.density_update()
        slab_ms, slab_n = s.timing_read(0)
        pen_ms, pen_n = s.timing_read(1)
        s.timing(False)
        b = 4 if prec == "fp32" else 8
        out[prec] = {"ms_per_density": dt * 1e3, "mde_steps_per_s": per_density / dt,
                     "step_roofline_frac": per_density / dt * b * (3 * n + 4 * nk) / HBM_PEAK,
                     "kernel_avg_us": {"slab": slab_ms / max(slab_n, 1) * 1e3,
                                       "pencil": pen_ms / max(pen_n, 1) * 1e3}}
        del s
    return {"metric": "128^3 blend MDE steps/s (FP32 propagator)",
            "value": out["fp32"]["mde_steps_per_s"], "unit": "MDE steps/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": out["fp32"]["ms_per_density"], "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f32",
            "data": "synthetic (seeded 0.1*N(0,1))",
            "config": {"workload": "128^3 orthorhombic ABC triblock (N=200, Ns=200) + A "
                                   "homopolymer (N=100, phi=0.15); one density! (500 MDE "
                                   "steps) per step; lanes: triblock q, q+, homopolymer q"},
            "fp32": out["fp32"], "fp64": out["fp64"]}


def grad_workload(P, args):
    """64^3 cubic diblock with fields 0.5 N(0,1): one FP64 evaluation of fhelm and its
    gradient (density! and the reverse sweeps; the gradient copied to the host)."""
    s, ch = build(P, "fp64")
    rng = np.random.default_rng(args.seed)
    for m in s.mids:
        s.fields[m] = 0.5 * rng.standard_normal(s.shape)
    dt_d = time_calls(s.density_update, s.synchronize, args.steps, args.warmup)
    dt_g = time_calls(lambda: P.fhelm_grad(s), s.synchronize, args.steps, args.warmup)
    return {"metric": "64^3 free-energy gradient evaluations/s (FP64)", "value": 1.0 / dt_g,
            "unit": "gradients/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": dt_g * 1e3, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded 0.5*N(0,1))",
            "config": {"workload": "64^3 cubic AB diblock N=100 fA=0.35 chiN=20 Ns=102; "
                                   "fhelm + d/d(w, cell, chi, f) per step"},
            "density_ms": dt_d * 1e3, "gradient_over_density": dt_g / dt_d}


def rqm4_workload(P, args):
    """One system's density! with RQM4 steps (src/system/mde.jl:83-104) against RK2 on the
    same system: 64^3 cubic, 64^2 and 128^2 square diblocks (N=100, fA=0.35, chiN=20,
    Ns=102), fields 0.3 N(0,1), FP64 and FP32."""
    cases = [((64, 64, 64), 3, "cubic", 16.0), ((64, 64, 1), 2, "square", 9.0),
             ((128, 128, 1), 2, "square", 12.0)]
    out = {}
    for prec in ("fp64", "fp32"):
        for dims, dim, crystal, a in cases:
            row = {}
            for mde in ("RK2", "RQM4"):
                cell = P.UnitCell(dim, crystal, a)
                A, B = P.Monomer(1), P.Monomer(2)
                ch = P.Diblock(A, B, 100, FA, 1.0, 1.0, 100)
                itx = P.FloryInteraction()
                itx.set_interaction(1, 2, CHI)
                s = P.FieldSystem(dims, cell, monomers=[A, B], precision=prec, mde=mde)
                s.add_species(ch, 1.0)
                s.add_interaction(itx)
                rng = np.random.default_rng(args.seed)
                for m in s.mids:
                    s.fields[m] = 0.3 * rng.standard_normal(s.shape)
                row[mde] = time_calls(s.density_update, s.synchronize, args.steps,
                                      args.warmup) * 1e3
                del s
            row["rqm4_over_rk2"] = row["RQM4"] / row["RK2"]
            out[f"{prec} {'x'.join(str(d) for d in dims if d > 1)}"] = row
    v = out["fp64 64x64x64"]["RQM4"]
    return {"metric": "64^3 RQM4 density! time (FP64)", "value": v, "unit": "ms",
            "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": v,
            "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
            "data": "synthetic (seeded 0.3*N(0,1))",
            "config": {"workload": "one diblock system (N=100, fA=0.35, chiN=20, Ns=102), "
                                   "RQM4 and RK2 density! per step"},
            "ms_per_density": out}


# config-2 SCFT: MomentumUpdater(:MOM, lam = 0.1) + DomainUpdater(nskip = 100, nper = 10),
# ftol = stol = 1e-5 (the reference's Anderson mixing diverges from this seed)
SCFT64 = dict(lam=0.1, nskip=100, nper=10, ftol=1e-5, stol=1e-5, nsteps=5000)


def scft64_run(P, precision, seed, ftol, stol=SCFT64["stol"]):
    s, ch = build(P, precision)
    for m, w in seeded_fields(s.shape, seed).items():
        s.fields[m] = w
    s.residuals_update()
    s.synchronize()
    t0 = time.perf_counter()
    r = P.scft(s, P.MomentumUpdater(method="MOM", lam=SCFT64["lam"]),
               domain=P.DomainUpdater(nskip=SCFT64["nskip"], nper=SCFT64["nper"]),
               opts=P.SCFTOptions(nsteps=SCFT64["nsteps"], ftol=ftol, stol=stol))
    s.synchronize()
    return s, r, time.perf_counter() - t0


def scft64_workload(P, args):
    """The 64^3 variable-cell SCFT from the BCC-biased seed, to convergence, in the FP64
    plain mode and with the FP32 propagator (stopped at its residual and stress floors,
    ftol 2e-4, stol 3e-5),
    with the FP32 run's and the FP32 propagator's errors against FP64."""
    s64, r64, t64 = scft64_run(P, "fp64", args.seed, SCFT64["ftol"])
    a64, f64 = float(s64.cell.params[0]), s64.free_energy()
    rho64 = {m: np.array(s64.density[m]) for m in s64.mids}
    s64.set_precision("fp32")
    s64.density_update()
    at = {"fhelm_abs": abs(s64.free_energy() - f64),
          "rho_max_abs": max(float(np.abs(np.asarray(s64.density[m]) - rho64[m]).max())
                             for m in s64.mids)}
    del s64
    s32, r32, t32 = scft64_run(P, "fp32", args.seed, 2e-4, 3e-5)
    a32 = float(s32.cell.params[0])
    amp = lambda r: np.abs(np.fft.rfftn(np.asarray(r)))
    err = {"fhelm_abs": abs(r32.fhelm - r64.fhelm), "a_rel": abs(a32 - a64) / a64,
           "rho_max_abs": max(float(np.abs(np.asarray(s32.density[m]) - rho64[m]).max())
                              for m in s32.mids),
           "rho_amp_rel": max(float(np.abs(amp(s32.density[m]) - amp(rho64[m])).max() /
                                    amp(rho64[m]).max()) for m in s32.mids)}
    # FP64 polish of the FP32 run: the same updaters continue with the FP64 propagator to
    # the FP64 run's tolerances
    s32.set_precision("fp64")
    t0 = time.perf_counter()
    rp = P.scft(s32, P.MomentumUpdater(method="MOM", lam=SCFT64["lam"]),
                domain=P.DomainUpdater(nskip=SCFT64["nskip"], nper=SCFT64["nper"]),
                opts=P.SCFTOptions(nsteps=SCFT64["nsteps"], ftol=SCFT64["ftol"],
                                   stol=SCFT64["stol"]))
    s32.synchronize()
    tp = time.perf_counter() - t0
    ap = float(s32.cell.params[0])
    pol = {"seconds": tp, "nsteps": rp.nsteps, "converged": rp.converged,
           "fhelm_abs_err_vs_fp64": abs(rp.fhelm - r64.fhelm), "a_rel_err_vs_fp64":
           abs(ap - a64) / a64,
           "rho_max_abs_err_vs_fp64": max(float(np.abs(np.asarray(s32.density[m]) -
                                                       rho64[m]).max()) for m in s32.mids)}
    res = lambda r, t, a: {"seconds": t, "nsteps": r.nsteps, "converged": r.converged,
                           "ferr": r.ferr, "serr": r.serr, "fhelm": r.fhelm, "a": a}
    return {"metric": "64^3 variable-cell SCFT solves/s (FP64 plain mode)", "value": 1.0 / t64,
            "unit": "solves/s", "n_gpus": 1, "steps": 1, "warmup": 0,
            "ms_per_step": t64 * 1e3, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f64",
            "data": "synthetic (seeded 0.1*N(0,1) + BCC cosine bias)",
            "config": {"workload": "64^3 cubic AB diblock N=100 fA=0.35 chiN=20 Ns=102, "
                                   "a0=4 Rg; scft! with MomentumUpdater(MOM, lam=0.1) + "
                                   "DomainUpdater(nskip=100, nper=10), ftol=stol=1e-5"},
            "fp64": res(r64, t64, a64), "fp32": res(r32, t32, a32),
            "fp32_run_error_vs_fp64": err, "fp32_propagator_at_fp64_state": at,
            "fp32_then_fp64_polish": pol}


def scft64_batch_workload(P, args):
    """A sweep of the 64^3 variable-cell SCFT over chi N: `--instances` systems (default 8
    here) with chi N = 18, 18.5, ... and seeds args.seed + k (BCC-biased fields), each
    scft!(MomentumUpdater(MOM, lam=0.1); DomainUpdater(nskip=100, nper=10)) to
    ftol = stol = 1e-5 in FP64, batched (one launch per contour step covers every
    instance, the cell updates in lockstep) against the first `--plain-sample` of them one
    at a time (default 1 here)."""
    count = args.instances if args.instances != 32 else 8
    sample = args.plain_sample if args.plain_sample != 8 else 1
    chis = [(18.0 + 0.5 * k) / N for k in range(count)]

    def make(k):
        s, ch = build(P, "fp64", args.seed + k, chi=chis[k])
        for m, w in seeded_fields(s.shape, args.seed + k).items():
            s.fields[m] = w
        return s

    systems = [make(k) for k in range(count)]
    b = P.Batch(systems)
    t0 = time.perf_counter()
    res = b.scft(updater=P.MomentumUpdater(method="MOM", lam=SCFT64["lam"]),
                 domain=P.DomainUpdater(nskip=SCFT64["nskip"], nper=SCFT64["nper"]),
                 nsteps=SCFT64["nsteps"], ftol=SCFT64["ftol"], stol=SCFT64["stol"])
    t_batch = time.perf_counter() - t0
    cells = [float(s.cell.params[0]) for s in systems]
    del b, systems
    t_plain, n_plain, cmp = 0.0, 0, []
    for k in range(sample):
        s = make(k)
        t0 = time.perf_counter()
        q = P.scft(s, P.MomentumUpdater(method="MOM", lam=SCFT64["lam"]),
                   domain=P.DomainUpdater(nskip=SCFT64["nskip"], nper=SCFT64["nper"]),
                   opts=P.SCFTOptions(nsteps=SCFT64["nsteps"], ftol=SCFT64["ftol"],
                                      stol=SCFT64["stol"]))
        t_plain += time.perf_counter() - t0
        n_plain += int(q.converged)
        cmp.append({"chiN": chis[k] * N, "nsteps_batch": res[k].nsteps, "nsteps_plain": q.nsteps,
                    "fhelm_abs": abs(res[k].fhelm - q.fhelm),
                    "a_rel": abs(cells[k] - float(s.cell.params[0])) / cells[k]})
        del s
    solved = sum(int(r.converged) for r in res)
    rate = solved / t_batch
    plain = n_plain / t_plain if t_plain > 0 else 0.0
    return {"metric": "batched 64^3 variable-cell SCFT solves/s (chi N sweep, FP64)",
            "value": rate, "unit": "solves/s", "n_gpus": 1, "steps": 1, "warmup": 0,
            "ms_per_step": t_batch * 1e3, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f64",
            "data": "synthetic (seeded 0.1*N(0,1) + BCC cosine bias)",
            "config": {"workload": f"{count} x 64^3 cubic AB diblock N=100 fA=0.35, chi N = "
                                   f"18 .. {chis[-1] * N:g}, a0=4 Rg; scft! with "
                                   "MomentumUpdater(MOM, lam=0.1) + DomainUpdater(nskip=100, "
                                   "nper=10), ftol=stol=1e-5", "instances": count},
            "solved": solved, "seconds": t_batch,
            "per_instance": [{"chiN": chis[k] * N, "nsteps": r.nsteps, "converged": r.converged,
                              "fhelm": r.fhelm, "a": cells[k]} for k, r in enumerate(res)],
            "plain_sample": sample, "plain_solved": n_plain, "plain_seconds": t_plain,
            "plain_solves_per_s": plain, "speedup_vs_plain": rate / plain if plain > 0 else None,
            "batch_vs_plain": cmp}


# variable-cell sweep: the config-3 instances (FP64) with a relaxing hexagonal cell
VC_DOMAIN = dict(nper=5, nskip=50, lam=0.1)
VC_PH2 = dict(nhist=20, iter_euler=1, lam_euler=0.01, nsteps=1500, ftol=1e-6)
VC_STOL = 1e-5


def vcell_sweep(P, count, plain_sample):
    """A phase-diagram sweep with the cell relaxed: `count` of the batched sweep's 2-D
    hexagonal instances, FP64, Euler mixing to ferr < 1e-2 on the fixed cell, then
    AndersonUpdater(nhist=20) with DomainUpdater(nper=5, nskip=50) to ferr < 1e-6 and
    serr < 1e-5; batched (each DomainUpdater round of all due systems in lockstep)
    against the same protocol one instance at a time on `plain_sample` of them."""
    cfg = sweep_instances(count)
    systems = [sweep_system(P, c, "fp64") for c in cfg]
    b = P.Batch(systems)
    t0 = time.perf_counter()
    r1 = b.scft(**PH1)
    r2 = b.scft(mask=[int(r.converged) for r in r1], domain=P.DomainUpdater(**VC_DOMAIN),
                stol=VC_STOL, **VC_PH2)
    t_batch = time.perf_counter() - t0
    solved = [i for i, r in enumerate(r2) if r.converged]
    cells = [float(s.cell.params[0]) for s in systems]
    moved = sum(1 for i in solved if abs(cells[i] - cfg[i][4]) > 1e-3 * cfg[i][4])
    del b, systems
    t_plain, n_plain, errs = 0.0, 0, []
    for i, c in enumerate(cfg[:plain_sample]):
        s = sweep_system(P, c, "fp64")
        t0 = time.perf_counter()
        ok = False
        try:
            q1 = P.scft(s, P.AndersonUpdater(nhist=10, iter_euler=10 ** 9, lam_euler=0.05),
                        opts=P.SCFTOptions(nsteps=PH1["nsteps"], ftol=PH1["ftol"]))
            if q1.converged:
                q2 = P.scft(s, P.AndersonUpdater(nhist=20), domain=P.DomainUpdater(**VC_DOMAIN),
                            opts=P.SCFTOptions(nsteps=VC_PH2["nsteps"], ftol=VC_PH2["ftol"],
                                               stol=VC_STOL))
                ok = q2.converged
        except P.PolyFieldsError:
            ok = False
        t_plain += time.perf_counter() - t0
        if ok:
            n_plain += 1
            if r2[i].converged:
                errs.append({"fhelm_abs": abs(r2[i].fhelm - q2.fhelm),
                             "a_rel": abs(cells[i] - float(s.cell.params[0])) / cells[i],
                             "nsteps_batch": r2[i].nsteps, "nsteps_plain": q2.nsteps})
    rate = len(solved) / t_batch
    plain = n_plain / t_plain if t_plain > 0 else 0.0
    return {"instances": count, "solved": len(solved), "cells_moved": moved,
            "seconds": t_batch, "batched_solves_per_s": rate, "plain_sample": plain_sample,
            "plain_solved": n_plain, "plain_seconds": t_plain, "plain_solves_per_s": plain,
            "speedup_vs_plain": rate / plain if plain > 0 else None, "batch_vs_plain": errs,
            "protocol": "Euler lam=0.05 to ferr<1e-2 (fixed cell), then Anderson nhist=20 + "
                        "DomainUpdater(nper=5, nskip=50) to ferr<1e-6, serr<1e-5; FP64"}


def sweep_vcell_workload(P, args):
    """vcell_sweep as its own JSON line: `--sweep-instances` instances (default 64 here),
    `--plain-sample` of them one at a time."""
    count = args.sweep_instances if args.sweep_instances != 256 else 64
    v = vcell_sweep(P, count, args.plain_sample)
    return {"metric": "batched variable-cell SCFT solves/s (2-D hexagonal sweep, FP64)",
            "value": v["batched_solves_per_s"], "unit": "solves/s", "n_gpus": 1, "steps": 1,
            "warmup": 0, "ms_per_step": v["seconds"] * 1e3, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f64",
            "data": "synthetic (seeded 0.1*N(0,1))",
            "config": {"workload": f"{count} hexagonal diblock instances of the batched "
                                   "sweep (grids 32^2-96^2), " + v["protocol"],
                       "instances": count},
            **{k: x for k, x in v.items() if k != "protocol"}}


def copy_workload(P, args):
    """Counter calibration: torch's device-to-device copy of a 68 MiB float32 buffer (the
    size of one 32-lane xy-slab launch's working set), so that rocprofv3 --pmc readings of
    this run give each counter's bytes per known byte read and written."""
    import torch
    nbytes = 68 * 2 ** 20
    x = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda:0")
    y = torch.empty_like(x)
    for _ in range(args.warmup):
        y.copy_(x)
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(args.steps):
        y.copy_(x)
    b.record()
    torch.cuda.synchronize()
    ms = a.elapsed_time(b) / args.steps
    return {"metric": "device copy bandwidth (read + write)", "value": 2 * nbytes / ms / 1e6,
            "unit": "GB/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f32", "data": "synthetic (ones)",
            "config": {"workload": f"copy of {nbytes} bytes per step"}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--instances", type=int, default=32)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--precision", default="fp32", choices=["fp32", "fp64"])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-sweep", action="store_true")
    ap.add_argument("--no-single", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--sweep-instances", type=int, default=256)
    ap.add_argument("--plain-sample", type=int, default=8)
    ap.add_argument("--schedule", default=None, metavar="SLAB,PENCIL",
                    help="density64: fix the 64^3 slab and pencil kernel forms (the numbers "
                         "of polyfields_hip.schedule_set) instead of autotuning")
    ap.add_argument("--workload", default="density64",
                    choices=["density64", "blend128", "grad64", "scft64", "scft64_batch",
                             "sweep_vcell", "rqm4", "copy"],
                    help="density64: the headline (64^3 batch); blend128 / grad64 / scft64: "
                         "the large blend, the free-energy gradient and the 64^3 "
                         "variable-cell SCFT, one JSON line each; rqm4: one system's "
                         "density! with RQM4 against RK2; copy: a plain device copy of "
                         "known size (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the SCFT backend runs on a single MI355X "
              "(sharding is described in DESIGN.md, not built).", file=sys.stderr)
        sys.exit(2)

    import polyfields_hip as P

    if args.workload != "density64":
        fn = {"blend128": blend_workload, "grad64": grad_workload,
              "scft64": scft64_workload, "scft64_batch": scft64_batch_workload,
              "sweep_vcell": sweep_vcell_workload, "rqm4": rqm4_workload,
              "copy": copy_workload}[args.workload]
        print(json.dumps(fn(P, args)))
        return

    fixed_schedule = []
    if args.schedule:
        slab, pencil = (int(x) for x in args.schedule.split(","))
        P.schedule_set(args.precision, (GRID, GRID, GRID), slab=slab, pencil=pencil)
        fixed_schedule = [{"nx": GRID, "ny": GRID, "nz": GRID, "slab": slab, "pencil": pencil}]
        args.no_autotune = True
    systems = []
    for k in range(args.instances):
        s, ch = build(P, args.precision, args.seed + k)
        for m, w in seeded_fields(s.shape, args.seed + k).items():
            s.fields[m] = w
        systems.append(s)
    batch = P.Batch(systems)
    nsteps_mde = 2 * ch.Ns
    for _ in range(args.warmup):
        batch.density_update()
    # kernel schedule for this grid, picked by the autotuner on this workload
    schedule = batch.autotune(reps=4) if not args.no_autotune else fixed_schedule

    # timed region: the captured launch sequence replayed once per step
    t0 = time.perf_counter()
    for _ in range(args.steps):
        batch.density_update()
    dt = time.perf_counter() - t0
    # the chain of propagate! calls alone (the sweeps of density!, without Q, Simpson and
    # pooling), timed the same way: SURVEY's isolated step chain
    batch.sweeps()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        batch.sweeps()
    dt_chain = time.perf_counter() - t0
    # kernel-level figures from a separate pass: HIP events on the launching streams
    # bracket one launch in 8, run 4 times back to back (the events' overhead spread over
    # the repeats); the launches run eagerly and one stream at a time, so each timed
    # launch has the device to itself.  The pass's densities are not used.
    batch.timing(True)
    for _ in range(max(2, args.steps // 2)):
        batch.density_update()
    slab_ms, slab_n = batch.timing_read(0)
    pen_ms, pen_n = batch.timing_read(1)
    slab_lanes, pen_lanes = batch.timing_lanes(0), batch.timing_lanes(1)
    batch.timing(False)

    ms_per_step = dt * 1e3 / args.steps
    value = nsteps_mde * args.instances * args.steps / dt
    s = 4 if args.precision == "fp32" else 8
    n = GRID ** 3
    nk = (GRID // 2 + 1) * GRID * GRID
    lanes = 2 * args.instances
    # algorithmic bytes per launch: per lane, the slab pass reads the half spectrum (2nk
    # reals) and LW (n), writes q_s (n) and the next half spectrum (2nk); the pencil pass
    # reads and writes the half spectrum (4nk).  A launch covers slab_lanes / pen_lanes
    # lanes (the batch runs large 3-D groups as two halves on two streams).
    slab_bytes = s * (2 * n + 4 * nk) * slab_lanes
    pen_bytes = s * (4 * nk) * pen_lanes
    slab_avg = slab_ms / max(slab_n, 1) * 1e-3
    pen_avg = pen_ms / max(pen_n, 1) * 1e-3
    dom = ("xy-slab pass", slab_bytes, slab_avg) if slab_ms >= pen_ms else \
        ("z-pencil pass", pen_bytes, pen_avg)
    achieved = dom[1] / dom[2] if dom[2] > 0 else 0.0
    b_step = s * (3 * n + 4 * nk)
    slab_form = next((c["slab"] for c in schedule if (c["nx"], c["ny"], c["nz"]) ==
                      (GRID, GRID, GRID)), -1)
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
        "config": {"workload": "64^3 cubic AB diblock N=100 fA=0.35 chiN=20 Ns=102; one "
                               "density! (204 MDE steps) of every instance per step",
                   "instances": args.instances, "lanes_in_flight": lanes,
                   "lanes_per_launch": {"slab": slab_lanes, "pencil": pen_lanes},
                   "schedule": schedule,
                   "precision": args.precision},
        "roofline": {"bound": "hbm", "kernel": dom[0], "achieved": achieved / 1e9,
                     "peak": HBM_PEAK / 1e9, "unit": "GB/s", "frac": achieved / HBM_PEAK,
                     "traffic": traffic_per_launch(dom[0], args, slab_lanes, pen_lanes,
                                                   slab_form),
                     "traffic_source": TRAFFIC_SOURCE,
                     "bytes_per_launch": dom[1], "avg_launch_us": dom[2] * 1e6},
        "step_roofline_frac": value * b_step / HBM_PEAK,
        "kernel_avg_us": {"slab": slab_avg * 1e6, "pencil": pen_avg * 1e6},
        # the propagator sweeps alone: MDE steps/s and their fraction of HBM peak on the
        # same B_step (density! adds Q, the Simpson pass and pooling to this)
        "step_chain": {"value": nsteps_mde * args.instances * args.steps / dt_chain,
                       "ms_per_step": dt_chain * 1e3 / args.steps,
                       "step_roofline_frac": nsteps_mde * args.instances * args.steps /
                       dt_chain * b_step / HBM_PEAK,
                       "what": "pf_batch_sweeps: the q and q+ sweeps of every instance "
                               "(exponential tables, xy-slab and z-pencil launches), "
                               "without Q, Simpson and pooling"},
    }
    if args.precision == "fp32":
        # the timing pass repeats launches, so the batch's densities are recomputed first
        batch.density_update()
        line["fp32_error_vs_fp64"] = precision_error(P, systems[:4])
    del batch, systems
    if args.precision == "fp32" and not args.no_fp64:
        line["fp64"] = fp64_rate(P, args, nsteps_mde)
    # the same workload with one instance in flight (two lanes per launch)
    if not args.no_single:
        line["single_instance_steps_per_s"] = single_instance(P, args, nsteps_mde)
    if not args.no_sweep:
        line["batched_scft"] = batched_scft(P, args.sweep_instances, args.plain_sample)
        # the same kind of sweep with the cell relaxed (64 instances, FP64, 4 one at a time)
        line["batched_vcell_scft"] = vcell_sweep(P, 64, 4)
    if not args.no_cpu:
        line["cpu_baseline"], state = cpu_baseline(args.seed)
        line["gpu_vs_oracle"] = gpu_vs_oracle(P, args.seed, state)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
