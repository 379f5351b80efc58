"""Throughput of the n-body hot path on one MI355X.

Default workload: Barnes-Hut on a Plummer sphere of N = 4,194,304 equal
masses (scale radius 1, G = 1 in model units), dt = 1e-3, FP32 pair forces
with FP64 accumulation, FP64 octree sums and integrator, at the largest
opening angle theta whose median relative force error against the FP64 direct
sum is at most 1e-4 (chosen on the initial state from a fixed candidate list,
then checked again after the timed steps).  A step is one iteration of the
reference's loop (BarnesHutAlgorithm.cpp:149-277): kick-drift, octree build,
forces, kick.  The line also carries the theta sweep {0.3, 0.5, 0.7, 1.0} of
the same input.

Every line states whether its error meets the 1e-4 bound (config.gate_met);
a line that misses it reports value null and the measured rate as
config.ungated_value.

Prints one JSON line on stdout; progress goes to stderr.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "n-body-simulation_amd"))

import nbody as nb  # noqa: E402

METRIC = "body-steps/s at median rel. force err ≤1e-4 vs FP64 direct sum"
FP32_PEAK_TFLOPS = 157.3
FP64_PEAK_TFLOPS = 78.6
BF16_PEAK_TFLOPS = 2500.0  # dense matrix-core peak
HBM_PEAK_GBS = 8000.0
FLOP_PER_INTERACTION = 20  # GPU Gems 3 convention, reference README.md:158
QUAD_FLOP_PER_PAIR = 49  # monopole + traceless quadrupole of a cell, same counting
MLP_FLOP_PER_PAIR = 18432  # 2 * (13*64 + 64*64 + 64*64 + 64*3) per network evaluation
ERR_SAMPLE = 16384
ERR_GATE = 1e-4
# opening angles tried for the default workload, largest first; the first whose
# median error on the initial state is at most ERR_GATE * THETA_MARGIN is
# refined by bisection towards the failing one before it
THETA_CANDIDATES = (1.0, 0.9, 0.8, 0.75, 0.7, 0.65, 0.6, 0.55, 0.5, 0.45, 0.42, 0.4, 0.38, 0.36,
                    0.34, 0.32, 0.3, 0.25, 0.2)
THETA_MARGIN = 0.97
THETA_BISECTIONS = 4  # refine between the first passing and the last failing candidate
THETA_SWEEP = (0.3, 0.5, 0.7, 1.0)

# Learned correction: training set (Plummer snapshots, accepted (cell, body)
# pairs of a per-body walk at theta 1.0, targets from the exact cell sums) and
# the plain Barnes-Hut opening angle it is compared with.
MLP_TRAIN = dict(snapshots=4, n=262144, targets=1800, pairs=2_000_000, theta=0.5, epochs=40,
                 batch=4096)
# corrected traversals tried: opening angles above the plain one's, and the
# gate tau (cells with e / d_group > tau get the network) as fractions of theta
MLP_THETA_STEPS = (0.0, 0.02, 0.04, 0.07, 0.1, 0.15)
MLP_TAU_FRACS = (0.0, 0.7, 0.85)

# Ensemble: 512 independent systems, sizes floor(2^U[10, 16]), alternating
# star + disk (dt = 1 h) and Plummer (dt = 1e-3); direct sum below the
# autotuned threshold, Barnes-Hut theta = 0.5 above it.
ENSEMBLE = dict(nsys=512, seed=0x4E424F44590004, lo_log2=10, hi_log2=16, theta=0.5,
                err_systems=32, err_targets=64)
# seconds of CPU loop time the reference CPU path is timed for
CPU_SAMPLE_S = 10.0

# HBM bytes per launch of the force kernel, keyed by (workload, theta), from
# rocprofv3 FETCH_SIZE and WRITE_SIZE runs of this script scaled by a device
# copy of known size (FETCH x 2.0, WRITE x 1.0):
#  - 4M Plummer, theta 0.37375: 5,775,771 KiB x 2 read + 528,162 KiB written
#    (profiles/bh_plummer_4m_hbm_counters.md);
#  - 1M disk, theta 0.5: 932,786 KiB x 2 read + 133,202 KiB written
#    (profiles/bh_disk_1m_hbm_counters.md);
#  - 1M disk, FP64 per-body traversal (k_bh_exact): 1,106,567 KiB x 2 read +
#    172,409 KiB written (profiles/bh_disk_1m_exact_kernel_stats.md).
TRAFFIC_BYTES = {
    ("bh_plummer_4m", 0.37375): 12.370e9,
    ("bh_disk_1m", 0.5): 2.047e9,
    ("bh_disk_1m_exact", 0.5): 2.443e9,
}

WORKLOADS = {
    # name: (generator, n, seed, algorithm, theta, dt, precision)
    "bh_disk_1m": ("disk", 1 << 20, 0x4E424F44590001, "BarnesHut", 0.5, 1.0 / 24, nb.PREC_FP32C),
    # the same in the reference's FP64 arithmetic and per-body traversal order
    # (bit-identical to the oracle; the application's default precision)
    "bh_disk_1m_exact": ("disk", 1 << 20, 0x4E424F44590001, "BarnesHut", 0.5, 1.0 / 24,
                         nb.PREC_FP64),
    # FP64 pairs, split sums in a fixed order (within 1e-12 of the reference sum)
    "naive_disk_16k": ("disk", 16384, 0x4E424F44590000, "naive", 0.5, 1.0 / 24,
                       nb.PREC_FP64_FAST),
    # the reference's arithmetic and summation order, bit-identical to the oracle
    "naive_disk_16k_exact": ("disk", 16384, 0x4E424F44590000, "naive", 0.5, 1.0 / 24,
                             nb.PREC_FP64),
    "naive_disk_16k_fp32": ("disk", 16384, 0x4E424F44590000, "naive", 0.5, 1.0 / 24, nb.PREC_FP32C),
    # theta None: the largest candidate meeting the error bound (choose_theta)
    "bh_plummer_4m": ("plummer", 1 << 22, 0x4E424F44590002, "BarnesHut", None, 1e-3, nb.PREC_FP32C),
    # the learned far-field correction against plain Barnes-Hut, both at the
    # error bound (run_learned)
    "bh_plummer_4m_mlp": ("plummer", 1 << 22, 0x4E424F44590003, "BarnesHut", None, 1e-3,
                          nb.PREC_FP32C),
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def gen(kind, n, seed):
    return nb.gen_star_disk(n, seed) if kind == "disk" else nb.gen_plummer(n, seed)


def median_rel_err(ctx, n, rng):
    """Median relative error of the current accelerations against the FP64
    direct sum (reference arithmetic, on the GPU) over sampled targets."""
    ax, ay, az = ctx.accel()
    t = np.arange(n, dtype=np.uint32) if n <= 65536 else \
        np.sort(rng.choice(n, ERR_SAMPLE, replace=False)).astype(np.uint32)
    rx, ry, rz = ctx.forces_direct_targets(t)
    e = np.sqrt((ax[t] - rx) ** 2 + (ay[t] - ry) ** 2 + (az[t] - rz) ** 2) / \
        np.sqrt(rx ** 2 + ry ** 2 + rz ** 2)
    return float(np.median(e)), int(t.size)


def cpu_baseline(d, algorithm, theta, dt):
    """The reference's CPU path (oracle/ restatement, OpenMP on all host cores),
    a few loop iterations of the same workload (about 10 s of CPU time)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc
    n = d["m"].size
    threads = orc.get_threads()
    if algorithm == "BarnesHut":
        iters = max(1, (4 << 20) // n)  # about 10 s at 1 M bodies, one iteration at 4 M
    else:
        # a 2-iteration probe sizes the run to about CPU_SAMPLE_S seconds
        probe = orc.simulate(algorithm, d["x"], d["y"], d["z"], d["vx"], d["vy"], d["vz"], d["m"],
                             dt, 2 * dt)
        per_iter = probe["loop_ms"] / max(1, probe["steps"])
        iters = int(min(1000, max(4, CPU_SAMPLE_S * 1e3 / max(per_iter, 1e-3))))
    log(f"cpu baseline: {algorithm} n={n}, {iters} loop iterations on {threads} threads ...")
    r = orc.simulate(algorithm, d["x"], d["y"], d["z"], d["vx"], d["vy"], d["vz"], d["m"], dt,
                     iters * dt, theta=theta)
    secs = r["loop_ms"] / 1000.0
    return {"value": n * r["steps"] / secs, "unit": "body-steps/s", "cores": threads,
            "cpu": cpu_model(), "kind": "port",
            "sample": f"{r['steps']} loop iteration(s) (kick-drift, forces"
                      f"{', octree build' if algorithm == 'BarnesHut' else ''}, kick) of the same "
                      f"n={n} input, FP64, t=0 force untimed; {secs:.1f} s"}


def copy_calibration(mib=256, reps=3):
    """A plain device-to-device copy of known size (hipMemcpy's blit kernel),
    run under rocprofv3 --pmc to scale FETCH_SIZE / WRITE_SIZE to bytes."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    nbytes = mib << 20
    src, dst = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(src), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(ctypes.byref(dst), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMemset(src, 1, ctypes.c_size_t(nbytes)) == 0
    for _ in range(reps):
        assert hip.hipMemcpy(dst, src, ctypes.c_size_t(nbytes), 3) == 0  # device to device
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(src)
    hip.hipFree(dst)
    print(json.dumps({"copy_bytes": nbytes, "reps": reps}), flush=True)


def snapshot_cost(reps=40, rounds=3):
    """Cost of a visualised step of the default workload (every step stored),
    three ways, alternated in one process: nb_steps alone; nb_steps then
    synchronous nb_get_state + nb_get_accel into reused buffers and a copy
    into a time series (the reference's loop); nb_steps with the asynchronous
    snapshot begun after each step and stored on a second thread during the
    next one (the application's loop)."""
    import ctypes
    import threading
    kind, n, seed, algorithm, theta, dt, prec = WORKLOADS["bh_disk_1m"]
    d = gen(kind, n, seed)
    ctx = nb.Context(n, theta=theta, precision=prec)
    ctx.set_state(d["x"], d["y"], d["z"], d["vx"], d["vy"], d["vz"], d["m"])
    ctx.bh_build()
    ctx.bh_forces()
    lib, h = ctx.lib, ctx.h
    P = ctypes.POINTER(ctypes.c_double)
    buf = [np.empty(n) for _ in range(9)]
    series = [np.empty(n) for _ in range(9)]

    def step():
        rc = lib.nb_steps(h, nb.ALGO_BARNES_HUT, dt, 1)
        if rc:
            raise nb.NBodyError(rc, "nb_steps")

    def plain():
        for _ in range(reps):
            step()

    def sync_reads():
        for _ in range(reps):
            step()
            assert lib.nb_get_state(h, *[b.ctypes.data_as(P) for b in buf[:6]]) == 0
            assert lib.nb_get_accel(h, *[b.ctypes.data_as(P) for b in buf[6:]]) == 0
            for dst, src in zip(series, buf):
                np.copyto(dst, src)

    def async_snapshots():
        worker = None
        ptrs = (P * 9)()

        def store(views):
            for dst, src in zip(series, views):
                np.copyto(dst, src)
        for r in range(reps):
            step()
            if r:
                assert lib.nb_snapshot_wait(h, ptrs) == 0
                views = [np.ctypeslib.as_array(ptrs[i], shape=(n,)) for i in range(9)]
                if worker:
                    worker.join()
                worker = threading.Thread(target=store, args=(views,))
                worker.start()
            assert lib.nb_snapshot_begin(h) == 0
        assert lib.nb_snapshot_wait(h, ptrs) == 0
        if worker:
            worker.join()

    ways = {"steps_only": plain, "sync_reads": sync_reads, "async_snapshot": async_snapshots}
    for f in ways.values():  # warm-up, also allocates the snapshot slots
        f()
    ms = {k: [] for k in ways}
    for _ in range(rounds):
        for k, f in ways.items():
            t0 = time.perf_counter()
            f()
            ms[k].append((time.perf_counter() - t0) / reps * 1e3)
    ctx.close()
    print(json.dumps({"snapshot_cost": "ms per step, every step stored", "bodies": n,
                      "workload": "bh_disk_1m", "reps": reps, "rounds": rounds,
                      **{k: round(float(np.median(v)), 3) for k, v in ms.items()}}), flush=True)


def cpu_model():
    """The host CPU's model name (first /proc/cpuinfo entry)."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def gated(value, err):
    """(value, gate_met): a rate counts only at median error <= ERR_GATE."""
    met = err is not None and err <= ERR_GATE
    return (round(value, 1) if met else None), met


def err_targets(n):
    """The sampled error targets: every body up to 65,536, else a seeded
    sample of ERR_SAMPLE."""
    if n <= 65536:
        return np.arange(n, dtype=np.uint32)
    return np.sort(np.random.default_rng(1).choice(n, ERR_SAMPLE, replace=False)).astype(np.uint32)


def choose_theta(ctx, d, dt):
    """The largest opening angle whose median error on the initial state is at
    most ERR_GATE * THETA_MARGIN: the first passing candidate of
    THETA_CANDIDATES (largest first), refined by THETA_BISECTIONS bisections
    towards the failing candidate before it (one build + traversal each,
    against one FP64 direct sum over the sampled targets)."""
    n = d["m"].size
    ctx.set_state(d["x"], d["y"], d["z"], d["vx"], d["vy"], d["vz"], d["m"])
    t = err_targets(n)
    rx, ry, rz = ctx.forces_direct_targets(t)
    den = np.sqrt(rx ** 2 + ry ** 2 + rz ** 2)
    trials = []

    def passes(th):
        ctx.set_params(theta=th)
        ctx.bh_build()
        ax, ay, az = ctx.bh_forces()
        e = np.sqrt((ax[t] - rx) ** 2 + (ay[t] - ry) ** 2 + (az[t] - rz) ** 2) / den
        err = float(np.median(e))
        trials.append({"theta": th, "median_rel_err": err})
        log(f"theta {th}: initial-state median rel err {err:.3e}")
        return err <= ERR_GATE * THETA_MARGIN

    failed = None
    for th in THETA_CANDIDATES:
        if passes(th):
            lo, hi = th, failed
            for _ in range(THETA_BISECTIONS if hi is not None else 0):
                mid = round(0.5 * (lo + hi), 5)
                if passes(mid):
                    lo = mid
                else:
                    hi = mid
            return lo, trials
        failed = th
    raise RuntimeError(f"no candidate opening angle meets {ERR_GATE}: {trials}")


def energy_cost(n=1 << 20, reps=3):
    """Time nb_energy (the --energy diagnostics, nBodyAlgorithm.cpp:11-86) on
    the 1M-body star + disk: FP64 pair terms in the reference's order, N^2/2
    pairs, then the host sums.  Reports wall ms per call; the kernel's own
    time comes from rocprofv3 --kernel-trace of this mode."""
    d = gen("disk", n, WORKLOADS["bh_disk_1m"][2])
    with nb.Context(n) as ctx:
        ctx.set_state(d["x"], d["y"], d["z"], d["vx"], d["vy"], d["vz"], d["m"])
        ctx.energy()  # warm-up
        t0 = time.perf_counter()
        for _ in range(reps):
            e = ctx.energy()
        ms = (time.perf_counter() - t0) / reps * 1e3
    pairs = n * (n - 1) / 2
    print(json.dumps({"energy_cost": "ms per nb_energy call (kernel, copies, host sums)",
                      "bodies": n, "pairs": pairs, "ms": round(ms, 3), "reps": reps,
                      "virial": e["virial"]}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="bh_plummer_4m",
                    choices=sorted(WORKLOADS) + ["ensemble_512"])
    ap.add_argument("--theta", type=float, default=None,
                    help="opening angle (default: the workload's; for bh_plummer_4m the largest "
                         "candidate meeting the error bound)")
    ap.add_argument("--group-max", type=int, default=0,
                    help="bodies per traversal group (0: the library's default)")
    ap.add_argument("--pack-max", type=int, default=0,
                    help="sibling groups are packed inside cells of at most this many bodies "
                         "(0: the library's default)")
    ap.add_argument("--quadrupole", action="store_true",
                    help="Barnes-Hut cells with their quadrupole moments (FP32C traversal)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-sweep", action="store_true",
                    help="skip the theta sweep of the Barnes-Hut workloads")
    ap.add_argument("--theta-sweep", default=",".join(str(t) for t in THETA_SWEEP),
                    help="comma-separated opening angles: also time the Barnes-Hut workload at "
                         "each (same input), reported under config.theta_sweep")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 256 MiB device copy (HBM counter calibration)")
    ap.add_argument("--snapshot-cost", action="store_true",
                    help="only time a visualised step of the 1M-body disk workload: synchronous "
                         "reads against the asynchronous snapshot")
    ap.add_argument("--mlp-train-theta", type=float, default=MLP_TRAIN["theta"],
                    help="opening angle of the per-body walks that collect the training pairs")
    ap.add_argument("--energy-cost", action="store_true",
                    help="only time the energy diagnostics on the 1M-body disk")
    args = ap.parse_args()
    if args.energy_cost:
        energy_cost()
        return
    if args.snapshot_cost:
        snapshot_cost()
        return
    if args.copy_calibration:
        copy_calibration()
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for "
              f"--gpus {args.gpus}).", file=sys.stderr)
        sys.exit(2)
    if args.workload == "ensemble_512":
        run_ensemble(args)
        return
    kind, n, seed, algorithm, theta, dt, prec = WORKLOADS[args.workload]
    if args.theta is not None:
        theta = args.theta
    if args.workload.endswith("_mlp"):
        run_learned(args)
        return
    log(f"device: {nb.device_name()}")
    log(f"workload {args.workload}: generating {kind} n={n} ...")
    d = gen(kind, n, seed)
    ctx = nb.Context(n, theta=theta if theta is not None else 0.5, precision=prec,
                     quadrupole=int(args.quadrupole), group_max=args.group_max,
                     pack_max=args.pack_max)
    theta_search = None
    if algorithm == "BarnesHut" and theta is None:
        theta, theta_search = choose_theta(ctx, d, dt)
        ctx.set_params(theta=theta)
        log(f"opening angle: {theta}")
    secs, phases, err, sample, inter = timed_run(ctx, d, algorithm, dt, args.steps, args.warmup)
    raw_value = n * args.steps / secs
    value, gate_met = gated(raw_value, err)

    # roofline of the dominant kernel: the force kernel
    accel_ms = phases["Acceleration Kernel Time"] / args.steps
    if algorithm == "BarnesHut":
        # quadrupole cell pairs count 49 flop (the direct sum's counting of
        # their arithmetic), monopole cells and bodies 20
        qpairs = float(ctx.quadrupole_pairs()) if args.quadrupole else 0.0
        flops = FLOP_PER_INTERACTION * (float(inter.sum()) - qpairs) + QUAD_FLOP_PER_PAIR * qpairs
        if prec == nb.PREC_FP32C:
            peak, bound, kernel = FP32_PEAK_TFLOPS, "valu", "k_bh_group"
        else:  # the FP64 per-body traversal in the reference's order
            peak, bound, kernel = FP64_PEAK_TFLOPS, "valu", "k_bh_exact"
    else:
        flops = FLOP_PER_INTERACTION * float(n) * float(n)
        peak = FP32_PEAK_TFLOPS if prec == nb.PREC_FP32C else FP64_PEAK_TFLOPS
        kernel = {nb.PREC_FP32C: "k_direct_f32c", nb.PREC_FP64: "k_direct_f64",
                  nb.PREC_FP64_FAST: "k_direct_f64f"}[prec]
        bound = "valu"
    achieved = flops / (accel_ms * 1e-3) / 1e12
    # compute-bound: "valu" is the vector ALU (the pair maths), "mfma" the
    # matrix cores (the network); the peak is that unit's datasheet figure
    roof = {"bound": bound, "achieved": round(achieved, 3), "peak": peak, "unit": "TFLOP/s",
            "frac": round(achieved / peak, 4), "traffic": TRAFFIC_BYTES.get((args.workload, theta)),
            "kernel": kernel, "kernel_ms": round(accel_ms, 4), "flop_per_launch": flops}
    out = {
        "metric": METRIC, "value": value, "unit": "body-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(1e3 * secs / args.steps, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32+f64" if prec == nb.PREC_FP32C else "f64",
        "data": "synthetic",
        "config": {"workload": args.workload, "bodies": n, "distribution": kind,
                   "algorithm": algorithm, "theta": theta if algorithm == "BarnesHut" else None,
                   "cells": "quadrupole" if args.quadrupole else "monopole",
                   "group_max": ctx.params.group_max or 64, "pack_max": ctx.params.pack_max or 256,
                   "dt": dt, "median_rel_err": err, "err_sample": sample, "gate": ERR_GATE,
                   "gate_met": gate_met, "ungated_value": round(raw_value, 1),
                   "phase_ms_per_step": {k: round(v / args.steps, 4) for k, v in phases.items()},
                   "mean_interactions": None if inter is None else float(inter.mean())},
        "roofline": roof,
    }
    if theta_search is not None:
        out["config"]["theta_search"] = theta_search
    if algorithm == "BarnesHut" and not args.no_sweep and args.theta_sweep:
        sweep = []
        for th in (float(v) for v in args.theta_sweep.split(",")):
            ctx.set_params(theta=th)
            ssecs, sph, serr, _, sinter = timed_run(ctx, d, algorithm, dt, args.steps, args.warmup)
            sms = sph["Acceleration Kernel Time"] / args.steps
            sq = float(ctx.quadrupole_pairs()) if args.quadrupole else 0.0
            sflop = FLOP_PER_INTERACTION * (float(sinter.sum()) - sq) + QUAD_FLOP_PER_PAIR * sq
            svalue = n * args.steps / ssecs
            sweep.append({"theta": th, "value": gated(svalue, serr)[0],
                          "ungated_value": round(svalue, 1), "median_rel_err": serr,
                          "gate_met": gated(svalue, serr)[1],
                          "mean_interactions": float(sinter.mean()), "force_ms": round(sms, 4),
                          "frac": round(sflop / (sms * 1e-3) / 1e12 / peak, 4)})
            log(f"theta {th}: {svalue:.4g} body-steps/s, err {serr:.3e}")
        out["config"]["theta_sweep"] = sweep
    ctx.close()
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(d, algorithm, theta, dt)
    print(json.dumps(out), flush=True)


def run_learned(args):
    """Barnes-Hut with the learned far-field correction against plain
    Barnes-Hut on the 4M Plummer sphere, both at median error <= ERR_GATE.
    Plain: the largest gated theta (choose_theta).  Corrected: every (theta,
    tau) of theta_plain + MLP_THETA_STEPS and tau = theta * MLP_TAU_FRACS is
    evaluated once on the initial state (error against the same FP64 targets,
    and the build + traversal time of that evaluation); the fastest that
    meets ERR_GATE * THETA_MARGIN is then timed like the plain run."""
    kind, n, seed, algorithm, _, dt, prec = WORKLOADS[args.workload]
    log(f"device: {nb.device_name()}")
    net, train_info = train_correction(args.mlp_train_theta)
    log(f"workload {args.workload}: generating {kind} n={n} ...")
    d = gen(kind, n, seed)
    ctx = nb.Context(n, theta=0.5, precision=prec)
    theta_p, search_p = choose_theta(ctx, d, dt)
    ctx.set_params(theta=theta_p)
    psecs, pph, perr, _, pinter = timed_run(ctx, d, algorithm, dt, args.steps, args.warmup)
    pvalue = n * args.steps / psecs
    log(f"plain theta {theta_p}: {pvalue:.4g} body-steps/s, err {perr:.3e}")

    # the corrected search on the initial state
    ctx.set_state(d["x"], d["y"], d["z"], d["vx"], d["vy"], d["vz"], d["m"])
    t = err_targets(n)
    rx, ry, rz = ctx.forces_direct_targets(t)
    den = np.sqrt(rx ** 2 + ry ** 2 + rz ** 2)
    trials, best = [], None
    for dth in MLP_THETA_STEPS:
        th = round(theta_p + dth, 4)
        for frac in MLP_TAU_FRACS:
            tau = round(th * frac, 4)
            ctx.set_params(theta=th)
            ctx.set_correction(net, tau)
            ctx.bh_build()
            build_ms = ctx.phase_ms()["Octree creation"]
            ax, ay, az = ctx.bh_forces()
            force_ms = ctx.phase_ms()["Acceleration Kernel Time"]
            evals = ctx.correction_pairs()
            e = np.sqrt((ax[t] - rx) ** 2 + (ay[t] - ry) ** 2 + (az[t] - rz) ** 2) / den
            err = float(np.median(e))
            rec = {"theta": th, "tau": tau, "median_rel_err": err, "build_ms": round(build_ms, 3),
                   "force_ms": round(force_ms, 3), "network_evals": int(evals)}
            trials.append(rec)
            log(f"corrected theta {th} tau {tau}: err {err:.3e}, force {force_ms:.2f} ms, "
                f"{evals / n:.1f} evaluations per body")
            if err <= ERR_GATE * THETA_MARGIN and (
                    best is None or build_ms + force_ms < best["build_ms"] + best["force_ms"]):
                best = rec
    out_lc = dict(train_info, plain_theta=theta_p, plain_value=round(pvalue, 1),
                  plain_median_rel_err=perr, plain_mean_interactions=float(pinter.mean()),
                  plain_theta_search=search_p, corrected_trials=trials)
    value, raw, err, theta, phases, sample = None, None, None, None, {}, 0
    if best is not None:
        theta = best["theta"]
        ctx.set_params(theta=theta)
        ctx.set_correction(net, best["tau"])
        secs, phases, err, sample, _ = timed_run(ctx, d, algorithm, dt, args.steps, args.warmup)
        raw = n * args.steps / secs
        value = gated(raw, err)[0]
        evals = ctx.correction_pairs()
        out_lc.update(tau=best["tau"], network_evals_per_step=int(evals),
                      speedup_vs_plain=round(raw / pvalue, 4))
        log(f"corrected theta {theta} tau {best['tau']}: {raw:.4g} body-steps/s, err {err:.3e}")
    else:
        out_lc.update(speedup_vs_plain=None)
        log("no corrected (theta, tau) met the bound")
    ctx.close()
    accel_ms = phases.get("Acceleration Kernel Time", 0.0) / args.steps
    flops = MLP_FLOP_PER_PAIR * float(out_lc.get("network_evals_per_step", 0))
    achieved = flops / (accel_ms * 1e-3) / 1e12 if accel_ms else 0.0
    out = {
        "metric": METRIC, "value": value, "unit": "body-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(1e3 * n / raw, 4) if raw else None,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "bf16+f32+f64", "data": "synthetic",
        "config": {"workload": args.workload, "bodies": n, "distribution": kind,
                   "algorithm": algorithm, "theta": theta, "dt": dt, "median_rel_err": err,
                   "err_sample": sample, "gate": ERR_GATE,
                   "gate_met": err is not None and err <= ERR_GATE,
                   "ungated_value": None if raw is None else round(raw, 1),
                   "phase_ms_per_step": {k: round(v / args.steps, 4) for k, v in phases.items()},
                   "learned_correction": out_lc},
        "roofline": {"bound": "mfma", "achieved": round(achieved, 3), "peak": BF16_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": round(achieved / BF16_PEAK_TFLOPS, 4),
                     "traffic": None, "kernel": "k_bh_group<true, false>",
                     "kernel_ms": round(accel_ms, 4), "flop_per_launch": flops},
    }
    print(json.dumps(out), flush=True)


def run_ensemble(args):
    """512 independent systems in one persistent force launch per step."""
    c = ENSEMBLE
    log(f"device: {nb.device_name()}")
    t0 = time.perf_counter()
    sizes, st, dts = nb.gen_ensemble(c["nsys"], c["seed"], c["lo_log2"], c["hi_log2"])
    n = int(sizes.sum())
    log(f"ensemble: {c['nsys']} systems, {n} bodies (generated in {time.perf_counter() - t0:.1f} s)")
    e = nb.Ensemble(sizes, theta=c["theta"])
    keys = ("x", "y", "z", "vx", "vy", "vz", "m")

    def reset():
        e.set_state(*[st[k] for k in keys], dts)
        e.forces()  # t = 0 forces, untimed as in the reference

    reset()
    t1 = time.perf_counter()
    tune, tune_ms = e.autotune(steps=2)
    log(f"autotune: {tune} at {tune_ms:.3f} ms/step ({time.perf_counter() - t1:.1f} s)")
    reset()
    if args.warmup:
        e.steps(args.warmup)
    log(f"timing {args.steps} steps ...")
    t0 = time.perf_counter()
    e.steps(args.steps)
    secs = time.perf_counter() - t0
    phases = e.phase_ms()
    value = n * args.steps / secs
    inter = e.interactions().astype(np.float64)
    uses_bh = e.uses_bh()
    # error: sampled targets of sampled systems vs the FP64 direct sum
    x, y, z, vx, vy, vz = e.state()
    ax, ay, az = e.accel()
    off = np.concatenate([[0], np.cumsum(sizes, dtype=np.int64)])
    rng = np.random.default_rng(1)
    errs = []
    for s in rng.choice(c["nsys"], c["err_systems"], replace=False):
        sl = slice(off[s], off[s + 1])
        ns = int(sizes[s])
        ctx = nb.Context(ns)
        ctx.set_state(x[sl], y[sl], z[sl], vx[sl], vy[sl], vz[sl], st["m"][sl])
        t = np.sort(rng.choice(ns, min(ns, c["err_targets"]), replace=False)).astype(np.uint32)
        rx, ry, rz = ctx.forces_direct_targets(t)
        ctx.close()
        gx, gy, gz = ax[sl][t], ay[sl][t], az[sl][t]
        errs.append(np.sqrt((gx - rx) ** 2 + (gy - ry) ** 2 + (gz - rz) ** 2) /
                    np.sqrt(rx ** 2 + ry ** 2 + rz ** 2))
    errs = np.concatenate(errs)
    err = float(np.median(errs))
    log(f"median rel err {err:.3e} over {errs.size} targets of {c['err_systems']} systems")
    accel_ms = phases["Acceleration Kernel Time"] / args.steps
    flops = FLOP_PER_INTERACTION * float(inter.sum())
    achieved = flops / (accel_ms * 1e-3) / 1e12
    gvalue, gate_met = gated(value, err)
    out = {
        "metric": METRIC, "value": gvalue, "unit": "body-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(1e3 * secs / args.steps, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": "ensemble_512", "systems": c["nsys"], "bodies": n,
                   "gate": ERR_GATE, "gate_met": gate_met, "ungated_value": round(value, 1),
                   "sizes": f"floor(2^U[{c['lo_log2']},{c['hi_log2']}])",
                   "distribution": "alternating star+disk (dt 1 h) and plummer (dt 1e-3)",
                   "theta": c["theta"], "tuning": tune, "tuned_ms_per_step": round(tune_ms, 4),
                   "bh_systems": int(uses_bh.sum()), "median_rel_err": err,
                   "err_sample": int(errs.size),
                   "phase_ms_per_step": {k: round(v / args.steps, 4) for k, v in phases.items()},
                   "mean_interactions": float(inter.mean())},
        "roofline": {"bound": "valu", "achieved": round(achieved, 3),
                     "peak": FP32_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": round(achieved / FP32_PEAK_TFLOPS, 4),
                     "traffic": TRAFFIC_BYTES.get(("ensemble_512", c["theta"])),
                     "kernel": "k_ens_forces",
                     "kernel_ms": round(accel_ms, 4), "flop_per_launch": flops},
    }
    e.close()
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = ensemble_cpu_baseline(sizes, st, dts, uses_bh, off)
    print(json.dumps(out), flush=True)


def ensemble_cpu_baseline(sizes, st, dts, uses_bh, off):
    """The reference CPU path (oracle restatement) runs one system at a time:
    one loop iteration per system, Barnes-Hut or naive as the tuned split
    chose, over systems in random order until about CPU_SAMPLE_S seconds of
    loop time, scaled to body-steps/s."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc
    c = ENSEMBLE
    order = np.random.default_rng(2).permutation(c["nsys"])
    bodies, ms, used = 0, 0.0, 0
    for s in order:
        sl = slice(off[s], off[s + 1])
        algo = "BarnesHut" if uses_bh[s] else "naive"
        r = orc.simulate(algo, st["x"][sl], st["y"][sl], st["z"][sl], st["vx"][sl], st["vy"][sl],
                         st["vz"][sl], st["m"][sl], dts[s], dts[s], theta=c["theta"])
        bodies += int(sizes[s]) * r["steps"]
        ms += r["loop_ms"]
        used += 1
        if ms >= CPU_SAMPLE_S * 1e3:
            break
    return {"value": bodies / (ms * 1e-3), "unit": "body-steps/s", "cores": orc.get_threads(),
            "cpu": cpu_model(), "kind": "port",
            "sample": f"1 loop iteration of each of {used} systems in random order "
                      f"({bodies} body-steps), FP64, naive below the tuned threshold and "
                      f"Barnes-Hut above; {ms / 1e3:.1f} s"}


def timed_run(ctx, d, algorithm, dt, steps, warmup):
    """Set the initial state, t = 0 force (untimed, as in the reference),
    `warmup` untimed steps, then `steps` timed loop iterations.  Returns
    (seconds, phase ms summed over the timed steps, median error, sample size,
    interactions per body)."""
    n = d["m"].size
    ctx.set_state(d["x"], d["y"], d["z"], d["vx"], d["vy"], d["vz"], d["m"])
    if algorithm == "BarnesHut":
        ctx.bh_build()
        ctx.bh_forces()
    else:
        ctx.forces_direct()
    if warmup:
        ctx.steps(algorithm, dt, warmup)
    log(f"timing {steps} steps ...")
    t0 = time.perf_counter()
    ctx.steps(algorithm, dt, steps)  # returns after the stream is idle
    t1 = time.perf_counter()
    phases = ctx.phase_ms()
    inter = ctx.interactions().astype(np.float64) if algorithm == "BarnesHut" else None
    err, sample = median_rel_err(ctx, n, np.random.default_rng(1))
    log(f"median rel err {err:.3e} over {sample} targets")
    return t1 - t0, phases, err, sample, inter


def train_correction(theta=None):
    """Train the correction network on the GPU (outside the timed region):
    pairs from MLP_TRAIN['snapshots'] Plummer snapshots, MLP_TRAIN['pairs'] in all,
    accepted by per-body walks at `theta` (default MLP_TRAIN['theta'])."""
    c = dict(MLP_TRAIN)
    if theta is not None:
        c["theta"] = theta
    t0 = time.perf_counter()
    Xs, Ts = [], []
    per = c["pairs"] // c["snapshots"]
    for s in range(c["snapshots"]):
        d = nb.gen_plummer(c["n"], 0x4E424F44591000 + s)
        ctx = nb.Context(c["n"], theta=c["theta"], precision=nb.PREC_FP32C)
        ctx.set_state(d["x"], d["y"], d["z"], d["vx"], d["vy"], d["vz"], d["m"])
        ctx.bh_build()
        t = np.sort(np.random.default_rng(s).choice(c["n"], c["targets"], replace=False))
        pr = ctx.mlp_pairs(c["theta"], t.astype(np.uint32), per)
        Xs.append(pr["features"])
        Ts.append(pr["residuals"])
        ctx.close()
    X, T = np.concatenate(Xs), np.concatenate(Ts)
    t1 = time.perf_counter()
    net = nb.Mlp(seed=7)
    losses = net.train(X, T, epochs=c["epochs"], batch=c["batch"], seed=9)
    t2 = time.perf_counter()
    var = float(np.mean(T.astype(np.float64) ** 2))
    log(f"correction trained: {X.shape[0]} pairs ({t1 - t0:.1f} s), {c['epochs']} epochs in "
        f"{t2 - t1:.1f} s, loss {losses[0]:.3e} -> {losses[-1]:.3e} (target mean square {var:.3e})")
    return net, dict(train_pairs=int(X.shape[0]), train_snapshots=c["snapshots"],
                     train_theta=c["theta"],
                     train_bodies=c["n"], train_epochs=c["epochs"],
                     train_seconds=round(t2 - t1, 3), train_loss=float(losses[-1]),
                     train_target_mean_square=var)


if __name__ == "__main__":
    main()
