"""MC sweeps/s of the checkerboard Monte Carlo sweep on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c4_tchpsc_65536]

A step is one MC sweep (N trial moves, the step loop of SC's
Updater::simulate, scOOP/mc/updater.cpp:206-230) of the whole system, on a
synthetic, seeded state resident in HBM.  Prints one JSON line.

The default workload is 65,536 TCHPSC (L=4, two chiral patches) at packing
fraction 0.25, T*=1: the system the delayed-acceptance metric is quoted on.
The CPU baseline is the FP64 oracle's sequential Metropolis (one core) on a
bounded sample of sweeps of the same state.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "MC sweeps/sec at N=65k PSC, surrogate-DA vs exact, ensemble avgs within 1%"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def read_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 PMC
    summary, if present (profiles/<kernel>_pmc.txt, line 'bytes_per_launch:')."""
    p = os.path.join(ROOT, "profiles", "%s_pmc.txt" % kernel)
    if not os.path.exists(p):
        return None
    for line in open(p):
        if line.startswith("bytes_per_launch:"):
            try:
                return float(line.split(":", 1)[1].split()[0])
            except ValueError:
                return None
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="c4_tchpsc_65536")
    ap.add_argument("--r", type=int, default=2, help="checkerboard stencil radius (1: 2^3 colours, 2: 3^3)")
    ap.add_argument("--trial-mode", type=int, default=1,
                    help="trials per active cell: 0 = its occupancy, 1 = dithered mean occupancy")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--seed", type=int, default=2025)
    ap.add_argument("--copy-probe", action="store_true",
                    help="also run a plain 256 MiB device copy (calibrates rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path (see DESIGN.md, sharding); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import sc_amd
    from sc_amd import prepare
    from sc_amd import workloads as W

    temper, transmx, rot_half = 1.0, 0.3, 20.0 / 180 * np.pi / 4   # 0.3 sigma moves, <= 10 deg rotations
    w = W.workload(a.workload)
    n = len(w["types"])
    g = sc_amd.GpuSystem(0)
    t0 = time.time()
    prepare.initial_state(g, w, seed=a.seed, log=log)
    log("state ready in %.1fs: N=%d box %.3f E/N %.4f" % (time.time() - t0, n, w["box"][0], g.allToAll() / n))

    if a.copy_probe:
        g.copy_probe(256 << 20, reps=3)
        log("copy probe: 3 x k_copy of 256 MiB")
    g.set_trial_mode(a.trial_mode)
    g.sweep(a.warmup, temper, transmx, rot_half, r=a.r, seed=a.seed, counter=1)
    g.set_profiling(True)
    t1 = time.perf_counter()
    st = g.sweep(a.steps, temper, transmx, rot_half, r=a.r, seed=a.seed, counter=10 ** 6)
    t2 = time.perf_counter()
    wall = t2 - t1
    dev_ms, launches, kern_ms = g.last_timing()
    # sweeps actually performed: trial moves / N (trial mode 1 makes N trials
    # per sweep on average, not exactly)
    done = (st["trans_try"] + st["rot_try"]) / float(n)
    sweeps_per_s = done / wall
    e_n = g.allToAll() / n
    log("exact sweep: %.2f sweeps/s (%.3f ms/sweep wall, %.3f ms device), %d pass launches, kernel %.3f ms" % (
        sweeps_per_s, 1e3 * wall / a.steps, dev_ms / a.steps, launches, kern_ms))

    # roofline of the dominant kernel (k_sweep_pass): algorithmic bytes per
    # launch = every record of the stencils read once (N * 44 B: 40-byte
    # record + 4-byte index), the active cells written back (N/ncol * 40 B)
    # and the cell CSR (4 B per cell)
    _, start, _, nc = g.build_cells(r=a.r)
    ncell = int(np.prod(nc))
    ncol = (a.r + 1) ** 3
    bytes_launch = n * 44.0 + n / ncol * 40.0 + 4.0 * (ncell + 1)
    avg_launch_s = kern_ms / 1e3 / max(launches, 1)
    achieved = bytes_launch / avg_launch_s / 1e9
    traffic = read_traffic("k_sweep_pass")

    cpu = None
    if not a.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle as po
        box, pos, d, q = g.get_config()
        tp = po.Topo(w["top"])
        s = po.State(box, pos, d, q, w["types"])
        tc = time.perf_counter()
        s1, _, _, _ = po.mc_run(tp, s, 1, temper, transmx, rot_half, seed=9, energies=False)
        one = time.perf_counter() - tc
        k = max(1, int(a.cpu_seconds / max(one, 1e-3)))
        tc = time.perf_counter()
        po.mc_run(tp, s1, k, temper, transmx, rot_half, seed=10, energies=False)
        el = time.perf_counter() - tc
        cpu = {"value": k / el, "unit": "sweeps/s", "cores": 1, "kind": "port",
               "sample": "%d sequential Metropolis sweeps (FP64 oracle, one core) of the same %d-particle state"
                         % (k, n)}
        log("cpu baseline: %.4f sweeps/s (%d sweeps in %.1fs)" % (k / el, k, el))

    out = {
        "metric": METRIC,
        "value": sweeps_per_s,
        "unit": "sweeps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": 1e3 * wall / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 pair terms, fp64 sums",
        "data": "synthetic seeded state (RSA + compression), no dataset",
        "config": {"workload": a.workload, "n": n, "eta": w["eta"], "box": float(w["box"][0]), "temper": temper,
                   "translation": transmx, "max_rotation_deg": 10.0,
                   "checkerboard": "%dx%dx%d" % ((a.r + 1,) * 3), "cells": list(nc), "sweep": "exact",
                   "trial_mode": a.trial_mode, "trial_moves_per_particle": done / a.steps},
        "exact_sweeps_per_s": sweeps_per_s,
        "acceptance": {"translation": st["trans_acc"] / max(1, st["trans_try"]),
                       "rotation": st["rot_acc"] / max(1, st["rot_try"]),
                       "left_cell": st["left_cell"] / max(1, st["trans_try"])},
        "energy_per_particle": e_n,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "k_sweep_pass", "bytes_per_launch": bytes_launch,
                     "avg_launch_ms": 1e3 * avg_launch_s, "launches": launches},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    g.close()


if __name__ == "__main__":
    main()
