"""MC sweeps/s of the checkerboard Monte Carlo sweep on one MI355X, exact and
with the learned surrogate's delayed acceptance.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c4_tchpsc_65536]

A step is one MC sweep (N trial moves, the step loop of SC's
Updater::simulate, scOOP/mc/updater.cpp:206-230) of the whole system, on a
synthetic, seeded state resident in HBM.  Prints one JSON line.

The default workload is 65,536 TCHPSC (L=4, two chiral patches) at packing
fraction 0.25, T*=1: the system the delayed-acceptance metric is quoted on.
Both sweeps start from the same state: first K exact sweeps, then the
surrogate is trained on the GPU (untimed: 10M exact pair energies over
segment distances 0-2.5 sigma, uniform orientations), then K
delayed-acceptance sweeps.  `value` is the delayed-acceptance rate; the exact
rate sits beside it.  The CPU baseline is the FP64 oracle's sequential
Metropolis (one core) on a bounded sample of sweeps of the same state.
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


def read_traffic(name):
    """HBM bytes per launch from the committed rocprofv3 PMC summary
    profiles/<name>_pmc.txt (line 'bytes_per_launch:'), if present."""
    p = os.path.join(ROOT, "profiles", "%s_pmc.txt" % name)
    if not os.path.exists(p):
        return None
    for line in open(p):
        if line.startswith("bytes_per_launch:"):
            try:
                return float(line.split(":", 1)[1].split()[0])
            except ValueError:
                return None
    return None


def host_cpu():
    """model name and CPU counts of this host (the CPU baseline's cores)"""
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        aff = len(os.sched_getaffinity(0))
    except AttributeError:
        aff = os.cpu_count()
    return {"cpu_model": model, "logical_cpus": os.cpu_count(), "affinity": aff}


def reference_calibration(nsweeps=300, binary=None):
    """The compiled reference's own sequential program (oracle/_ref/scref,
    built from its unmodified sources with -O2 -ffp-contract=off) against the
    oracle's sampler on the same 1,000-PSC state (the plain NVT system: eta
    0.2, T* 1, 0.3 sigma moves, <= 10 deg rotations), one core each, on this
    host: the speed of the port relative to the reference.  None when the
    binary is absent."""
    import subprocess
    import tempfile
    import numpy as np
    import sc_amd
    from sc_amd import workloads as W
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po
    binary = os.path.abspath(binary or po.REF_BIN)
    if not os.path.exists(binary):
        return None
    w = W.workload("c1_psc_1000")
    h = sc_amd.GpuSystem(-1)
    h.load_topology(w["top"])
    pos, d, q = h.generate_rsa(w["types"], w["box"], seed=31)
    tp, st = po.Topo(w["top"]), po.State(w["box"], pos, d, q, w["types"])
    t0 = time.perf_counter()
    po.mc_run(tp, st, nsweeps, 1.0, 0.3, 20.0 / 180 * np.pi / 4, seed=17, energies=False)
    port = nsweeps / (time.perf_counter() - t0)
    with tempfile.TemporaryDirectory() as dd:
        open(os.path.join(dd, "options"), "w").write(
            W.options_text(nsweeps=nsweeps, transmx=0.15, rotmx=20.0, temper=1.0, seed=4321))
        open(os.path.join(dd, "top.init"), "w").write(w["top"])
        open(os.path.join(dd, "config.init"), "w").write(W.config_text(w["box"], pos, d, q))
        r = subprocess.run([binary, "run", os.path.join(dd, "wall.txt")], cwd=dd, stdout=subprocess.DEVNULL,
                           stderr=subprocess.DEVNULL, timeout=600)
        if r.returncode != 0:
            return None
        ref = nsweeps / float(open(os.path.join(dd, "wall.txt")).read().split()[0])
    return {"workload": "1,000 PSC, eta 0.2, T* 1 (the plain NVT system), same seeded state", "sweeps": nsweeps,
            "cores": 1, "reference_sweeps_per_s": ref, "port_sweeps_per_s": port, "reference_over_port": ref / port,
            "reference_build": "scOOP sources, g++ -O2 -ffp-contract=off -fno-fast-math (TotalEMatrix, pair list)"}


def npt_leg(a, transmx, rot_half):
    """NPT self-assembly of bonded chains: SPA head + two PSC tails, harmonic
    bonds, P* = 0.5, T* = 1.1, one isotropic volume move (two whole-system
    energies) after every sweep; exact sweep, the cheap-terms screen and the
    network screen (trained on the tails' PSC pair).  A step is one sweep
    plus one volume move."""
    import numpy as np
    import sc_amd
    from sc_amd import workloads as W
    gn = sc_amd.GpuSystem(0)
    gn.set_block_waves(a.waves)
    gn.set_trial_mode(a.trial_mode)
    gn.set_visit_factor(a.visit)
    P_, T_ = 0.5, 1.1
    nch = a.npt_chains
    boxc, posc, dc, qc, tc_ = W.chain_lattice(nch, 0.25, a.seed)
    gn.load_topology(W.chain_topology(nch))
    nb_ = len(tc_)

    def npt_steps(k, mode, counter):
        gn.set_surrogate(mode)
        acc, tot = 0, {"sweep_s": 0.0, "volume_s": 0.0}
        for s_ in range(k):
            t_ = time.perf_counter()
            st_ = gn.sweep(1, T_, transmx, rot_half, r=a.r, seed=a.seed + 3, counter=counter + s_)
            t2_ = time.perf_counter()
            for kk, v in st_.items():
                tot[kk] = tot.get(kk, 0) + v
            ok, bx = gn.volume_move(1, P_, T_, a.npt_edge, seed=a.seed + 4, counter=counter + s_)
            tot["sweep_s"] += t2_ - t_
            tot["volume_s"] += time.perf_counter() - t2_
            acc += ok
        return acc, tot, bx

    npt = {"chains": nch, "particles": nb_, "press": P_, "temper": T_, "volume_move": "isotropic, every sweep",
           "edge_mx": a.npt_edge, "start": "straight chains on a lattice at packing fraction 0.25"}
    gn.set_config(boxc, posc, dc, qc, tc_)
    npt_steps(a.npt_warmup, 0, 0)
    bx0, p0, d0_, q0_ = gn.get_config()
    npt["eta_start"] = nch * W.chain_volume() / float(np.prod(bx0))
    legs_ = [("exact", 0)] + ([] if a.no_da else [("cheap_terms_screen", 2), ("network_screen", 1)])
    for label, mode in legs_:
        gn.set_config(bx0, p0, d0_, q0_, tc_)
        if mode == 1:   # the tails' pair (PSC-PSC); pairs with a head screen with 0
            from sc_amd import surrogate
            npt["surrogate_training"] = surrogate.train(gn, 2, nsamples=a.samples, epochs=a.epochs,
                                                        seed=a.seed + 5, log=log)
            gn.set_config(bx0, p0, d0_, q0_, tc_)
        npt_steps(5, mode, 10 ** 5)
        gn.set_config(bx0, p0, d0_, q0_, tc_)
        t1 = time.perf_counter()
        acc, tot, bx = npt_steps(a.npt_steps, mode, 10 ** 6)
        dt = time.perf_counter() - t1
        tries = max(1, tot["trans_try"] + tot["rot_try"])
        npt[label] = {"sweeps_per_s": a.npt_steps / dt, "ms_per_step": 1e3 * dt / a.npt_steps,
                      "ms_sweep": 1e3 * tot["sweep_s"] / a.npt_steps, "ms_volume_move": 1e3 * tot["volume_s"] / a.npt_steps,
                      "volume_acceptance": acc / float(a.npt_steps),
                      "move_acceptance": (tot["trans_acc"] + tot["rot_acc"]) / tries,
                      "bond_rejected": int(tot["bond_rejected"]), "left_cell": int(tot["left_cell"]),
                      "eta_end": nch * W.chain_volume() / float(np.prod(bx)),
                      "energy_per_particle": gn.allToAll() / nb_}
        log("npt chains %s: %.2f steps/s (sweep + volume move), volume acc %.3f, eta %.4f -> %.4f" % (
            label, npt[label]["sweeps_per_s"], npt[label]["volume_acceptance"], npt["eta_start"],
            npt[label]["eta_end"]))
    if a.npt_chainprob > 0:
        # the exact sweep plus whole-molecule moves (chainMove) at the reference's usual option values
        gn.set_config(bx0, p0, d0_, q0_, tc_)
        mx_, ang_ = 2.0 * a.npt_chainmmx, a.npt_chainrmx / 180.0 * np.pi / 2.0   # sim.h:362, 366
        gn.set_chain_moves(a.npt_chainprob, mx_, ang_)
        gn.set_chain_visit_factor(a.npt_chain_visit)
        npt_steps(5, 0, 10 ** 5)
        gn.set_config(bx0, p0, d0_, q0_, tc_)
        gn.set_chain_moves(a.npt_chainprob, mx_, ang_)
        t1 = time.perf_counter()
        acc, tot, bx = npt_steps(a.npt_steps, 0, 10 ** 6)
        dt = time.perf_counter() - t1
        lab = "exact_with_chain_moves"
        npt[lab] = {"chainprob": a.npt_chainprob, "chainmmx": a.npt_chainmmx, "chainrmx_deg": a.npt_chainrmx,
                    "extent": gn.chain_extent(), "colour_period": gn.chain_pass_info()[0],
                    "chain_visit_factor": gn.chain_visit_info(),
                    "lds_stage": gn.chain_pass_info()[1], "sweeps_per_s": a.npt_steps / dt,
                    "ms_per_step": 1e3 * dt / a.npt_steps, "ms_sweep": 1e3 * tot["sweep_s"] / a.npt_steps,
                    "ms_volume_move": 1e3 * tot["volume_s"] / a.npt_steps,
                    "volume_acceptance": acc / float(a.npt_steps),
                    "chain_moves_per_sweep": (tot["chain_disp_try"] + tot["chain_rot_try"]) / float(a.npt_steps),
                    "chain_displace_acceptance": tot["chain_disp_acc"] / max(1, tot["chain_disp_try"]),
                    "chain_rotate_acceptance": tot["chain_rot_acc"] / max(1, tot["chain_rot_try"]),
                    "chain_rejected": int(tot["chain_rejected"]),
                    "eta_end": nch * W.chain_volume() / float(np.prod(bx)), "energy_per_particle": gn.allToAll() / nb_}
        if "exact" in npt:
            npt[lab]["ms_chain_moves_per_sweep"] = npt[lab]["ms_sweep"] - npt["exact"]["ms_sweep"]
        log("npt chains %s: %.2f steps/s, %.0f chain moves per sweep (displace acc %.3f, rotate acc %.3f; "
            "period %d, visit factor %d, LDS stage %s, cells skipped %d)" % (
                lab, npt[lab]["sweeps_per_s"], npt[lab]["chain_moves_per_sweep"],
                npt[lab]["chain_displace_acceptance"], npt[lab]["chain_rotate_acceptance"], npt[lab]["colour_period"],
                npt[lab]["chain_visit_factor"], npt[lab]["lds_stage"], tot["chain_cells_skipped"]))
        npt[lab]["chain_cells_skipped"] = int(tot["chain_cells_skipped"])
        gn.set_chain_moves(0.0, 0.0, 0.0)
        gn.set_chain_visit_factor(1)
    gn.close()
    return npt


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--workload", default="c4_tchpsc_65536")
    ap.add_argument("--r", type=int, default=1, help="checkerboard stencil radius (1: 2^3 colours, 2: 3^3)")
    ap.add_argument("--waves", type=int, default=0,
                    help="waves per trial in the sweep kernels (1, 2, 4: one trial at a time; 0: the speculative "
                         "sweep, four waves with a trial each)")
    ap.add_argument("--trial-mode", type=int, default=1,
                    help="trials per active cell: 0 = its occupancy, 1 = dithered mean occupancy")
    ap.add_argument("--visit", type=int, default=8,
                    help="visit factor: a sweep visits 1/v of the colours with v times the trials per cell "
                         "(8: one colour per sweep, 8 x the mean occupancy per active cell)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--seed", type=int, default=2025)
    ap.add_argument("--copy-probe", action="store_true",
                    help="also run a plain 256 MiB device copy (calibrates rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    ap.add_argument("--samples", type=int, default=10_000_000, help="surrogate training pairs")
    ap.add_argument("--epochs", type=int, default=10)
    ap.add_argument("--no-da", action="store_true", help="exact sweep only")
    ap.add_argument("--no-exact", action="store_true", help="delayed-acceptance sweep only")
    ap.add_argument("--no-pt", action="store_true", help="skip the batched parallel-tempering leg")
    ap.add_argument("--pt-sweeps", type=int, default=100)
    ap.add_argument("--equil", type=int, default=4000, help="untimed exact sweeps before the legs")
    ap.add_argument("--no-npt", action="store_true", help="skip the NPT chain leg")
    ap.add_argument("--no-psc", action="store_true", help="skip the 65,536-PSC leg")
    ap.add_argument("--psc-equil", type=int, default=1000, help="untimed exact sweeps of the PSC leg")
    ap.add_argument("--npt-chainprob", type=float, default=0.1,
                    help="chain-move probability per step of the NPT leg's chain-move run (0: no such run)")
    ap.add_argument("--npt-chainmmx", type=float, default=0.1, help="options' chainmmx (displacement, doubled)")
    ap.add_argument("--npt-chainrmx", type=float, default=5.0, help="options' chainrmx in degrees")
    ap.add_argument("--npt-chain-visit", type=int, default=27,
                    help="chain-move visit factor (the largest divisor of the coarse colour count not above it)")
    ap.add_argument("--npt-chains", type=int, default=43690, help="3-bead chains of the NPT leg (131,070 beads)")
    ap.add_argument("--npt-steps", type=int, default=200, help="timed NPT steps (one sweep + one volume move)")
    ap.add_argument("--npt-warmup", type=int, default=50)
    ap.add_argument("--npt-edge", type=float, default=0.1, help="edge_mx of the isotropic volume move")
    ap.add_argument("--npt-only", action="store_true",
                    help="run the NPT chain leg alone and print its result (profiling runs)")
    ap.add_argument("--ensemble", type=int, default=1000,
                    help="untimed sweeps after each timed leg over which E/N is averaged")
    a = ap.parse_args()
    if a.npt_only:
        import numpy as np
        print(json.dumps({"npt_chains": npt_leg(a, 0.3, 20.0 / 180 * np.pi / 4)}), flush=True)
        return
    if a.gpus != 1:
        print("bench.py runs on one GPU only: the sweep has no multi-GPU path (spatial sharding is designed, not built); "
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
    g.set_block_waves(a.waves)
    g.set_visit_factor(a.visit)
    if a.equil:
        t0 = time.time()
        g.sweep(a.equil, temper, transmx, rot_half, r=a.r, seed=a.seed + 1, counter=5 * 10 ** 6)
        log("equilibrated: %d exact sweeps in %.1fs, E/N %.4f" % (a.equil, time.time() - t0, g.allToAll() / n))
    box0, pos0, d0, q0 = g.get_config()
    _, start, _, nc = g.build_cells(r=a.r)
    ncell = int(np.prod(nc))
    ncol = (a.r + 1) ** 3
    # algorithmic bytes per sweep-pass launch (k_sweep_spec, the speculative
    # sweep; k_sweep_pass with --waves 1/2/4): every record of the stencils
    # read once (N * 44 B: 40-byte record + 4-byte index), the active cells
    # written back (N/ncol * 40 B) and the cell CSR (4 B per cell)
    bytes_launch = n * 44.0 + n / ncol * 40.0 + 4.0 * (ncell + 1)

    def leg(label, mode):
        g.set_config(box0, pos0, d0, q0, w["types"])
        g.set_surrogate(mode)
        g.set_profiling(False)
        g.sweep(a.warmup, temper, transmx, rot_half, r=a.r, seed=a.seed, counter=1)
        g.set_profiling(True)
        t1 = time.perf_counter()
        st = g.sweep(a.steps, temper, transmx, rot_half, r=a.r, seed=a.seed, counter=10 ** 6)
        t2 = time.perf_counter()
        wall = t2 - t1
        dev_ms, launches, kern_ms = g.last_timing()
        # sweeps actually performed: trial moves / N (trial mode 1 makes N
        # trials per sweep on average, not exactly)
        done = (st["trans_try"] + st["rot_try"]) / float(n)
        sps = done / wall
        avg_launch_s = kern_ms / 1e3 / max(launches, 1)
        achieved = bytes_launch / avg_launch_s / 1e9
        tries = max(1, st["trans_try"] + st["rot_try"])
        acc = st["trans_acc"] + st["rot_acc"]
        kern = "k_sweep_spec" if a.waves == 0 else "k_sweep_pass"
        kname = kern + {0: "<DA_OFF>", 1: "<DA_NET>", 2: "<DA_REP>"}[mode]
        r = {"sweeps_per_s": sps, "ms_per_step": 1e3 * wall / a.steps, "wall_s": wall, "device_ms": dev_ms,
             "trial_moves_per_particle": done / a.steps,
             "acceptance": {"translation": st["trans_acc"] / max(1, st["trans_try"]),
                            "rotation": st["rot_acc"] / max(1, st["rot_try"]),
                            "left_cell": st["left_cell"] / max(1, st["trans_try"]), "overall": acc / tries},
             "energy_per_particle": g.allToAll() / n,
             "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                          "frac": achieved / HBM_PEAK_GBS,
                          "traffic": read_traffic(kern + {0: "_exact", 1: "_da_net", 2: "_da_rep"}[mode])
                          if a.waves == 0 else None,
                          "kernel": kname, "bytes_per_launch": bytes_launch, "avg_launch_ms": 1e3 * avg_launch_s,
                          "launches": launches}}
        # ensemble average of E/N over the following untimed sweeps (every 25)
        es = []
        for k in range(a.ensemble // 25):
            g.sweep(25, temper, transmx, rot_half, r=a.r, seed=a.seed, counter=2 * 10 ** 6 + 25 * k)
            es.append(g.allToAll() / n)
        if es:
            es = np.array(es)
            nb = min(8, len(es))
            blocks = es[: len(es) // nb * nb].reshape(nb, -1).mean(1)
            r["ensemble_energy_per_particle"] = {"mean": float(es.mean()), "stderr": float(blocks.std(ddof=1) /
                                                                                        np.sqrt(nb)) if nb > 1 else None,
                                                 "sweeps": len(es) * 25}
        if mode:
            r["screen1_pass"] = st["stage1_acc"] / max(1, st["stage1_try"])
            passed = st["stage1_acc"]
            if mode == 1:
                # the network runs for the trials past screen 1 with enough entries in range
                # (stage2_try of them); the others reach the exact stage unscreened
                r["network_screened"] = st["stage2_try"] / max(1, st["stage1_try"])
                r["screen2_pass"] = st["stage2_acc"] / max(1, st["stage2_try"])
                passed = st["stage1_acc"] - st["stage2_try"] + st["stage2_acc"]
            r["exact_stage_accept"] = acc / max(1, passed)
        log("%s sweep: %.2f sweeps/s (%.3f ms/sweep wall, %.3f ms device), %d pass launches, kernel %.3f ms, "
            "acceptance %.4f" % (label, sps, 1e3 * wall / a.steps, dev_ms / a.steps, launches, kern_ms, acc / tries))
        return r

    exact = None if a.no_exact else leg("exact", 0)
    da = None
    rep = None
    train = None
    if not a.no_da:
        rep = leg("delayed-acceptance, cheap-terms screen only", 2)
        from sc_amd import surrogate
        ptype = int(w["types"][0])
        train = surrogate.train(g, ptype, nsamples=a.samples, epochs=a.epochs, seed=a.seed, log=log)
        da = leg("delayed-acceptance, network screen", 1)
    main_leg = da if da is not None else exact

    # the same two sweeps on 65,536 one-patch PSC (L = 3, eta 0.25), the
    # system the metric's name quotes
    pscres = None
    if not a.no_psc:
        from sc_amd import surrogate
        wp = W.workload("psc_65536")
        gq = sc_amd.GpuSystem(0)
        gq.set_trial_mode(a.trial_mode)
        gq.set_block_waves(a.waves)
        prepare.initial_state(gq, wp, seed=a.seed + 7)
        gq.set_visit_factor(a.visit)
        gq.sweep(a.psc_equil, temper, transmx, rot_half, r=a.r, seed=a.seed + 8, counter=5 * 10 ** 6)
        bp = gq.get_config()
        pscres = {"workload": "psc_65536", "n": len(wp["types"]), "eta": wp["eta"], "box": float(wp["box"][0]),
                  "equilibration_sweeps": a.psc_equil}
        for label, mode in (("exact", 0), ("delayed_acceptance", 1)):
            gq.set_config(*bp, wp["types"])
            if mode == 1:
                pscres["surrogate_training"] = surrogate.train(gq, int(wp["types"][0]), nsamples=a.samples,
                                                               epochs=a.epochs, seed=a.seed + 9, log=log)
                gq.set_config(*bp, wp["types"])
            gq.set_surrogate(mode)
            gq.sweep(a.warmup, temper, transmx, rot_half, r=a.r, seed=a.seed, counter=1)
            t1 = time.perf_counter()
            st = gq.sweep(a.steps, temper, transmx, rot_half, r=a.r, seed=a.seed, counter=10 ** 6)
            dt = time.perf_counter() - t1
            tries = st["trans_try"] + st["rot_try"]
            pscres[label] = {"sweeps_per_s": tries / float(len(wp["types"])) / dt,
                             "acceptance": (st["trans_acc"] + st["rot_acc"]) / max(1, tries),
                             "energy_per_particle": gq.allToAll() / len(wp["types"])}
            log("psc 65536 %s: %.2f sweeps/s, acceptance %.4f" % (label, pscres[label]["sweeps_per_s"],
                                                                  pscres[label]["acceptance"]))
        pscres["da_over_exact"] = pscres["delayed_acceptance"]["sweeps_per_s"] / pscres["exact"]["sweeps_per_s"]
        gq.close()

    # batched parallel tempering: 64 replicas of uneven size in one context,
    # autotuned launch shape, exchanges every 10 sweeps
    ptres = None
    if not a.no_pt:
        from sc_amd import pt
        gp = sc_amd.GpuSystem(0)
        gp.set_block_waves(a.waves)
        gp.set_visit_factor(a.visit)   # for every replica whose colour count it divides
        t0 = time.time()
        nrep_n, temps, _ = pt.c3_batch(gp)
        tb = time.time() - t0
        r_t, m_t, rates = gp.autotune(3, temps, transmx, rot_half, seed=a.seed)
        pt.run(gp, temps, 10, exchange_every=10, seed=a.seed, counter=0)   # warm-up
        t1 = time.perf_counter()
        temps2, tot, acc, tried = pt.run(gp, temps, a.pt_sweeps, exchange_every=10, seed=a.seed, counter=10 ** 6)
        dt = time.perf_counter() - t1
        moves = sum(x["trans_try"] + x["rot_try"] for x in tot)
        ptres = {"replicas": int(len(nrep_n)), "particles": int(nrep_n.sum()),
                 "n_min": int(nrep_n.min()), "n_max": int(nrep_n.max()),
                 "temperatures": [float(temps[0]), float(temps[-1])], "build_s": tb,
                 "autotune": {"r": r_t, "trial_mode": m_t, "sweeps_per_s_r1m0_r1m1_r2m0_r2m1": rates.tolist()},
                 "sweeps": a.pt_sweeps, "exchange_every": 10, "wall_s": dt,
                 "batch_sweeps_per_s": moves / float(nrep_n.sum()) / dt, "particle_moves_per_s": moves / dt,
                 "exchange_acceptance": float(acc.sum()) / max(1, int(tried.sum())),
                 "launches_per_sweep": (8 // a.visit if 8 % a.visit == 0 else 8) if r_t == 1
                 else (27 // a.visit if 27 % a.visit == 0 else 27)}
        log("batched tempering: %d replicas / %d particles, %.2f batch sweeps/s, %.3e moves/s, exchanges %d/%d" % (
            len(nrep_n), nrep_n.sum(), ptres["batch_sweeps_per_s"], ptres["particle_moves_per_s"], acc.sum(),
            tried.sum()))
        gp.close()

    npt = None if a.no_npt else npt_leg(a, transmx, rot_half)

    cpu = None
    if not a.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle as po
        box, pos, d, q = box0, pos0, d0, q0
        tp = po.Topo(w["top"])
        s = po.State(box, pos, d, q, w["types"])
        tc = time.perf_counter()
        s1, _, _, _ = po.mc_run(tp, s, 1, temper, transmx, rot_half, seed=9, energies=False)
        one = time.perf_counter() - tc
        k = max(1, int(a.cpu_seconds / max(one, 1e-3)))
        # one independent replica per host core (the oracle's C sampler
        # releases the GIL), each running k sequential Metropolis sweeps
        from concurrent.futures import ThreadPoolExecutor
        # one replica per core of this process's CPU share: the host's
        # affinity set, capped by the share a one-GPU job gets on these
        # boxes (OMP_NUM_THREADS, 16)
        hc = host_cpu()
        share = int(os.environ.get("OMP_NUM_THREADS") or 16)
        cores = max(1, min(share, hc["affinity"] or 1))

        def replica(i):
            t = time.perf_counter()
            po.mc_run(tp, s1, k, temper, transmx, rot_half, seed=10 + i, energies=False)
            return time.perf_counter() - t

        tc = time.perf_counter()
        with ThreadPoolExecutor(max_workers=cores) as ex:
            per = list(ex.map(replica, range(cores)))
        el = time.perf_counter() - tc
        agg = cores * k / el
        cpu = {"value": agg, "unit": "sweeps/s", "cores": cores, "kind": "port",
               "per_core": k / float(np.mean(per)),
               "sample": "%d independent replicas (one per core of the CPU share) x %d sequential Metropolis sweeps "
                         "(FP64 oracle, cell list) of the same %d-particle state; value = aggregate" % (cores, k, n),
               "host": dict(hc, cpu_share=share),
               "note": "the compiled reference cannot run this system: its TotalEMatrix holds two N(N-1)/2 FP64 "
                       "matrices (69 GB at N=65,536 per process)"}
        cal = reference_calibration()
        if cal:
            cpu["reference_calibration"] = cal
            log("reference calibration: reference %.1f, port %.1f sweeps/s (1 core, 1,000 PSC)" % (
                cal["reference_sweeps_per_s"], cal["port_sweeps_per_s"]))
        log("cpu baseline: %.4f sweeps/s aggregate over %d threads (%.4f per core; %d sweeps each in %.1fs)" % (
            agg, cores, cpu["per_core"], k, el))

    out = {
        "metric": METRIC,
        "value": main_leg["sweeps_per_s"],
        "unit": "sweeps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": main_leg["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 pair terms, fp64 sums" + (", bf16 mfma surrogate (fp32 accumulate)" if da else ""),
        "data": "synthetic seeded state (RSA + compression), no dataset; surrogate trained on synthetic exact pairs",
        "config": {"workload": a.workload, "n": n, "eta": w["eta"], "box": float(w["box"][0]), "temper": temper,
                   "translation": transmx, "max_rotation_deg": 10.0,
                   "checkerboard": "%dx%dx%d" % ((a.r + 1,) * 3), "cells": list(nc),
                   "sweep": "delayed-acceptance" if da else "exact", "trial_mode": a.trial_mode,
                   "visit_factor": a.visit},
        "exact_sweeps_per_s": exact["sweeps_per_s"] if exact else None,
        "da_sweeps_per_s": da["sweeps_per_s"] if da else None,
        "da_over_exact": (da["sweeps_per_s"] / exact["sweeps_per_s"]) if (da and exact) else None,
        "exact": exact,
        "delayed_acceptance": da,
        "delayed_acceptance_without_network": rep,
        "ensemble_rel_diff_vs_exact": {
            k: (abs(v["ensemble_energy_per_particle"]["mean"] / exact["ensemble_energy_per_particle"]["mean"] - 1.0)
                if (v and exact and "ensemble_energy_per_particle" in v and "ensemble_energy_per_particle" in exact)
                else None) for k, v in (("delayed_acceptance", da), ("delayed_acceptance_without_network", rep))},
        "surrogate_training": train,
        "batched_tempering": ptres,
        "npt_chains": npt,
        "psc_65536": pscres,
        "roofline": main_leg["roofline"],
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    g.close()


if __name__ == "__main__":
    main()
