"""The checkerboard sweep on the two-patch chiral and the cylindrical-patch
species, through the C ABI, against the FP64 oracle.

The sweep's trial energies have their own code for these species: the
patch-pair enumeration of the TCHPSC four-term sum and the chiral axes
(SpheroCylinder::operator(), paire.h:1141-1170; Particle::init,
particle.cpp:50-78), taken from a per-pass LDS cache of the home cell's
derived vectors.  Systems: 1,000 TCHPSC at packing fraction 0.25 (the
delayed-acceptance benchmark species: two patches 90 deg apart, chirality 10
deg, L = 4) and 1,000 CPSC at 0.3 (the whole-system energy species, L = 3),
each compressed and equilibrated on the GPU, with a network trained on its
own pair.  Every sweep is 2x2x2 with the dithered trial count (the
benchmark's settings).

Tolerances:
  * energy bookkeeping: allToAll(end) - allToAll(start) equals the sum of the
    accepted FP64 trial differences of FP32 pair terms within 1e-8 |E|;
  * allToAll against the FP64 oracle at the end: 1e-5 relative;
  * cached and uncached derived vectors: the same trajectory bit for bit;
  * ensembles (exact, cheap-terms screen, network screen) against the
    oracle's sequential Metropolis from the same state: E/N and mean cluster
    size within 3 combined block errors + 1%, S2 within 3 errors + 0.03
    (S2 of an isotropic fluid is O(N^-1/2) and decorrelates slowly).
"""
import numpy as np
import pytest

import sc_amd
from sc_amd import prepare
from sc_amd import surrogate
from sc_amd import workloads as W
import pyoracle as po

pytestmark = pytest.mark.gpu
ROT_HALF = 20.0 / 180 * np.pi / 4
SYSTEMS = {
    # name: (workload, n, packing fraction, reference type of the network)
    "tchpsc": ("c4_tchpsc_65536", 1000, 0.25, 3),
    "cpsc": ("c2_cpsc_32768", 1000, 0.30, 2),
}


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def prepared():
    """name -> (workload, (box, pos, dir, pd0) after 3000 exact sweeps, trained network weights)"""
    out = {}
    for name, (wl, n, eta, ptype) in SYSTEMS.items():
        say("prepare %s: %d particles, eta %.2f" % (name, n, eta))
        w = W.workload(wl, n_override=n)
        w["box"] = w["box"] * (w["eta"] / eta) ** (1.0 / 3.0)
        w["eta"] = eta
        g = sc_amd.GpuSystem(0)
        prepare.initial_state(g, w, seed=11)
        g.set_trial_mode(1)
        g.sweep(3000, 1.0, 0.3, ROT_HALF, r=1, seed=12, counter=0)
        state = g.get_config()
        surrogate.train(g, ptype, nsamples=1 << 21, epochs=6, seed=13, log=say)
        wts = g.mlp_get_weights()
        say("%s: E/N after equilibration %.4f" % (name, g.allToAll() / n))
        g.close()
        out[name] = (w, state, wts)
    return out


def gpu_system(prepared, name, mode, home_cache=None):
    w, (box, pos, d, q), wts = prepared[name]
    g = sc_amd.GpuSystem(0)
    g.load_topology(w["top"])
    if mode == 1:
        g.mlp_init(SYSTEMS[name][3], seed=1)
        g.mlp_set_weights(wts)
        g.set_network_min_entries(0)   # the network screens every trial that passes the first screen
    g.set_config(box, pos, d, q, w["types"])
    g.set_trial_mode(1)
    g.set_surrogate(mode)
    if home_cache is not None:
        g.set_home_cache(home_cache)
    return g


@pytest.mark.parametrize("name", list(SYSTEMS))
@pytest.mark.parametrize("mode", [0, 1, 2])
def test_sweep_energy_bookkeeping_and_oracle_energy(prepared, name, mode):
    say("case: %s sweep, surrogate mode %d: drift and allToAll vs oracle" % (name, mode))
    w = prepared[name][0]
    n = len(w["types"])
    g = gpu_system(prepared, name, mode)
    e0 = g.allToAll()
    s = g.sweep(30, 1.0, 0.3, ROT_HALF, r=1, seed=21 + mode, counter=0)
    e1 = g.allToAll()
    drift = e1 - e0 - s["energy_change"]
    tries = s["trans_try"] + s["rot_try"]
    acc = s["trans_acc"] + s["rot_acc"]
    say("E %.6f -> %.6f, sum dE %.6f, drift %.3e (%.2e of |E|); %d trials (%.3f N/sweep), acceptance %.3f, "
        "screen 1 %d/%d, screen 2 %d/%d" % (e0, e1, s["energy_change"], drift, abs(drift) / abs(e1), tries,
                                             tries / (30.0 * n), acc / tries, s["stage1_acc"], s["stage1_try"],
                                             s["stage2_acc"], s["stage2_try"]))
    assert abs(drift) <= 1e-8 * abs(e1)
    assert abs(tries / (30.0 * n) - 1.0) < 0.05
    assert 0 < acc < tries
    if mode:
        assert s["stage1_try"] > 0 and 0 < s["stage1_acc"] < s["stage1_try"]
        assert (s["stage2_try"] > 0) == (mode == 1)
    b, p, dd, qq = g.get_config()
    g.close()
    tp, st = po.Topo(w["top"]), po.State(b, p, dd, qq, w["types"])
    eo = po.energy_all(tp, st, use_cells=True)
    say("allToAll gpu %.8f oracle %.8f rel %.2e" % (e1, eo, abs(e1 - eo) / abs(eo)))
    assert abs(e1 - eo) <= 1e-5 * abs(eo)


@pytest.mark.parametrize("name", list(SYSTEMS))
@pytest.mark.parametrize("mode", [0, 1])
def test_home_cache_reproduces_uncached_derivation(prepared, name, mode):
    """The LDS cache of the home cell's derived vectors (patch sides, second
    patch, chiral axes) gives the trajectory of deriving them at every trial:
    the same counters and the same configuration bit for bit."""
    say("case: %s sweep, mode %d, derived-vector cache on and off" % (name, mode))
    out = []
    for cap in (64, 0):
        g = gpu_system(prepared, name, mode, home_cache=cap)
        s = g.sweep(10, 1.0, 0.3, ROT_HALF, r=1, seed=5, counter=0)
        out.append((s, g.get_config()))
        g.close()
    (s1, c1), (s2, c2) = out
    for k in ("trans_acc", "trans_try", "rot_acc", "rot_try", "left_cell", "stage1_acc", "stage1_try",
              "stage2_acc", "stage2_try"):
        assert s1[k] == s2[k], k
    say("accepted %d / %d, energy change %.9f vs %.9f" % (s1["trans_acc"] + s1["rot_acc"],
                                                          s1["trans_try"] + s1["rot_try"], s1["energy_change"],
                                                          s2["energy_change"]))
    assert abs(s1["energy_change"] - s2["energy_change"]) <= 1e-9 * max(1.0, abs(s1["energy_change"]))
    for a, b in zip(c1, c2):
        np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("name", list(SYSTEMS))
@pytest.mark.parametrize("mode", [0, 1, 2])
@pytest.mark.parametrize("r,visit", [(1, 8), (2, 9)])
def test_speculative_sweep_reproduces_sequential(prepared, name, mode, r, visit):
    """A trial per wave, kept in order up to the first dependence on an
    accepted trial of the round, gives the one-trial-at-a-time trajectory:
    the same counters and the same configuration bit for bit.  r = 1 at
    visit factor 8 (~128 trials per cell visit) and r = 2 at visit factor 9
    (cells of a few particles, so a round often tries one particle twice)."""
    say("case: %s sweep, mode %d, r %d, visit %d, speculative vs one trial at a time" % (name, mode, r, visit))
    out = []
    for nw in (0, 1):
        g = gpu_system(prepared, name, mode)
        g.set_visit_factor(visit)
        g.set_block_waves(nw)
        s = g.sweep(20, 1.0, 0.3, ROT_HALF, r=r, seed=7, counter=0)
        out.append((s, g.get_config()))
        g.close()
    (s1, c1), (s2, c2) = out
    for k in ("trans_acc", "trans_try", "rot_acc", "rot_try", "left_cell", "stage1_acc", "stage1_try",
              "stage2_acc", "stage2_try"):
        assert s1[k] == s2[k], k
    say("accepted %d / %d, energy change %.9f vs %.9f" % (s1["trans_acc"] + s1["rot_acc"],
                                                          s1["trans_try"] + s1["rot_try"], s1["energy_change"],
                                                          s2["energy_change"]))
    assert abs(s1["energy_change"] - s2["energy_change"]) <= 1e-9 * max(1.0, abs(s1["energy_change"]))
    for a, b in zip(c1, c2):
        np.testing.assert_array_equal(a, b)


def test_speculative_sweep_with_tables_in_global_memory():
    """Five particle types (TCHPSC copies), more than the sweep keeps in LDS
    beside the stage: the speculative sweep reads the pair and type tables
    from global memory (the other instantiation of the kernel) and gives the
    one-trial-at-a-time trajectory bit for bit."""
    say("case: 5-type TCHPSC system (tables in global memory), speculative vs one trial at a time")
    fields = W.TYPES["tchpsc"].split()[2:]
    top = W.top_init(["%s %d %s" % ("ABCDE"[k], k + 1, " ".join(fields)) for k in range(5)], [200] * 5)
    types = np.repeat(np.arange(1, 6), 200).astype(np.int32)
    box = np.full(3, (1000 * W.particle_volume(1.0, 4.0) / 0.15) ** (1.0 / 3.0))
    out = []
    for nw in (0, 1):
        g = sc_amd.GpuSystem(0)
        g.load_topology(top)
        pos, d, q = g.generate_rsa(types, box, 21)
        g.set_config(box, pos, d, q, types)
        g.set_trial_mode(1)
        g.set_visit_factor(8)
        g.set_block_waves(nw)
        s = g.sweep(20, 1.0, 0.3, ROT_HALF, r=1, seed=7, counter=0)
        out.append((s, g.get_config()))
        g.close()
    (s1, c1), (s2, c2) = out
    say("accepted %d / %d" % (s1["trans_acc"] + s1["rot_acc"], s1["trans_try"] + s1["rot_try"]))
    for k in ("trans_acc", "trans_try", "rot_acc", "rot_try", "left_cell"):
        assert s1[k] == s2[k], k
    assert s1["trans_acc"] + s1["rot_acc"] > 0
    for a, b in zip(c1, c2):
        np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("name", list(SYSTEMS))
def test_network_threshold_is_the_same_in_every_variant(prepared, name):
    """With the network screening only trials of at least the default number
    of (pair, state) entries, the speculative sweep, the 4-wave sweep (the
    network on the last wave beside phase C) and the 1-wave sweep give the
    same trajectory, and the network runs for fewer trials than the first
    screen passes."""
    say("case: %s, network screen at the default entry threshold, speculative / 4 / 1 waves" % name)
    out = []
    for nw in (0, 4, 1):
        g = gpu_system(prepared, name, 1)
        g.set_network_min_entries(-1)
        g.set_visit_factor(8)
        g.set_block_waves(nw)
        s = g.sweep(10, 1.0, 0.3, ROT_HALF, r=1, seed=9, counter=0)
        out.append((s, g.get_config()))
        g.close()
    s0 = out[0][0]
    say("first screen passed %d of %d, network ran for %d (passed %d)" % (s0["stage1_acc"], s0["stage1_try"],
                                                                        s0["stage2_try"], s0["stage2_acc"]))
    assert s0["stage2_try"] < s0["stage1_acc"]
    for s1, c1 in out[1:]:
        for k in ("trans_acc", "trans_try", "rot_acc", "rot_try", "left_cell", "stage1_acc", "stage1_try",
                  "stage2_acc", "stage2_try"):
            assert s1[k] == s0[k], k
        assert abs(s1["energy_change"] - s0["energy_change"]) <= 1e-9 * max(1.0, abs(s0["energy_change"]))
        for a, b in zip(out[0][1], c1):
            np.testing.assert_array_equal(a, b)


def _blocked(x, nb=8):
    b = np.asarray(x)[: len(x) // nb * nb].reshape(nb, -1).mean(1)
    return b.mean(), b.std(ddof=1) / np.sqrt(nb)


@pytest.fixture(scope="module")
def oracle_ensembles(prepared):
    """name -> [40 x 3] (E/N, S2, mean cluster size) of the oracle's sequential
    Metropolis from the prepared state: 200 sweeps, then 40 blocks of 25."""
    out = {}
    for name in SYSTEMS:
        w, (box, pos, d, q), _ = prepared[name]
        n = len(w["types"])
        tp, st = po.Topo(w["top"]), po.State(box, pos, d, q, w["types"])
        say("oracle ensemble %s: 200 + 40 x 25 sweeps" % name)
        cur, drift, _, _ = po.mc_run(tp, st, 200, 1.0, 0.3, ROT_HALF, seed=77)
        assert abs(drift) <= 1e-9
        rows = []
        for blk in range(40):
            cur, drift, _, ef = po.mc_run(tp, cur, 25, 1.0, 0.3, ROT_HALF, seed=300 + blk)
            assert abs(drift) <= 1e-9
            ncl = len(np.unique(po.cluster_labels(tp, cur)))
            rows.append((ef / n, po.nematic_s2(cur.dir), n / ncl))
        out[name] = np.array(rows)
    return out


@pytest.mark.parametrize("name,mode", [("tchpsc", 0), ("tchpsc", 1), ("tchpsc", 2), ("cpsc", 0), ("cpsc", 1)])
def test_sweep_ensemble_matches_cpu_oracle(prepared, oracle_ensembles, name, mode):
    say("case: %s ensemble, surrogate mode %d, vs the oracle" % (name, mode))
    w = prepared[name][0]
    n = len(w["types"])
    g = gpu_system(prepared, name, mode)
    g.sweep(200, 1.0, 0.3, ROT_HALF, r=1, seed=8, counter=0)
    rows = []
    for blk in range(40):
        g.sweep(25, 1.0, 0.3, ROT_HALF, r=1, seed=8, counter=1000 + 25 * blk)
        _, ncl = g.cluster_labels()
        rows.append((g.allToAll() / n, g.nematic_order()[0], n / ncl))
    g.close()
    rows = np.array(rows)
    ec = oracle_ensembles[name]
    for c, label in enumerate(("E/N", "S2", "mean cluster size")):
        mg, sg = _blocked(rows[:, c])
        mc, scc = _blocked(ec[:, c])
        say("%s mode %d %s: gpu %.5f +- %.5f, oracle %.5f +- %.5f" % (name, mode, label, mg, sg, mc, scc))
        margin = 0.03 if label == "S2" else 0.01 * abs(mc)
        assert abs(mg - mc) <= 3.0 * np.hypot(sg, scc) + margin, label
