"""Replica batches (parallel tempering) and the autotuner on the GPU.

  * a batch sweep, speculative or one trial at a time, is bit-identical,
    replica by replica, to sweeping each replica alone with its replica seed
    (positions, axes, patch directions, acceptance counts and the FP64 energy
    change);
  * batch energies equal the single-system whole-system energies of the same
    configurations (1e-6 relative: loading a configuration renormalises the
    axes in FP64 and rounds them to FP32 again, which moves their last bits);
  * the exchange step follows min(1, exp((1/T_k - 1/T_k+1)(E_k - E_k+1))):
    always for change > 0, with frequency exp(change) (within 3 binomial
    standard errors) otherwise;
  * the autotuner restores the state it times on.
"""
import numpy as np
import pytest

import sc_amd
from sc_amd import pt
from sc_amd import workloads as W

pytestmark = pytest.mark.gpu
ROT_HALF = 20.0 / 180 * np.pi / 4


def say(*a):
    print(*a, flush=True)


def small_batch(g, sizes=(700, 1300, 400), etas=(0.18, 0.2, 0.15), seed=5):
    """three PSC/CPSC replicas of uneven size, loaded into g"""
    total = int(sum(sizes))
    g.load_topology(pt.c3_topology(total))
    v = W.particle_volume(1.0, 3.0)
    pos, d, q, ty, boxes = [], [], [], [], []
    for k, (n, e) in enumerate(zip(sizes, etas)):
        b = (n * v / e) ** (1.0 / 3.0)
        t = np.array([1] * (n // 2) + [2] * (n - n // 2), np.int32)
        p_, d_, q_ = g.generate_rsa(t, np.full(3, b), seed + k)
        pos.append(p_), d.append(d_), q.append(q_), ty.append(t), boxes.append([b, b, b])
    g.set_config(np.array(boxes[0]), np.concatenate(pos), np.concatenate(d), np.concatenate(q), np.concatenate(ty))
    g.batch_set(np.array(sizes, np.int32), np.array(boxes))
    return (pos, d, q, ty), np.array(boxes), np.array(sizes)


@pytest.mark.parametrize("nw,r,visit", [(0, 2, 1), (1, 2, 1), (0, 1, 8)])
def test_batch_sweep_bit_identical_to_single_replicas(nw, r, visit):
    """nw 0: the batch's speculative sweep; 1: one trial at a time (the single
    systems sweep speculatively); visit 8: a sweep of one colour with 8x the
    trials per cell, in the batch and alone."""
    say("case: batch sweep (block waves %d, r %d, visit %d) vs single-replica sweeps, 3 replicas" % (nw, r, visit))
    g = sc_amd.GpuSystem(0)
    (pos, d, q, ty), boxes, sizes = small_batch(g)
    temps = np.array([0.9, 1.0, 1.2])
    g.set_trial_mode(1)
    g.set_block_waves(nw)
    g.set_visit_factor(visit)
    st = g.batch_sweep(6, temps, 0.3, ROT_HALF, r=r, seed=77, counter=11)
    top = pt.c3_topology(int(sizes.sum()))
    for k in range(3):
        bp, bd, bq = g.batch_get_config(k)
        s = sc_amd.GpuSystem(0)
        s.load_topology(top)
        s.set_config(boxes[k], pos[k], d[k], q[k], ty[k])
        s.set_trial_mode(1)
        s.set_visit_factor(visit)
        ss = s.sweep(6, temps[k], 0.3, ROT_HALF, r=r, seed=sc_amd.GpuSystem.replica_seed(77, k), counter=11)
        _, sp, sd, sq = s.get_config()
        say("replica %d: batch acc %d/%d, single acc %d/%d, dE %.12f vs %.12f" % (
            k, st[k]["trans_acc"] + st[k]["rot_acc"], st[k]["trans_try"] + st[k]["rot_try"],
            ss["trans_acc"] + ss["rot_acc"], ss["trans_try"] + ss["rot_try"], st[k]["energy_change"],
            ss["energy_change"]))
        for key in ("trans_acc", "trans_try", "rot_acc", "rot_try", "left_cell"):
            assert st[k][key] == ss[key], key
        np.testing.assert_array_equal(bp, sp)
        np.testing.assert_array_equal(bd, sd)
        np.testing.assert_array_equal(bq, sq)
        assert abs(st[k]["energy_change"] - ss["energy_change"]) <= 1e-9 * max(1.0, abs(ss["energy_change"]))
        s.close()
    g.close()


def test_batch_energies_match_single_systems():
    say("case: batch energies vs single systems")
    g = sc_amd.GpuSystem(0)
    (pos, d, q, ty), boxes, sizes = small_batch(g, seed=9)
    g.batch_sweep(3, np.array([1.0, 1.1, 0.9]), 0.3, ROT_HALF, r=2, seed=4, counter=0)
    e = g.batch_energies()
    top = pt.c3_topology(int(sizes.sum()))
    for k in range(3):
        bp, bd, bq = g.batch_get_config(k)
        s = sc_amd.GpuSystem(0)
        s.load_topology(top)
        s.set_config(boxes[k], bp, bd, bq, ty[k])
        ek = s.allToAll()
        say("replica %d: batch %.10f single %.10f" % (k, e[k], ek))
        assert abs(e[k] - ek) <= 1e-6 * max(1.0, abs(ek))
        s.close()
    g.close()


def test_exchange_criterion():
    """Ladder order by temperature; change = (1/Ta - 1/Tb)(Ea - Eb)."""
    say("case: replica exchange criterion")
    g = sc_amd.GpuSystem(0)
    small_batch(g, seed=2)
    temps = np.array([1.2, 0.8, 1.0])          # ladder: replica 1 < replica 2 < replica 0
    # pair (rank 0, rank 1) = replicas (1, 2): change = (1/0.8 - 1/1.0) (E1 - E2)
    e = np.array([0.0, 10.0, 12.8])             # change = 0.25 * (-2.8) = -0.7
    acc = 0
    trials = 4000
    for c in range(trials):
        t2, a = g.batch_exchange(temps, e, 0, seed=3, counter=c)
        assert a[1] == 0                          # parity 0: only ladder pair 0 is tried
        if a[0]:
            acc += 1
            np.testing.assert_array_equal(t2, [1.2, 1.0, 0.8])
        else:
            np.testing.assert_array_equal(t2, temps)
    p = np.exp(-0.7)
    say("exchange acceptance %.4f, expected %.4f" % (acc / trials, p))
    assert abs(acc / trials - p) <= 3 * np.sqrt(p * (1 - p) / trials)
    # a favourable swap is always taken (parity 1: ladder pair 1 = replicas (2, 0))
    e2 = np.array([-5.0, 0.0, 0.0])             # change = (1/1.0 - 1/1.2) (0 - (-5)) > 0
    t3, a3 = g.batch_exchange(temps, e2, 1, seed=3, counter=0)
    assert a3[1] == 1 and a3[0] == 0
    np.testing.assert_array_equal(t3, [1.0, 0.8, 1.2])
    g.close()


def test_autotune_restores_state_and_pt_run():
    say("case: autotuner on a batch, then a short tempering run")
    g = sc_amd.GpuSystem(0)
    small_batch(g, sizes=(900, 1600, 600, 1200), etas=(0.18, 0.2, 0.15, 0.17), seed=12)
    temps = np.array([0.8, 0.95, 1.1, 1.3])
    e0 = g.batch_energies()
    r, m, rates = g.autotune(3, temps, 0.3, ROT_HALF, seed=5)
    say("autotune: r=%d mode=%d rates %s" % (r, m, np.round(rates, 1).tolist()))
    assert r in (1, 2) and m in (0, 1) and np.all(rates > 0)
    np.testing.assert_allclose(g.batch_energies(), e0, rtol=1e-10, atol=1e-10)
    temps2, tot, acc, tried = pt.run(g, temps, 40, exchange_every=10, seed=8)
    say("tempering: temperatures %s, exchanges %s / %s" % (np.round(temps2, 3).tolist(), acc.tolist(),
                                                           tried.tolist()))
    assert sorted(temps2.tolist()) == sorted(temps.tolist())
    assert tried.tolist() == [2, 2, 2]   # 4 exchange steps, even and odd ladder pairs alternating
    for t in tot:
        assert t["trans_try"] + t["rot_try"] > 0
    g.close()


def test_pt_run_with_carried_energies_matches_full_evaluation():
    """The exchange step on each replica's energy carried through the sweeps
    (evaluated in full at the first exchange only) against a full evaluation
    at every exchange, from the same state and seed: the same temperatures,
    exchange counts and configurations, and the carried energies meet the
    full ones to 1e-8 of |E|."""
    say("case: tempering run, carried replica energies vs full evaluation at every exchange")
    out = []
    for refresh in (1, 1000):
        g = sc_amd.GpuSystem(0)
        small_batch(g, sizes=(900, 1600, 600, 1200), etas=(0.18, 0.2, 0.15, 0.17), seed=12)
        temps = np.array([0.8, 0.95, 1.1, 1.3])
        e0 = g.batch_energies()
        temps2, tot, acc, tried = pt.run(g, temps, 80, exchange_every=10, seed=8, r=1, refresh_every=refresh)
        e1 = g.batch_energies()
        carried = e0 + np.array([t["energy_change"] for t in tot])
        say("refresh every %d: temperatures %s, exchanges %s / %s, carried - full %s" % (
            refresh, np.round(temps2, 3).tolist(), acc.tolist(), tried.tolist(), (carried - e1).tolist()))
        np.testing.assert_allclose(carried, e1, rtol=1e-8, atol=1e-8)
        out.append((temps2, acc, tried, [g.batch_get_config(k) for k in range(4)]))
        g.close()
    (ta, aa, ra, ca), (tb, ab, rb, cb) = out
    assert ta.tolist() == tb.tolist() and aa.tolist() == ab.tolist() and ra.tolist() == rb.tolist()
    for x, y in zip(ca, cb):
        for u, v in zip(x, y):
            np.testing.assert_array_equal(u, v)
