"""The benchmark configurations at their full size, through the C ABI.

Sizes the oracle cannot sample are checked by properties that hold at any
size, and by the oracle's energy of the GPU's own configurations:

  * C4, 65,536 TCHPSC at packing fraction 0.25 (the delayed-acceptance
    benchmark, bench settings: 2x2x2 checkerboard, visit factor 8): energy
    bookkeeping of the exact sweep and of both delayed-acceptance modes
    (allToAll(end) - allToAll(start) = sum of accepted dE within 1e-8 |E|),
    allToAll against the FP64 oracle (1e-5), and the network-screened
    sweep's ensemble (E/N, S2, mean cluster size) against the exact sweep's
    from one equilibrated state, each run first checked stationary;
  * C3, 64 replicas of 500-8,000 PSC/CPSC (253,424 particles): per-replica
    energy bookkeeping of one batch sweep, a replica's energy against the
    oracle, the exchange step's bookkeeping, and the smallest replica's E/N
    against the oracle's sequential Metropolis at its temperature;
  * C5, 43,690 bonded chains (131,070 beads) under NPT (P* 0.5, T* 1.1, a
    volume move after every sweep): the energy the volume moves carry
    through sweeps and volume moves against a full evaluation (1e-8), and
    that against the oracle (1e-5).

Reference: Updater::simulate's drift bookkeeping (updater.cpp:122-150),
MoveCreator::pressureMove (movecreator.cpp:330-421), replicaExchangeMove
(movecreator.cpp:552-795).
"""
import numpy as np
import pytest

import sc_amd
from sc_amd import prepare
from sc_amd import pt
from sc_amd import surrogate
from sc_amd import workloads as W
import pyoracle as po

pytestmark = pytest.mark.gpu
ROT_HALF = 20.0 / 180 * np.pi / 4


def say(*a):
    print(*a, flush=True)


def _blocked(x, nb=8):
    b = np.asarray(x)[: len(x) // nb * nb].reshape(nb, -1).mean(1)
    return b.mean(), b.std(ddof=1) / np.sqrt(nb)


# ------------------------------------------------------------------- C4
def test_c4_speculative_sweep_reproduces_sequential_from_compression():
    """The speculative sweep (a trial per wave, kept up to the first that
    depends on an accepted trial of its round) against one trial at a time,
    on the full C4 system from its RSA start: the compression that prepares
    it (r = 2, cells of a few particles, overlapping rods: rounds often try
    one particle twice) and 200 sweeps at the bench's r = 1, visit factor 8.
    The same counters and the same configuration bit for bit."""
    say("case: C4 65,536 TCHPSC, compression + 200 sweeps, speculative vs one trial at a time")
    w = W.workload("c4_tchpsc_65536")
    out = []
    for nw in (0, 1):
        g = sc_amd.GpuSystem(0)
        g.set_block_waves(nw)
        prepare.initial_state(g, w, seed=2025)
        g.set_trial_mode(1)
        g.set_visit_factor(8)
        s = g.sweep(200, 1.0, 0.3, ROT_HALF, r=1, seed=5, counter=10 ** 6)
        out.append((s, g.get_config(), g.allToAll()))
        g.close()
    (s1, c1, e1), (s2, c2, e2) = out
    say("E/N %.6f vs %.6f, accepted %d vs %d" % (e1 / len(w["types"]), e2 / len(w["types"]),
                                               s1["trans_acc"] + s1["rot_acc"], s2["trans_acc"] + s2["rot_acc"]))
    for k in ("trans_acc", "trans_try", "rot_acc", "rot_try", "left_cell"):
        assert s1[k] == s2[k], k
    for a, b in zip(c1, c2):
        np.testing.assert_array_equal(a, b)


@pytest.fixture(scope="module")
def c4():
    say("prepare C4: 65,536 TCHPSC, eta 0.25, 6000 exact sweeps")
    w = W.workload("c4_tchpsc_65536")
    g = sc_amd.GpuSystem(0)
    prepare.initial_state(g, w, seed=2025)
    g.set_trial_mode(1)
    g.set_visit_factor(8)
    g.sweep(6000, 1.0, 0.3, ROT_HALF, r=1, seed=5, counter=10 ** 6)
    state = g.get_config()
    surrogate.train(g, 3, nsamples=4_000_000, epochs=6, seed=3, log=say)
    wts = g.mlp_get_weights()
    say("C4 E/N %.5f" % (g.allToAll() / len(w["types"])))
    g.close()
    return w, state, wts


def c4_system(c4, mode):
    w, (box, pos, d, q), wts = c4
    g = sc_amd.GpuSystem(0)
    g.load_topology(w["top"])
    if mode == 1:
        g.mlp_init(3, seed=1)
        g.mlp_set_weights(wts)
    g.set_config(box, pos, d, q, w["types"])
    g.set_trial_mode(1)
    g.set_visit_factor(8)
    g.set_surrogate(mode)
    return g


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_c4_energy_bookkeeping_and_oracle_energy(c4, mode):
    say("case: C4 65,536 TCHPSC, surrogate mode %d: bookkeeping over 100 sweeps, allToAll vs oracle" % mode)
    w = c4[0]
    n = len(w["types"])
    g = c4_system(c4, mode)
    e0 = g.allToAll()
    s = g.sweep(100, 1.0, 0.3, ROT_HALF, r=1, seed=11 + mode, counter=0)
    e1 = g.allToAll()
    drift = e1 - e0 - s["energy_change"]
    tries = s["trans_try"] + s["rot_try"]
    say("E %.5f -> %.5f, sum dE %.5f, drift %.3e (%.2e of |E|), %.4f N trials per sweep, acceptance %.4f" % (
        e0, e1, s["energy_change"], drift, abs(drift) / abs(e1), tries / (100.0 * n),
        (s["trans_acc"] + s["rot_acc"]) / tries))
    assert abs(drift) <= 1e-8 * abs(e1)
    assert abs(tries / (100.0 * n) - 1.0) < 0.01
    b, p, dd, qq = g.get_config()
    g.close()
    eo = po.energy_all(po.Topo(w["top"]), po.State(b, p, dd, qq, w["types"]), use_cells=True)
    say("allToAll gpu %.6f oracle %.6f rel %.2e" % (e1, eo, abs(e1 - eo) / abs(eo)))
    assert abs(e1 - eo) <= 1e-5 * abs(eo)


def test_c4_network_screen_ensemble_matches_exact(c4):
    """E/N, S2 and mean cluster size of the network-screened sweep equal the
    exact sweep's (3 combined block errors + 1%; S2 + 0.03 absolute) from the
    same equilibrated state, over 8,000 sweeps each (E/N of this system
    fluctuates on a scale of ~1,000 sweeps); each run is stationary over its
    window (halves agree within 3 errors + 1%)."""
    say("case: C4 ensemble, network screen vs exact, 8,000 sweeps each")
    w = c4[0]
    n = len(w["types"])
    res = {}
    for mode in (0, 1):
        g = c4_system(c4, mode)
        g.sweep(500, 1.0, 0.3, ROT_HALF, r=1, seed=31, counter=0)
        rows = []
        for blk in range(32):
            g.sweep(250, 1.0, 0.3, ROT_HALF, r=1, seed=31, counter=1000 + 250 * blk)
            _, ncl = g.cluster_labels()
            rows.append((g.allToAll() / n, g.nematic_order()[0], n / ncl))
        g.close()
        rows = np.array(rows)
        for c, label in enumerate(("E/N", "S2", "cluster size")):
            m1, s1 = _blocked(rows[:16, c], 4)
            m2, s2 = _blocked(rows[16:, c], 4)
            say("mode %d %s: first half %.5f +- %.5f, second half %.5f +- %.5f" % (mode, label, m1, s1, m2, s2))
            margin = 0.03 if label == "S2" else 0.01 * abs(m1)
            assert abs(m1 - m2) <= 3.0 * np.hypot(s1, s2) + margin, ("not stationary", mode, label)
        res[mode] = rows
    for c, label in enumerate(("E/N", "S2", "cluster size")):
        me, se = _blocked(res[0][:, c])
        md, sd = _blocked(res[1][:, c])
        say("%s: exact %.5f +- %.5f, network screen %.5f +- %.5f" % (label, me, se, md, sd))
        margin = 0.03 if label == "S2" else 0.01 * abs(me)
        assert abs(md - me) <= 3.0 * np.hypot(se, sd) + margin, label


# ------------------------------------------------------------------- C3
@pytest.fixture(scope="module")
def c3():
    say("prepare C3: 64 replicas, N 500-8000, PSC+CPSC, eta 0.15-0.35")
    g = sc_amd.GpuSystem(0)
    n, temps, boxes = pt.c3_batch(g)
    g.set_trial_mode(1)
    yield g, n, temps, boxes
    g.close()


def _replica_state(g, k, n, boxes):
    p, d, q = g.batch_get_config(k)
    t = np.array([1] * (n[k] // 2) + [2] * (n[k] - n[k] // 2), np.int32)
    return po.State(boxes[k], p, d, q, t)


def test_c3_replica_bookkeeping_energy_and_exchanges(c3):
    g, n, temps, boxes = c3
    say("case: C3 %d replicas / %d particles: per-replica drift, oracle energy, exchanges" % (len(n), n.sum()))
    e0 = g.batch_energies()
    st = g.batch_sweep(10, temps, 0.3, ROT_HALF, r=1, seed=41, counter=0)
    e1 = g.batch_energies()
    drift = np.array([e1[k] - e0[k] - st[k]["energy_change"] for k in range(len(n))])
    rel = np.abs(drift) / np.maximum(1.0, np.abs(e1))
    tries = np.array([s["trans_try"] + s["rot_try"] for s in st], float)
    say("max drift %.3e (replica %d, %.2e of |E|); trials per sweep / N: %.4f .. %.4f" % (
        np.abs(drift).max(), int(np.argmax(rel)), rel.max(), (tries / (10.0 * n)).min(), (tries / (10.0 * n)).max()))
    assert rel.max() <= 1e-8
    assert np.all(np.abs(tries / (10.0 * n) - 1.0) < 0.1)
    tp = po.Topo(pt.c3_topology(int(n.sum())))
    for k in (int(np.argmin(n)), int(np.argmax(n))):
        eo = po.energy_all(tp, _replica_state(g, k, n, boxes), use_cells=True)
        say("replica %d (N %d, T %.3f): E gpu %.6f oracle %.6f" % (k, n[k], temps[k], e1[k], eo))
        assert abs(e1[k] - eo) <= 1e-5 * max(1.0, abs(eo))
    # exchange bookkeeping: temperatures stay a permutation of the ladder, the
    # attempted pairs alternate parity, acceptances within the attempts
    t2, tot, acc, tried = pt.run(g, temps, 40, exchange_every=10, seed=43, counter=100)
    say("exchanges: %d/%d accepted; per pair attempts %s" % (acc.sum(), tried.sum(), tried.tolist()))
    np.testing.assert_allclose(np.sort(t2), np.sort(temps), rtol=0, atol=0)
    assert np.all(acc <= tried) and tried.sum() == 2 * (len(n) // 2) + 2 * ((len(n) - 1) // 2)
    assert 0 < acc.sum() < tried.sum()
    e3 = g.batch_energies()
    assert np.all(np.isfinite(e3))


def test_c3_smallest_replica_matches_oracle(c3):
    """The smallest replica of the batch, swept with all the others (no
    exchanges), samples its temperature's ensemble: E/N equals the oracle's
    sequential Metropolis at that temperature from the same state within 3
    combined block errors + 1%."""
    g, n, temps, boxes = c3
    k = int(np.argmin(n))
    say("case: C3 replica %d (N %d, T %.3f) E/N vs the oracle" % (k, n[k], temps[k]))
    tp = po.Topo(pt.c3_topology(int(n.sum())))
    st = _replica_state(g, k, n, boxes)
    cur, _, _, _ = po.mc_run(tp, st, 200, temps[k], 0.3, ROT_HALF, seed=7)
    ref = []
    for blk in range(40):
        cur, _, _, ef = po.mc_run(tp, cur, 25, temps[k], 0.3, ROT_HALF, seed=60 + blk)
        ref.append(ef / n[k])
    g.batch_sweep(200, temps, 0.3, ROT_HALF, r=1, seed=51, counter=10 ** 5)
    gpu = []
    for blk in range(40):
        g.batch_sweep(25, temps, 0.3, ROT_HALF, r=1, seed=51, counter=2 * 10 ** 5 + 25 * blk)
        gpu.append(g.batch_energies()[k] / n[k])
    mg, sg = _blocked(gpu)
    mo, so = _blocked(ref)
    say("E/N batch %.5f +- %.5f, oracle %.5f +- %.5f" % (mg, sg, mo, so))
    assert abs(mg - mo) <= 3.0 * np.hypot(sg, so) + 0.01 * abs(mo)


# ------------------------------------------------------------------- C5
def test_c5_npt_carried_energy_and_oracle():
    say("case: C5 43,690 chains (131,070 beads), NPT P* 0.5 T* 1.1, 40 steps of sweep + volume move")
    nch = 43690
    top = W.chain_topology(nch)
    box, pos, d, q, t = W.chain_lattice(nch, 0.25, 2025)
    g = sc_amd.GpuSystem(0)
    g.load_topology(top)
    g.set_config(box, pos, d, q, t)
    g.set_trial_mode(1)
    acc = 0
    worst, longest, age = 0.0, 0, 0
    since_full = 0   # sweeps and volume moves the carried energy has been carried through
    for k in range(40):
        g.sweep(1, 1.1, 0.3, ROT_HALF, r=1, seed=61, counter=k)
        if k > 0:   # carried through this sweep since the last volume move
            ec, age = g.carried_energy()
            ea = g.allToAll()
            since_full += 1
            worst = max(worst, abs(ec - ea) / abs(ea))
            longest = max(longest, since_full)
        a, bx = g.volume_move(1, 0.5, 1.1, 0.1, seed=62, counter=k)
        acc += a
        ec, age = g.carried_energy()
        if age == 0:   # a full evaluation (refresh or accepted move) resets the carry
            since_full = 0
    ea = g.allToAll()
    say("carried vs full: worst %.3e of |E| over carries of up to %d sweeps + volume moves; volume moves "
        "accepted %d/40, box %.4f, last age %d" % (worst, longest, acc, bx[0], age))
    assert longest >= 2 and worst <= 1e-8
    assert 0 < acc < 40
    b, p, dd, qq = g.get_config()
    g.close()
    eo = po.energy_all(po.Topo(top), po.State(b, p, dd, qq, t), use_cells=True)
    say("allToAll gpu %.6f oracle %.6f rel %.2e" % (ea, eo, abs(ea - eo) / abs(eo)))
    assert abs(ea - eo) <= 1e-5 * abs(eo)
