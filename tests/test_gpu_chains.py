"""Bonded chains on the GPU (through the C ABI) against the FP64 oracle and
the reference's chain golden vectors.

Tolerances as for atomic systems: pair energies 1e-9 relative in the FP64
device mode (bond terms are always FP64; the axes are stored in FP32 and the
oracle reads the same stored values), 1e-4 relative / 1e-5 absolute in FP32;
whole-system energies 1e-5 relative; per-particle energies 1e-5.  The sweep
keeps bonds exact (FP64) and rejects a move whose bonded partner lies outside
the staged stencil; its ensemble is compared with the oracle's sequential
Metropolis on the same chains.
"""
import os

import numpy as np
import pytest

import sc_amd
from sc_amd import surrogate
from sc_amd import workloads as W
import pyoracle as po

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
SETS = ("c5", "angles", "bondd", "bondh", "sp")
ROT_HALF = 20.0 / 180 * np.pi / 4


def say(*a):
    print(*a, flush=True)


def dev_state(g, types):
    """The device's state for the oracle.  Axes are stored in FP32; they are
    renormalised in FP64 as the device does when it reads them (the
    reference's axes are unit vectors, inicializer.cpp:209-225, and a sphere's
    axis places its bond point for bondh)."""
    b, p, d, q = g.get_config()
    d = d / np.linalg.norm(d, axis=1)[:, None]
    return po.State(b, p, d, q, types)


def golden(g, name):
    z = np.load(os.path.join(GOLD, "chain_golden.npz"))
    top = str(z[name + "_top"])
    st0 = po.State.from_real(z[name + "_box"], z[name + "_pos"], z[name + "_dir"], z[name + "_pd0"],
                             z[name + "_types"])
    g.load_topology(top)
    g.set_config(st0.box, st0.pos, st0.dir, st0.pd0, st0.types)
    return z, po.Topo(top), dev_state(g, st0.types)


@pytest.mark.parametrize("name", SETS)
def test_chain_pair_energies(name):
    say("case: chain pair energies", name)
    g = sc_amd.GpuSystem(0)
    z, tp, st = golden(g, name)
    i, j = z[name + "_i"], z[name + "_j"]
    eo = po.pair_energies(tp, st, i, j)
    bo = po.bond_energies(tp, st, i, j)
    bonded = bo != 0.0
    e64 = g.p2p(i, j, fp64=True)
    rel = np.abs(e64 - eo) / np.maximum(1.0, np.abs(eo))
    say("fp64: %d pairs (%d bonded), max rel %.3e, bonded max rel %.3e" % (
        len(i), bonded.sum(), rel.max(), rel[bonded].max()))
    assert bonded.sum() >= 20
    assert np.all(rel[bonded] <= 1e-9)
    assert np.mean(rel <= 1e-9) > 0.999
    # the reference's own values (positions quantised to the device's fixed point)
    er = z[name + "_e"]
    assert np.mean(np.abs(e64 - er) <= 1e-6 * np.maximum(1.0, np.abs(er))) > 0.995
    e32 = g.p2p(i, j)
    err = np.abs(e32 - eo)
    ok = (err <= 1e-4 * np.abs(eo)) | (err <= 1e-5)
    say("fp32: %d outside 1e-4 rel / 1e-5 abs" % (~ok).sum())
    assert ok.mean() > 0.99
    assert np.all(ok[bonded] | (err[bonded] <= 1e-4 * np.maximum(1.0, np.abs(eo[bonded]))))
    g.close()


@pytest.mark.parametrize("name", SETS)
def test_chain_energy_all_and_per_particle(name):
    say("case: chain energy_all / oneToAll", name)
    g = sc_amd.GpuSystem(0)
    z, tp, st = golden(g, name)
    eg = g.allToAll()
    eo = po.energy_all(tp, st, use_cells=False)
    say("E_all gpu %.10f oracle %.10f reference %.10f" % (eg, eo, float(z[name + "_eall"])))
    assert abs(eg - eo) <= 1e-5 * max(1.0, abs(eo))
    pp = g.energy_per_particle()
    po_ = po.energy_one_each(tp, st, np.arange(st.n))
    say("per particle: max rel %.2e" % (np.abs(pp - po_) / np.maximum(np.abs(po_), 1e-3)).max())
    np.testing.assert_allclose(pp, po_, rtol=1e-5, atol=1e-5)   # 1e-5 kT where terms cancel
    for t in (0, 1, st.n // 2, st.n - 1):
        assert g.oneToAll(t) == pytest.approx(po_[t], rel=1e-5, abs=1e-5)
    g.close()


def chain_system(g, nchains, eta, seed):
    top = W.chain_topology(nchains)
    box, pos, d, q, t = W.chain_lattice(nchains, eta, seed)
    g.load_topology(top)
    g.set_config(box, pos, d, q, t)
    return po.Topo(top), dev_state(g, t)


def test_chain_trial_commit_and_volume():
    """oneToAllTrial / update / allToAll on chain beads, and the whole-system
    energy of a trial box (volume moves scale every position, bonds included)."""
    say("case: chain trial/commit protocol, trial box")
    g = sc_amd.GpuSystem(0)
    tp, st = chain_system(g, 200, 0.2, 4)
    rng = np.random.default_rng(3)
    for t in (0, 1, 2, 301, 599):
        e_old = g.oneToAll(t)
        p = st.pos[t] + rng.normal(size=3) * 0.25 / st.box
        d = st.dir[t] + rng.normal(size=3) * 0.2
        d /= np.linalg.norm(d)
        q = np.cross(d, rng.normal(size=3))
        q /= np.linalg.norm(q)
        e_new = g.oneToAllTrial(t, p, d, q)
        st2 = po.State(st.box, st.pos.copy(), st.dir.copy(), st.pd0.copy(), st.types)
        st2.pos[t], st2.dir[t], st2.pd0[t] = p, d, q
        want = po.energy_one_each(tp, st2, [t])[0]
        say("bead %d: E_old %.6f, trial %.6f, oracle %.6f" % (t, e_old, e_new, want))
        assert e_new == pytest.approx(want, rel=1e-4, abs=1e-4)
        e0 = g.allToAll()
        g.update_particle(t, p, d, q)
        e1 = g.allToAll()
        assert e1 - e0 == pytest.approx(e_new - e_old, rel=1e-4, abs=1e-4)
        st = dev_state(g, st.types)
    for s in (0.97, 1.03):
        nb = st.box * s
        eg = g.allToAllTrial(nb)
        eo = po.energy_all(tp, po.State(nb, st.pos, st.dir, st.pd0, st.types), use_cells=True)
        say("trial box x%.2f: gpu %.8f oracle %.8f" % (s, eg, eo))
        assert abs(eg - eo) <= 1e-5 * max(1.0, abs(eo))
    g.close()


def test_chain_sweep_drift_and_volume_moves():
    """Energy bookkeeping of the chain sweep (E_end - E_start = sum of
    accepted dE) and of NPT volume moves; bond rejections are counted."""
    say("case: chain sweep drift + volume moves")
    g = sc_amd.GpuSystem(0)
    tp, st = chain_system(g, 400, 0.2, 6)
    e0 = g.allToAll()
    s = g.sweep(30, 1.1, 0.3, ROT_HALF, r=2, seed=7, counter=0)
    e1 = g.allToAll()
    drift = e1 - e0 - s["energy_change"]
    say("sweep: E %.6f -> %.6f, sum dE %.6f, drift %.3e, acc %d/%d %d/%d, left %d, bond-rejected %d" % (
        e0, e1, s["energy_change"], drift, s["trans_acc"], s["trans_try"], s["rot_acc"], s["rot_try"],
        s["left_cell"], s["bond_rejected"]))
    assert abs(drift) <= 1e-8 * max(1.0, abs(e1))
    assert s["trans_acc"] > 0 and s["rot_acc"] > 0
    acc = 0
    for k in range(20):
        a, box = g.volume_move(1, 0.5, 1.1, 2.0, seed=9, counter=k)
        acc += a
        g.sweep(1, 1.1, 0.3, ROT_HALF, r=2, seed=7, counter=100 + k)
    stg = dev_state(g, st.types)
    eo = po.energy_all(tp, stg, use_cells=True)
    eg = g.allToAll()
    b = stg.box
    say("after %d/20 volume moves: box %.4f, E gpu %.8f oracle %.8f" % (acc, b[0], eg, eo))
    assert abs(eg - eo) <= 1e-5 * max(1.0, abs(eo))
    assert acc > 0
    g.close()


@pytest.mark.parametrize("r", [1, 2])
def test_speculative_chain_sweep_reproduces_sequential(r):
    """The speculative sweep (a trial per wave, kept in order up to the first
    dependence on an accepted trial, chain partners always dependent) gives
    the one-trial-at-a-time trajectory of a bonded system bit for bit."""
    say("case: chain sweep, speculative vs one trial at a time, r =", r)
    out = []
    for nw in (0, 1):
        g = sc_amd.GpuSystem(0)
        chain_system(g, 400, 0.2, 6)
        g.set_trial_mode(1)
        g.set_visit_factor(8 if r == 1 else 9)
        g.set_block_waves(nw)
        s = g.sweep(10, 1.1, 0.3, ROT_HALF, r=r, seed=17, counter=0)
        out.append((s, g.get_config()))
        g.close()
    (s1, c1), (s2, c2) = out
    say("accepted %d / %d, bond-rejected %d" % (s1["trans_acc"] + s1["rot_acc"], s1["trans_try"] + s1["rot_try"],
                                                s1["bond_rejected"]))
    for k in ("trans_acc", "trans_try", "rot_acc", "rot_try", "left_cell", "bond_rejected"):
        assert s1[k] == s2[k], k
    assert abs(s1["energy_change"] - s2["energy_change"]) <= 1e-9 * max(1.0, abs(s1["energy_change"]))
    for x, y in zip(c1, c2):
        np.testing.assert_array_equal(x, y)


def _blocked(x, nb=8):
    b = x[: len(x) // nb * nb].reshape(nb, -1).mean(1)
    return b.mean(), b.std(ddof=1) / np.sqrt(nb)


def _bond_energy_per_bond(tp, st):
    n = st.n
    i = np.arange(1, n)
    nxt = tp.chain[i, 1] > 0
    return po.bond_energies(tp, st, i[nxt], i[nxt] - 1).mean()


@pytest.fixture(scope="module")
def chain_nvt_oracle():
    """60 chains (180 beads) at eta 0.15, T* = 1.1: the oracle's sequential
    Metropolis after 400 sweeps, then E/N, the mean bond energy and the mean
    cluster size at the end of 40 blocks of 25 sweeps."""
    nch = 60
    top = W.chain_topology(nch)
    tp = po.Topo(top)
    box, pos, d, q, t = W.chain_lattice(nch, 0.15, 11)
    st = po.State(box, pos, d, q, t)
    cur, _, _, _ = po.mc_run(tp, st, 400, 1.1, 0.3, ROT_HALF, seed=21)
    ec = []
    for blk in range(40):
        cur, _, _, ef = po.mc_run(tp, cur, 25, 1.1, 0.3, ROT_HALF, seed=200 + blk)
        ec.append((ef / st.n, _bond_energy_per_bond(tp, cur), st.n / len(np.unique(po.cluster_labels(tp, cur)))))
    return nch, (box, pos, d, q, t), np.array(ec)


@pytest.mark.parametrize("mode", [0, 2, 1])
def test_chain_ensemble_matches_cpu_oracle(chain_nvt_oracle, mode):
    """E/N, the mean bond energy and the mean cluster size (molecules +
    attraction) of the GPU checkerboard sweep -- exact (0), with the
    cheap-terms screen (2) or with the network screen trained on the tails'
    PSC pair, heads screening with 0 (1) -- and of the oracle's sequential
    Metropolis agree within 3 combined standard errors + 1 %."""
    say("case: chain ensemble GPU vs CPU oracle, surrogate mode", mode)
    nch, (box, pos, d, q, t), ec = chain_nvt_oracle
    g = sc_amd.GpuSystem(0)
    tp = po.Topo(W.chain_topology(nch))
    g.load_topology(W.chain_topology(nch))
    if mode == 1:
        g.set_config(box, pos, d, q, t)
        res = surrogate.train(g, 2, nsamples=1_000_000, epochs=4, seed=3)
        say("network: held-out rmse %.4f (target std %.4f)" % (res["val_rmse"], res["target_std"]))
    g.set_config(box, pos, d, q, t)
    g.set_surrogate(mode)
    g.set_network_min_entries(0)   # the network screens every trial that passes the first screen
    n = len(t)
    g.sweep(400, 1.1, 0.3, ROT_HALF, r=2, seed=5, counter=0)
    obs = []
    tot = {}
    for blk in range(40):
        s = g.sweep(25, 1.1, 0.3, ROT_HALF, r=2, seed=5, counter=1000 + 25 * blk)
        for k, v in s.items():
            tot[k] = tot.get(k, 0) + v
        stg = dev_state(g, t)
        _, ncl = g.cluster_labels()
        obs.append((g.allToAll() / n, _bond_energy_per_bond(tp, stg), n / ncl))
    obs = np.array(obs)
    say("bond-rejected moves %d; screens %d/%d, %d/%d" % (tot["bond_rejected"], tot["stage1_acc"], tot["stage1_try"],
                                                          tot["stage2_acc"], tot["stage2_try"]))
    if mode:
        assert tot["stage1_try"] > 0 and tot["stage1_acc"] < tot["stage1_try"]
    if mode == 1:
        assert tot["stage2_try"] > 0
    for c, name in enumerate(("E/N", "bond energy", "mean cluster size")):
        mg, sg = _blocked(obs[:, c])
        mc, scc = _blocked(ec[:, c])
        say("%s: gpu %.5f +- %.5f, cpu %.5f +- %.5f" % (name, mg, sg, mc, scc))
        assert abs(mg - mc) <= 3.0 * np.hypot(sg, scc) + 0.01 * abs(mc), name
    g.close()


def test_chain_cluster_labels_match_oracle():
    """Molecules are clusters by themselves; attraction joins them."""
    say("case: chain cluster labels")
    g = sc_amd.GpuSystem(0)
    tp, st = chain_system(g, 300, 0.25, 13)
    g.sweep(40, 0.7, 0.3, ROT_HALF, r=2, seed=3, counter=0)
    lab, ncl = g.cluster_labels()
    ref = po.cluster_labels(tp, dev_state(g, st.types))
    say("clusters: gpu %d, oracle %d" % (ncl, len(np.unique(ref))))
    np.testing.assert_array_equal(lab, ref)
    assert ncl <= 300
    g.close()


def test_batches_refuse_chains():
    g = sc_amd.GpuSystem(0)
    chain_system(g, 10, 0.1, 1)
    with pytest.raises(RuntimeError):
        g.batch_set(np.array([15, 15], np.int32), np.full((2, 3), 20.0))
    g.close()


def test_chain_npt_ensemble_matches_cpu_oracle():
    """NPT (P* = 0.5, T* = 1.1, one isotropic volume move per sweep) of 60
    chains: packing fraction and E/N of the GPU sweep + device volume moves
    against the oracle's sequential Metropolis + pressureMove, both continuing
    from one oracle-compressed state: equal within 3 combined standard errors
    + 1 %."""
    say("case: chain NPT ensemble GPU vs CPU oracle")
    nch, press, temper, edge = 60, 0.5, 1.1, 0.3
    top = W.chain_topology(nch)
    tp = po.Topo(top)
    box, pos, d, q, t = W.chain_lattice(nch, 0.15, 17)
    st, _, _, _ = po.mc_run_npt(tp, po.State(box, pos, d, q, t), 1500, temper, 0.3, ROT_HALF, 5, 1, press, edge)
    vchain = nch * W.chain_volume()
    n = st.n
    # oracle
    cur, drift, _, _ = po.mc_run_npt(tp, st, 400, temper, 0.3, ROT_HALF, 7, 1, press, edge)
    assert abs(drift) < 1e-8
    ec = []
    vacc = [0, 0]
    for blk in range(40):
        cur, _, s, ef = po.mc_run_npt(tp, cur, 25, temper, 0.3, ROT_HALF, 300 + blk, 1, press, edge)
        vacc[0] += s[4]
        vacc[1] += s[5]
        ec.append((vchain / np.prod(cur.box), ef / n))
    ec = np.array(ec)
    # device
    g = sc_amd.GpuSystem(0)
    g.load_topology(top)
    g.set_config(st.box, st.pos, st.dir, st.pd0, st.types)
    gacc = 0
    obs = []
    for sw in range(1400):
        g.sweep(1, temper, 0.3, ROT_HALF, r=2, seed=11, counter=sw)
        a, b = g.volume_move(1, press, temper, edge, seed=12, counter=sw)
        if sw >= 400:
            gacc += a
            if (sw - 400) % 25 == 24:
                obs.append((vchain / np.prod(b), g.allToAll() / n))
    obs = np.array(obs)
    say("volume acceptance: gpu %.3f, cpu %.3f" % (gacc / 1000.0, vacc[0] / max(1, vacc[1])))
    for c, name in enumerate(("packing fraction", "E/N")):
        mg, sg = _blocked(obs[:, c])
        mc, scc = _blocked(ec[:, c])
        say("%s: gpu %.5f +- %.5f, cpu %.5f +- %.5f" % (name, mg, sg, mc, scc))
        assert abs(mg - mc) <= 3.0 * np.hypot(sg, scc) + 0.01 * abs(mc), name
    g.close()
