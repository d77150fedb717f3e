"""Parity of the HIP path (through the C ABI) with the FP64 oracle.

Tolerances:
  * pair energies, FP64 device mode: identical arithmetic to the oracle up to
    libm rounding -> 1e-9 relative for >= 99.9% of pairs (the rest are
    degenerate-branch pairs, counted and bounded);
  * pair energies, FP32: 1e-5 relative for >= 99.5% of the pairs with
    0.1 <= |E| < 1e3 kT (all within 5e-5), 2e-6 kT absolute below, 1e-3
    relative above;
  * per-particle energies: 1e-5 relative (1e-5 kT where terms cancel);
  * sweep energy bookkeeping: 1e-8 |E| (every kernel rounds a pair term
    alike, so only the order of the FP64 sums differs);
  * whole-system energy (FP32 pair terms, FP64 reduction): 1e-5 relative;
  * cell lists, neighbour lists: bit-exact;
  * overlap flags: identical except pairs within 1e-9 sigma of contact.
"""
import os

import numpy as np
import pytest

import sc_amd
from sc_amd import workloads as W
import pyoracle as po

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def dev():
    return 0


def gpu_state(g, top, box, pos_frac, d, q, types):
    g.load_topology(top)
    g.set_config(box, pos_frac, d, q, types)
    b, p, dd, qq = g.get_config()
    return po.Topo(top), po.State(b, p, dd, qq, types)


def mixed(g):
    z = np.load(os.path.join(GOLD, "pair_golden.npz"))
    st0 = po.State.from_real(z["mixed_box"], z["mixed_pos"], z["mixed_dir"], z["mixed_pd0"], z["mixed_types"])
    tp, st = gpu_state(g, str(z["mixed_top"]), st0.box, st0.pos, st0.dir, st0.pd0, st0.types)
    return z, tp, st


def test_pair_energies_fp64_mode(dev):
    say("case: pair energies fp64 mode, mixed 12-type system")
    g = sc_amd.GpuSystem(dev)
    z, tp, st = mixed(g)
    i, j = z["mixed_i"], z["mixed_j"]
    eg = g.p2p(i, j, fp64=True)
    eo = po.pair_energies(tp, st, i, j)
    rel = np.abs(eg - eo) / np.maximum(1.0, np.abs(eo))
    bad = rel > 1e-9
    say("fp64: %d pairs, %d beyond 1e-9, max rel %.3e" % (len(i), bad.sum(), rel.max()))
    assert bad.mean() < 1e-3
    # the golden reference values themselves (state quantised to the device's
    # fixed point, so a looser bound)
    er = z["mixed_e"]
    rel2 = np.abs(eg - er) / np.maximum(1.0, np.abs(er))
    assert np.mean(rel2 < 1e-6) > 0.995


def test_pair_energies_fp32(dev):
    """FP32 pair terms against the FP64 oracle on the 59,918 mixed-species
    pairs, class by class.  The FP32 error of a pair term is that of its
    separation vector (~2e-7 sigma at the stencil's distances) times the
    term's slope, so the relative error is bounded only where |E'/E| is:
      * 0.1 <= |E| < 1e3 kT: >= 99.5% within 1e-5 relative, every one within
        5e-5 (measured with per-expression FMA contraction: 7 of 2,096
        beyond 1e-5, the largest 2.0e-5; the steepest terms of this class,
        |E'/E| ~ 100 per sigma, reach 2e-5 from the separation alone);
      * |E| < 0.1 kT (the WCA cutoff and the end of the attraction's cos^2
        switch, where E -> 0 with a finite or zero slope): 2e-6 kT absolute;
      * |E| >= 1e3 kT (segments closer than 0.56 sigma; Boltzmann weight
        below e^-1000): 1e-3 relative (WCA's d^-12 multiplies the distance's
        relative error by 12)."""
    say("case: pair energies fp32, mixed 12-type system")
    g = sc_amd.GpuSystem(dev)
    z, tp, st = mixed(g)
    i, j = z["mixed_i"], z["mixed_j"]
    eg = g.p2p(i, j)
    eo = po.pair_energies(tp, st, i, j)
    err = np.abs(eg - eo)
    a = np.abs(eo)
    mid, small, huge = (a >= 0.1) & (a < 1e3), a < 0.1, a >= 1e3
    rel = err / np.maximum(a, 1e-300)
    out = rel[mid] > 1e-5
    say("fp32: %d pairs; %d with 0.1 <= |E| < 1e3: %d beyond 1e-5 relative (max %.2e, quantiles 0.5/0.99/0.999 "
        "%.2e/%.2e/%.2e); %d with |E| < 0.1: max abs error %.2e; %d with |E| >= 1e3: max rel %.2e" % (
            len(i), mid.sum(), out.sum(), rel[mid].max(), np.quantile(rel[mid], 0.5), np.quantile(rel[mid], 0.99),
            np.quantile(rel[mid], 0.999), small.sum(), err[small].max(), huge.sum(), rel[huge].max()))
    for k in np.nonzero(mid)[0][out][:20]:
        say("  pair %d-%d types %d-%d: gpu %.9f oracle %.9f" % (i[k], j[k], st.types[i[k]], st.types[j[k]], eg[k],
                                                               eo[k]))
    assert out.mean() <= 5e-3
    assert rel[mid].max() <= 5e-5
    assert err[small].max() <= 2e-6
    assert rel[huge].max() <= 1e-3


def test_pair_energies_grid_all_type_pairs(dev):
    """Interactions-test geometry for all 144 type pairs, FP64 device mode
    against the oracle on the device's own (quantised) state."""
    say("case: grid pairs, 144 type pairs")
    z = np.load(os.path.join(GOLD, "pair_golden.npz"))
    box = np.array([10.0, 10.0, 10.0])
    nb = len(z["grid_pos"]) - 1
    g = sc_amd.GpuSystem(dev)
    nbad, ntot = 0, 0
    for k in range(len(z["grid_ta"])):
        ta, tb = int(z["grid_ta"][k]), int(z["grid_tb"][k])
        types = np.array([ta] + [tb] * nb, np.int32)
        st0 = po.State.from_real(box, z["grid_pos"], z["grid_dir"], z["grid_pd0"], types)
        tp, st = gpu_state(g, str(z["grid_tops"][k]), box, st0.pos, st0.dir, st0.pd0, types)
        i = np.zeros(nb, np.int32)
        j = np.arange(1, nb + 1, dtype=np.int32)
        eg = g.p2p(i, j, fp64=True)
        eo = po.pair_energies(tp, st, i, j)
        bad = np.abs(eg - eo) > 1e-9 * np.maximum(1.0, np.abs(eo))
        nbad += bad.sum()
        ntot += nb
    say("grid fp64: %d of %d pairs beyond 1e-9" % (nbad, ntot))
    assert nbad <= 1e-3 * ntot


@pytest.mark.parametrize("key", ["psc", "cpsc", "tchpsc"])
def test_energy_all_and_per_particle_golden_systems(dev, key):
    say("case: energy_all golden", key)
    z = np.load(os.path.join(GOLD, "total_golden.npz"))
    st0 = po.State.from_real(z[key + "_box"], z[key + "_pos"], z[key + "_dir"], z[key + "_pd0"], z[key + "_types"])
    g = sc_amd.GpuSystem(dev)
    tp, st = gpu_state(g, str(z[key + "_top"]), st0.box, st0.pos, st0.dir, st0.pd0, st0.types)
    eg = g.allToAll()
    eo = po.energy_all(tp, st, use_cells=False)
    say("E_all gpu %.10f oracle %.10f reference %.10f" % (eg, eo, float(z[key + "_eall"])))
    assert abs(eg - eo) <= 1e-5 * max(1.0, abs(eo))
    pp = g.energy_per_particle()
    po_ = po.energy_one_each(tp, st, np.arange(st.n))
    rel = np.abs(pp - po_) / np.maximum(np.abs(po_), 1e-3)
    say("per particle: max rel %.2e" % rel.max())
    # 1e-5 relative, or 1e-5 kT where a particle's terms cancel
    np.testing.assert_allclose(pp, po_, rtol=1e-5, atol=1e-5)
    for t in (0, 7, st.n - 1):
        assert g.oneToAll(t) == pytest.approx(po_[t], rel=1e-5, abs=1e-5)


def test_energy_all_config2_cpsc_32768(dev):
    """32,768 CPSC at packing fraction 0.3: FP32 pair terms with FP64 reduction
    within 1e-5 relative of the FP64 oracle; cell list bit-exact."""
    say("case: config 2, 32768 CPSC eta 0.3")
    w = W.workload("c2_cpsc_32768")
    g = sc_amd.GpuSystem(dev)
    g.load_topology(w["top"])
    pos, d, q = g.generate_rsa(w["types"], w["box"], seed=2024)
    tp, st = gpu_state(g, w["top"], w["box"], pos, d, q, w["types"])
    eg = g.allToAll()
    eo = po.energy_all(tp, st, use_cells=True)
    say("config2 E_all gpu %.8f oracle %.8f rel %.2e" % (eg, eo, abs(eg - eo) / abs(eo)))
    assert abs(eg - eo) <= 1e-5 * abs(eo)
    u = g.fixed_positions()
    for r in (1, 2):
        shift = (12345, 987654321, 3000000000)
        cell_of, start, srt, nc = g.build_cells(r=r, shift=shift)
        c2, s2, o2 = po.cell_list(u, nc, shift)
        np.testing.assert_array_equal(cell_of, c2)
        np.testing.assert_array_equal(start, s2)
        np.testing.assert_array_equal(srt, o2)


@pytest.mark.parametrize("n", [300, 700])
def test_cell_list_of_crowded_cells_bit_exact(dev, n):
    """Cells of more than 64 particles (a box of under two maxcut per axis
    holds one cell per axis): the in-cell order by reference index of the
    chunked rank sort (up to 512) and of the one-lane sort beyond, against
    the CPU's counting sort."""
    say("case: cell list, one cell of %d particles" % n)
    w = W.workload("c1_psc_1000", n_override=n)
    g = sc_amd.GpuSystem(dev)
    g.load_topology(w["top"])
    box = np.full(3, 1.5 * g.maxcut)
    pos = np.random.default_rng(n).random((n, 3))
    d = np.tile([0.0, 0.0, 1.0], (n, 1))
    q = np.tile([1.0, 0.0, 0.0], (n, 1))
    g.set_config(box, pos, d, q, w["types"])
    u = g.fixed_positions()
    cell_of, start, srt, nc = g.build_cells(r=1, shift=(777, 12345678, 99))
    c2, s2, o2 = po.cell_list(u, nc, (777, 12345678, 99))
    assert int(np.prod(nc)) == 1
    np.testing.assert_array_equal(cell_of, c2)
    np.testing.assert_array_equal(start, s2)
    np.testing.assert_array_equal(srt, o2)
    g.close()


def test_neighbor_pairs_bit_exact(dev):
    say("case: neighbour pairs")
    w = W.workload("c1_psc_1000", n_override=1500)
    g = sc_amd.GpuSystem(dev)
    g.load_topology(w["top"])
    pos, d, q = g.generate_rsa(w["types"], w["box"], seed=4)
    g.set_config(w["box"], pos, d, q)
    u = g.fixed_positions()
    for rc in (1.0, 2.5, g.maxcut):
        a = g.neighbor_pairs(rc)
        b = po.neighbor_pairs(u, w["box"], rc)
        np.testing.assert_array_equal(a, b)


def test_overlap_flags(dev):
    say("case: overlap flags")
    z = np.load(os.path.join(GOLD, "pair_golden.npz"))
    g = sc_amd.GpuSystem(dev)
    z, tp, st = mixed(g)
    i, j = z["mixed_i"], z["mixed_j"]
    f, dm = g.overlap_flags(i, j)
    # oracle distances: segment / sphere closest distance in FP64
    sig = np.array([tp.ia(st.types[a], st.types[b]).sigma for a, b in zip(i, j)])
    dist = np.zeros(len(i))
    for k, (a, b) in enumerate(zip(i, j)):
        dd = st.pos[a] - st.pos[b]
        r = st.box * (dd - np.rint(dd))
        ia = tp.ia(st.types[a], st.types[b])
        ga, gb = ia.geotype[0], ia.geotype[1]
        da = st.dir[a] / np.linalg.norm(st.dir[a])
        db = st.dir[b] / np.linalg.norm(st.dir[b])
        if ga >= 30 and gb >= 30:
            dist[k] = np.linalg.norm(r)
        elif ga >= 30 or gb >= 30:
            sc_dir, hl, rr = (db, ia.half_len[1], -r) if ga >= 30 else (da, ia.half_len[0], r)
            c = np.clip(sc_dir @ rr, -hl, hl)
            dist[k] = np.linalg.norm(-rr + sc_dir * c)
        else:
            dist[k] = np.linalg.norm(po.min_dist_segments(da, db, ia.half_len[0], ia.half_len[1], r))
    want = dist < sig
    near = np.abs(dist - sig) < 1e-9 * sig
    mism = (want != f.astype(bool)) & ~near
    say("overlap: %d pairs, %d overlapping, %d near contact, %d mismatches" % (len(i), want.sum(), near.sum(), mism.sum()))
    assert mism.sum() == 0


def test_trial_and_commit_protocol(dev):
    """E_old = oneToAll(t); E_new = oneToAllTrial(t, x'); update -> oneToAll = E_new
    (MoveCreator::partDisplace protocol, movecreator.cpp:947-994)"""
    say("case: trial/commit protocol")
    w = W.workload("c1_psc_1000", n_override=600)
    g = sc_amd.GpuSystem(dev)
    g.load_topology(w["top"])
    pos, d, q = g.generate_rsa(w["types"], w["box"], seed=8)
    tp, st = gpu_state(g, w["top"], w["box"], pos, d, q, w["types"])
    rng = np.random.default_rng(1)
    e_all0 = g.allToAll()
    for t in rng.choice(st.n, 10, replace=False):
        e_old = g.oneToAll(t)
        newpos = st.pos[t] + rng.normal(size=3) * 0.3 / st.box
        e_new = g.oneToAllTrial(t, newpos, st.dir[t], st.pd0[t])
        st2 = po.State(st.box, st.pos.copy(), st.dir, st.pd0, st.types)
        st2.pos[t] = newpos
        want = po.energy_one_each(tp, st2, [t])[0]
        assert e_new == pytest.approx(want, rel=1e-5, abs=1e-5)
        g.update_particle(t, newpos, st.dir[t], st.pd0[t])
        assert g.oneToAll(t) == pytest.approx(e_new, rel=1e-6, abs=1e-6)
        e_all1 = g.allToAll()
        assert abs((e_all1 - e_all0) - (e_new - e_old)) <= 1e-6 * max(1.0, abs(e_all1))
        e_all0 = e_all1
        _, p, _, _ = g.get_config()
        st.pos[:] = p


def test_sweep_energy_drift_and_determinism(dev):
    say("case: sweep drift/determinism")
    w = W.workload("c1_psc_1000")
    out = []
    for rep in range(2):
        g = sc_amd.GpuSystem(dev)
        g.load_topology(w["top"])
        pos, d, q = g.generate_rsa(w["types"], w["box"], seed=21)
        g.set_config(w["box"], pos, d, q)
        e0 = g.allToAll()
        s = g.sweep(20, temper=1.0, transmx=0.3, rot_half=20.0 / 180 * np.pi / 4, r=2, seed=99, counter=0)
        e1 = g.allToAll()
        drift = e1 - e0 - s["energy_change"]
        tries = s["trans_try"] + s["rot_try"]   # left_cell moves are counted in trans_try
        say("sweep: E %.6f -> %.6f, sum dE %.6f, drift %.3e, tries %d, acc %d/%d %d/%d left %d" % (
            e0, e1, s["energy_change"], drift, tries, s["trans_acc"], s["trans_try"], s["rot_acc"], s["rot_try"],
            s["left_cell"]))
        assert tries == 20 * 1000
        assert abs(drift) <= 1e-8 * max(1.0, abs(e1))
        _, p, dd, qq = g.get_config()
        out.append((p, dd, qq))
        g.close()
    for a, b in zip(out[0], out[1]):
        np.testing.assert_array_equal(a, b)


def test_knife_edge_pair_is_rounded_alike_everywhere(dev):
    """A pair whose FP32 term sits on a knife edge (tests/golden/
    knife_edge_pair.json: a last-bit change of a derived vector flips an
    intersection branch between 0.0437 and -0.6049 kT) has one FP32 value
    in every kernel: p2p, oneToAll and allToAll agree exactly, and the
    sweep's trials price it alike (the energy bookkeeping of sweeps that move
    the pair stays at FP64-sum rounding; with kernels that rounded it
    differently the first accepted move booked the 0.65 kT jump)."""
    import json
    say("case: knife-edge pair (2 TCHPSC of the 65,536 benchmark system)")
    z = json.load(open(os.path.join(GOLD, "knife_edge_pair.json")))
    w = W.workload("c4_tchpsc_65536")
    g = sc_amd.GpuSystem(dev)
    g.load_topology(w["top"].replace("C 65536", "C 2"))
    g.set_config(z["box"], z["pos"], z["dir"], z["pd0"], [3, 3])
    e32 = g.p2p([1], [0])[0]
    e64 = g.p2p([1], [0], fp64=True)[0]
    ea, e1, e0 = g.allToAll(), g.oneToAll(1), g.oneToAll(0)
    say("p2p FP32 %.6f FP64 %.6f allToAll %.6f oneToAll %.6f %.6f" % (e32, e64, ea, e1, e0))
    assert ea == e32 and e1 == e32 and e0 == e32
    g.set_visit_factor(1)
    worst = 0.0
    moved = 0
    for c in range(40):
        e0 = g.allToAll()
        s = g.sweep(1, 1.0, 0.02, 0.5 / 180 * np.pi, r=1, seed=3, counter=c)
        e1 = g.allToAll()
        moved += s["trans_acc"] + s["rot_acc"]
        worst = max(worst, abs(e1 - e0 - s["energy_change"]))
    say("40 sweeps, %d accepted moves, worst drift %.3e" % (moved, worst))
    g.close()
    assert moved > 0 and worst <= 1e-8


def _blocked(x, nb=8):
    b = x[: len(x) // nb * nb].reshape(nb, -1).mean(1)
    return b.mean(), b.std(ddof=1) / np.sqrt(nb)


def ensemble_check(g, n, ec, rot_half, label):
    """Run 200 + 40 x 25 sweeps on g and compare E/N, S2 and mean cluster size
    block averages with the oracle's: equal within 3 combined standard errors
    + 1% of the oracle's mean.  S2 of this isotropic fluid is O(N^-1/2) ~ 0.02
    and decorrelates over ~400 sweeps (10 degree rotations), so the block
    errors understate its spread and 1% of ~0 means nothing: its margin is
    0.03 absolute instead.  Returns the counters summed over the blocks."""
    g.sweep(200, 1.0, 0.3, rot_half, r=2, seed=5, counter=0)
    obs = []
    tot = {}
    for blk in range(40):
        s = g.sweep(25, 1.0, 0.3, rot_half, r=2, seed=5, counter=1000 + 25 * blk)
        for k, v in s.items():
            tot[k] = tot.get(k, 0) + v
        _, ncl = g.cluster_labels()
        obs.append((g.allToAll() / n, g.nematic_order()[0], n / ncl))
    obs = np.array(obs)
    for c, name in enumerate(("E/N", "S2", "mean cluster size")):
        mg, sg = _blocked(obs[:, c])
        mc, scc = _blocked(ec[:, c])
        say("%s %s: gpu %.5f +- %.5f, cpu %.5f +- %.5f" % (label, name, mg, sg, mc, scc))
        margin = 0.03 if name == "S2" else 0.01 * abs(mc)
        assert abs(mg - mc) <= 3.0 * np.hypot(sg, scc) + margin, name
    return tot


@pytest.mark.parametrize("mode", [0, 1])
def test_sweep_ensemble_matches_cpu_oracle(dev, ensemble_system, mode):
    """E/N, nematic S2 and mean cluster size of 500 PSC at eta 0.2, T*=1,
    from the checkerboard GPU sweep (trial mode 0 and 1) and from the
    sequential CPU oracle Metropolis: equal within 3 combined standard errors
    (+1%)."""
    say("case: ensemble GPU vs CPU oracle, trial mode", mode)
    w, (pos, d, q), ec, rot_half = ensemble_system
    g = sc_amd.GpuSystem(dev)
    g.load_topology(w["top"])
    g.set_config(w["box"], pos, d, q, w["types"])
    g.set_trial_mode(mode)
    n = len(w["types"])
    tot = ensemble_check(g, n, ec, rot_half, "mode %d" % mode)
    tries = tot["trans_try"] + tot["rot_try"]
    say("trials/sweep/N %.4f" % (tries / (1000.0 * n)))
    assert abs(tries / (1000.0 * n) - 1.0) < 0.02


def test_cluster_labels_and_nematic_order_match_oracle(dev):
    """Device cluster labels (FP64 pair energies <= -0.10, connected
    components) equal the oracle's, and S2 matches, on 1500 PSC after 60
    sweeps at T*=0.6 (clusters of several particles)."""
    say("case: cluster labels + nematic order vs oracle")
    w = W.workload("c1_psc_1000", n_override=1500)
    g = sc_amd.GpuSystem(dev)
    box = w["box"] * (0.2 / 0.3) ** (1.0 / 3.0)
    g.load_topology(w["top"])
    pos, d, q = g.generate_rsa(w["types"], box, seed=8)
    g.set_config(box, pos, d, q, w["types"])
    g.sweep(60, 0.6, 0.3, 20.0 / 180 * np.pi / 4, r=2, seed=8, counter=0)
    lab, ncl = g.cluster_labels()
    b, p, dd, qq = g.get_config()
    tp, st = po.Topo(w["top"]), po.State(b, p, dd, qq, w["types"])
    ref = po.cluster_labels(tp, st)
    sizes = np.bincount(np.unique(ref, return_counts=True)[1])
    say("clusters: gpu %d, oracle %d; size histogram %s" % (ncl, len(np.unique(ref)), sizes.tolist()))
    assert len(sizes) > 3
    np.testing.assert_array_equal(lab, ref)
    assert ncl == len(np.unique(ref))
    s2, director = g.nematic_order()
    s2r = po.nematic_s2(dd)
    say("S2 gpu %.10f oracle %.10f" % (s2, s2r))
    assert abs(s2 - s2r) <= 1e-6   # directions are FP32 on the device
    assert abs(np.linalg.norm(director) - 1.0) < 1e-9


@pytest.mark.parametrize("mode", [0, 2])
def test_block_waves_reproduce_wave_sweep(dev, mode):
    """The sweep with 2 or 4 waves per trial, and the speculative sweep (a
    trial per wave, kept in order up to the first dependence on an accepted
    one), follow the 1-wave sweep move for move: the same trial moves, pair
    lists and decisions; only the order of the FP64 sums differs (~1e-16
    relative), which could flip a Metropolis decision only for a uniform
    within that of its threshold."""
    say("case: 2- and 4-wave and speculative vs 1-wave sweep, surrogate mode", mode)
    w = W.workload("c1_psc_1000")
    out = []
    for nw in (1, 2, 4, 0):
        g = sc_amd.GpuSystem(dev)
        g.load_topology(w["top"])
        pos, d, q = g.generate_rsa(w["types"], w["box"], seed=23)
        g.set_config(w["box"], pos, d, q)
        g.set_block_waves(nw)
        g.set_surrogate(mode)
        g.set_network_min_entries(0)
        s = g.sweep(10, 1.0, 0.3, 20.0 / 180 * np.pi / 4, r=1, seed=5, counter=0)
        _, p, dd, qq = g.get_config()
        out.append((s, p, dd, qq))
        g.close()
    s1, p1, d1, q1 = out[0]
    for s4, p4, d4, q4 in out[1:]:
        for k in ("trans_acc", "trans_try", "rot_acc", "rot_try", "left_cell", "stage1_acc", "stage1_try"):
            assert s1[k] == s4[k], k
        assert abs(s1["energy_change"] - s4["energy_change"]) <= 1e-9 * max(1.0, abs(s1["energy_change"]))
        np.testing.assert_array_equal(p1, p4)
        np.testing.assert_array_equal(d1, d4)
        np.testing.assert_array_equal(q1, q4)


def test_volume_moves_one_edge(dev):
    """ptype 0 (pressureMove, movecreator.cpp:350-389): an accepted move changes
    one edge; the energy carried through sweeps and volume moves matches a
    full evaluation and the oracle at the final box."""
    say("case: anisotropic volume moves")
    w = W.workload("c1_psc_1000")
    g = sc_amd.GpuSystem(dev)
    g.load_topology(w["top"])
    pos, d, q = g.generate_rsa(w["types"], w["box"], seed=29)
    g.set_config(w["box"], pos, d, q)
    box = np.array(w["box"], float)
    acc = 0
    for k in range(40):
        g.sweep(1, 1.0, 0.3, 20.0 / 180 * np.pi / 4, r=1, seed=3, counter=k)
        a, b = g.volume_move(0, 2.0, 1.0, 1.0, seed=4, counter=k)
        changed = np.nonzero(b != box)[0]
        if a:
            acc += 1
            assert len(changed) == 1
        else:
            assert len(changed) == 0
        box = b
    b, p, dd, qq = g.get_config()
    np.testing.assert_array_equal(b, box)
    tp, st = po.Topo(w["top"]), po.State(b, p, dd, qq, w["types"])
    eo = po.energy_all(tp, st, use_cells=True)
    eg = g.allToAll()
    say("accepted %d/40, box %s, E gpu %.6f oracle %.6f" % (acc, np.round(b, 3).tolist(), eg, eo))
    assert acc > 0
    assert abs(eg - eo) <= 1e-5 * max(1.0, abs(eo))
    g.close()
