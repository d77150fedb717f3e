"""Whole-molecule moves on the GPU (through the C ABI): mol2others /
mol2othersTrial / update(Molecule) against the FP64 oracle and the
reference's values (tests/golden/chain_golden.npz), and the chain-move
passes of the sweep (MoveCreator::chainDisplace / chainRotate,
movecreator.cpp:1075-1305): rigid moves, energy bookkeeping, bit-for-bit
reproducibility, and the ensemble against the oracle's sequential Metropolis
with the same moves.

Tolerances: mol2others sums FP32 pair terms in FP64, 1e-4 relative / 1e-5
absolute against the oracle; rigidity to the fixed-point / FP32 storage
(1e-4 sigma, 1e-5 in the axes); ensemble averages within 3 combined
standard errors + 1 %.
"""
import os

import numpy as np
import pytest

import sc_amd
from sc_amd import workloads as W
import pyoracle as po

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
SETS = ("c5", "angles", "bondd", "bondh", "sp")
ROT_HALF = 20.0 / 180 * np.pi / 4


def say(*a):
    print(*a, flush=True)


def dev_state(g, types):
    b, p, d, q = g.get_config()
    d = d / np.linalg.norm(d, axis=1)[:, None]
    return po.State(b, p, d, q, types)


def rigid_pose(st, f, m, vol, rng):
    """the molecule displaced by 0.4 sigma and rotated by 25 deg about its centre"""
    p = st.pos[f:f + m]
    d = p - p[0]
    d -= np.rint(d)
    c = (d * vol[:, None]).sum(0) / vol.sum()
    o = (d - c) * st.box
    ax = rng.normal(size=3)
    ax /= np.linalg.norm(ax)
    a = np.radians(25.0)
    K = np.array([[0, -ax[2], ax[1]], [ax[2], 0, -ax[0]], [-ax[1], ax[0], 0]])
    R = np.eye(3) + np.sin(a) * K + (1 - np.cos(a)) * K @ K
    sh = rng.normal(size=3)
    sh *= 0.4 / np.linalg.norm(sh)
    pos = p[0] + c + (o @ R.T + sh) / st.box
    return np.mod(pos, 1.0), st.dir[f:f + m] @ R.T, st.pd0[f:f + m] @ R.T


@pytest.mark.parametrize("name", SETS)
def test_mol2others_trial_and_commit(name):
    say("case: mol2others", name)
    z = np.load(os.path.join(GOLD, "chain_golden.npz"))
    top = str(z[name + "_top"])
    st0 = po.State.from_real(z[name + "_box"], z[name + "_pos"], z[name + "_dir"], z[name + "_pd0"],
                             z[name + "_types"])
    g = sc_amd.GpuSystem(0)
    g.load_topology(top)
    g.set_config(st0.box, st0.pos, st0.dir, st0.pd0, st0.types)
    tp = po.Topo(top)
    st = dev_state(g, st0.types)
    first, size, ref = z[name + "_mfirst"], z[name + "_msize"], z[name + "_emol"]
    eo = po.mol2others(tp, st, first, size)
    eg = np.array([g.mol2others(f, m) for f, m in zip(first, size)])
    say("molecules %d: max |gpu - oracle| %.3e, max |gpu - reference| %.3e" % (
        len(first), np.abs(eg - eo).max(), np.abs(eg - ref).max()))
    np.testing.assert_allclose(eg, eo, rtol=1e-4, atol=1e-5)
    np.testing.assert_allclose(eg, ref, rtol=1e-3, atol=1e-4)
    # a one-particle "molecule" taken out of a chain: mol2others's empty
    # connectivity list leaves the bonds out (oneToAll has them)
    f = int(first[0])
    want1 = po.mol2others(tp, st, [f], [1])[0]
    assert g.mol2others(f, 1) == pytest.approx(want1, rel=1e-4, abs=1e-5)
    rng = np.random.default_rng(4)
    vol = np.array([tp.ia(t, t).volume for t in st.types])
    for k in range(min(4, len(first))):
        f, m = int(first[k]), int(size[k])
        pos, d, q = rigid_pose(st, f, m, vol[f:f + m], rng)
        e_trial = g.mol2othersTrial(f, m, pos, d, q)
        st2 = po.State(st.box, st.pos.copy(), st.dir.copy(), st.pd0.copy(), st.types)
        st2.pos[f:f + m], st2.dir[f:f + m], st2.pd0[f:f + m] = pos, d, q
        want = po.mol2others(tp, st2, [f], [m])[0]
        say("molecule at %d (%d particles): trial %.6f oracle %.6f" % (f, m, e_trial, want))
        assert e_trial == pytest.approx(want, rel=1e-4, abs=1e-5)
        # update(Molecule): the whole-system energy changes by the change of
        # mol2others plus that of the pairs inside the molecule (zero for an
        # exact rigid move; here the rounding of the stored state)
        e0, m0 = g.allToAll(), g.mol2others(f, m)
        inner0 = po.energy_all(tp, st, use_cells=False) - po.mol2others(tp, st, [f], [m])[0]
        g.update_mol(f, m, pos, d, q)
        st = dev_state(g, st.types)
        assert g.mol2others(f, m) == pytest.approx(e_trial, rel=1e-4, abs=1e-5)
        inner1 = po.energy_all(tp, st, use_cells=False) - po.mol2others(tp, st, [f], [m])[0]
        de = g.allToAll() - e0
        say("  dE_all %.6f, d mol2others %.6f, d inside %.3e" % (de, e_trial - m0, inner1 - inner0))
        assert abs(de - (e_trial - m0 + inner1 - inner0)) <= 2e-5 * abs(e0) + 1e-4
    g.close()


def lattice(g, nch, eta, seed):
    top = W.chain_topology(nch)
    box, pos, d, q, t = W.chain_lattice(nch, eta, seed)
    g.load_topology(top)
    g.set_config(box, pos, d, q, t)
    return po.Topo(top), (box, pos, d, q, t)


def offsets(b, p, d, q, f, m, vol):
    x = p[f:f + m] - p[f]
    x -= np.rint(x)
    c = (x * vol[:, None]).sum(0) / vol.sum()
    return np.vstack([(x - c) * b, d[f:f + m] / np.linalg.norm(d[f:f + m], axis=1)[:, None], q[f:f + m]]), p[f] + c


def kabsch(a, b):
    u, s, vt = np.linalg.svd(a.T @ b)
    dd = np.sign(np.linalg.det(vt.T @ u.T))
    return vt.T @ np.diag([1.0, 1.0, dd]) @ u.T


@pytest.mark.parametrize("extent,visit", [(0.0, 1), (7.0, 1), (0.0, 27), (7.0, 27)])
def test_chain_move_passes_are_rigid_reproducible_and_keep_the_energy(extent, visit):
    """Particle moves of length and angle 0 (they change nothing), chain moves
    at chainprob 0.25: every molecule keeps its shape, the energy changes by
    the sum of the accepted dE, a second run with the same seed gives the
    same state bit for bit, and the sweep makes chainprob N molecule trials.
    With the automatic extent each coarse cell's neighbourhood fits the LDS
    stage; an extent of 7 sigma makes the regions too large for it, so the
    passes read global memory.  Chain visit factor 1 (a pass per colour) and
    27 (the largest divisor of the colour count: one or a few passes, many
    trials each)."""
    say("case: chain-move passes, 8000 chains, extent", extent, "chain visit", visit)
    g = sc_amd.GpuSystem(0)
    nch = 8000
    tp, (box, pos, d, q, t) = lattice(g, nch, 0.2, 6)
    g.set_chain_moves(0.25, 0.6, 0.25, extent)
    g.set_chain_visit_factor(visit)
    rho = g.chain_extent()
    say("box %.3f, extent %.3f, maxcut %.3f" % (box[0], rho, tp.maxcut))
    assert box[0] >= 2 * (2 * rho + tp.maxcut)   # at least two coarse cells per axis
    e0 = g.allToAll()
    b0, p0, d0, q0 = g.get_config()
    s = g.sweep(3, 1.1, 0.0, 0.0, r=1, seed=9, counter=0)
    period, staged = g.chain_pass_info()
    vis = g.chain_visit_info()
    say("colour period %d, LDS stage %s, cells skipped %d, visit factor %d" % (period, staged,
                                                                               s["chain_cells_skipped"], vis))
    assert staged == (extent == 0.0)
    assert vis == (1 if visit == 1 else period ** 3)
    assert s["chain_cells_skipped"] == 0
    e1 = g.allToAll()
    n = len(t)
    tries = s["chain_disp_try"] + s["chain_rot_try"]
    say("chain tries %d (expected %.0f): displace %d/%d, rotate %d/%d, rejected %d; E %.4f -> %.4f, sum dE %.4f" % (
        tries, 0.25 * n * 3, s["chain_disp_acc"], s["chain_disp_try"], s["chain_rot_acc"], s["chain_rot_try"],
        s["chain_rejected"], e0, e1, s["energy_change"]))
    # a pass's trials are its cells' molecules times chainprob N / nchains
    # times the visit factor: with one colour per sweep (visit 27) a sweep's
    # count follows that colour's share of the molecules (~300 of 8000 here),
    # exact only on average over colours
    assert abs(tries - 0.25 * n * 3) <= (0.02 if vis == 1 else 0.1) * 0.25 * n * 3
    assert 0 < s["chain_disp_acc"] < s["chain_disp_try"] and 0 < s["chain_rot_acc"] < s["chain_rot_try"]
    assert s["chain_rejected"] < 0.2 * tries
    drift = (e1 - e0) - s["energy_change"]
    say("drift %.3e (%.2e of |E|)" % (drift, abs(drift) / abs(e0)))
    assert abs(drift) <= 1e-8 * abs(e0)
    b1, p1, d1, q1 = g.get_config()
    vol = np.array([tp.ia(x, x).volume for x in t[:3]])
    moved, worst, worst_r = 0, 0.0, 0.0
    for c in range(0, nch, 7):
        a0, c0 = offsets(b0, p0, d0, q0, 3 * c, 3, vol)
        a1, c1 = offsets(b1, p1, d1, q1, 3 * c, 3, vol)
        dist0 = np.linalg.norm(a0[:3, None] - a0[None, :3], axis=2)
        dist1 = np.linalg.norm(a1[:3, None] - a1[None, :3], axis=2)
        worst = max(worst, np.abs(dist1 - dist0).max())
        R = kabsch(a0, a1)
        worst_r = max(worst_r, np.abs(a0 @ R.T - a1).max())
        dc = c1 - c0
        dc -= np.rint(dc)
        if np.abs(a1 - a0).max() > 1e-4 or np.abs(dc * b0).max() > 1e-4:
            moved += 1
    say("molecules checked %d, moved %d; worst distance change %.2e, worst rotation residual %.2e" % (
        len(range(0, nch, 7)), moved, worst, worst_r))
    assert worst < 1e-4 and worst_r < 2e-5
    # every sweep visits at least 1/vis of the colours (all of them at vis 1)
    assert moved > 0.5 / vis * len(range(0, nch, 7))
    # the same run again
    g.set_config(box, pos, d, q, t)
    g.set_chain_moves(0.25, 0.6, 0.25, rho)
    g.set_chain_visit_factor(visit)
    g.sweep(3, 1.1, 0.0, 0.0, r=1, seed=9, counter=0)
    b2, p2, d2, q2 = g.get_config()
    assert np.array_equal(p1, p2) and np.array_equal(d1, d2) and np.array_equal(q1, q2)
    g.close()


def test_chain_moves_refuse_a_box_a_molecule_spans():
    say("case: chain moves in a small box")
    g = sc_amd.GpuSystem(0)
    lattice(g, 40, 0.15, 3)   # box 12.1 sigma < 2 x extent + maxcut
    g.set_chain_moves(0.1, 0.5, 0.2)
    with pytest.raises(sc_amd.SCError, match="own periodic image"):
        g.sweep(1, 1.1, 0.3, ROT_HALF, r=1, seed=1, counter=0)
    g.close()


def _blocked(x, nb=8):
    b = np.array_split(np.asarray(x), nb)
    m = np.array([v.mean() for v in b])
    return m.mean(), m.std(ddof=1) / np.sqrt(nb)


def _bond_energy_per_bond(tp, st):
    nb = st.n // 3
    i = np.concatenate([np.arange(nb) * 3 + 1, np.arange(nb) * 3 + 2])
    return po.bond_energies(tp, st, i, i - 1).mean()


def _end_to_end(st):
    x = st.pos[2::3] - st.pos[0::3]
    x -= np.rint(x)
    return np.linalg.norm(x * st.box, axis=1).mean()


@pytest.mark.parametrize("visit", [1, 27])
def test_chain_move_ensemble_matches_cpu_oracle(visit):
    """120 chains at eta 0.15, T* 1.1, particle moves plus chain moves
    (chainprob 0.1, displacement 0.5, rotations up to 2 x 0.2 rad): E/N, the
    mean bond energy and the head-to-tail distance of the GPU sweep and of
    the oracle's sequential Metropolis with the same moves; chain visit
    factor 1 and 27 (the largest divisor of the colour count)."""
    say("case: chain-move ensemble GPU vs CPU oracle, chain visit", visit)
    nch = 120
    top = W.chain_topology(nch)
    tp = po.Topo(top)
    box, pos, d, q, t = W.chain_lattice(nch, 0.15, 11)
    st = po.State(box, pos, d, q, t)
    cur, _, _, _ = po.mc_run_chains(tp, st, 300, 1.1, 0.3, ROT_HALF, 0.1, 0.5, 0.2, seed=21)
    ec = []
    cst = np.zeros(10)
    for blk in range(32):
        cur, drift, s, ef = po.mc_run_chains(tp, cur, 25, 1.1, 0.3, ROT_HALF, 0.1, 0.5, 0.2, seed=300 + blk)
        cst += s
        assert abs(drift) < 1e-8 * max(1.0, abs(ef))
        ec.append((ef / st.n, _bond_energy_per_bond(tp, cur), _end_to_end(cur)))
    ec = np.array(ec)
    say("oracle chain moves: displace %d/%d, rotate %d/%d" % tuple(cst[6:10]))
    g = sc_amd.GpuSystem(0)
    g.load_topology(top)
    g.set_config(box, pos, d, q, t)
    g.set_chain_moves(0.1, 0.5, 0.2)
    g.set_chain_visit_factor(visit)
    n = len(t)
    g.sweep(300, 1.1, 0.3, ROT_HALF, r=1, seed=5, counter=0)
    assert g.chain_pass_info()[1]   # the LDS-staged passes
    say("chain visit factor %d, colour period %d" % (g.chain_visit_info(), g.chain_pass_info()[0]))
    obs = []
    tot = {}
    for blk in range(32):
        s = g.sweep(25, 1.1, 0.3, ROT_HALF, r=1, seed=5, counter=1000 + 25 * blk)
        for k, v in s.items():
            tot[k] = tot.get(k, 0) + v
        stg = dev_state(g, t)
        obs.append((g.allToAll() / n, _bond_energy_per_bond(tp, stg), _end_to_end(stg)))
    obs = np.array(obs)
    say("gpu chain moves: displace %d/%d, rotate %d/%d, rejected %d" % (
        tot["chain_disp_acc"], tot["chain_disp_try"], tot["chain_rot_acc"], tot["chain_rot_try"],
        tot["chain_rejected"]))
    assert tot["chain_disp_acc"] > 0 and tot["chain_rot_acc"] > 0
    for c, name in enumerate(("E/N", "bond energy", "head-to-tail distance")):
        mg, sg = _blocked(obs[:, c])
        mc, scc = _blocked(ec[:, c])
        say("%s: gpu %.5f +- %.5f, cpu %.5f +- %.5f" % (name, mg, sg, mc, scc))
        assert abs(mg - mc) <= 3.0 * np.hypot(sg, scc) + 0.01 * abs(mc), name
    g.close()


def test_chain_move_npt_ensemble_matches_cpu_oracle():
    """NPT (P* = 0.5, T* = 1.1, one isotropic volume move per sweep) of 400
    chains with particle and chain moves (chainprob 0.1): packing fraction and
    E/N of the GPU against the oracle's sequential Metropolis with the same
    moves, both continuing from one oracle-compressed state: equal within 3
    combined standard errors + 1 %.  The chain moves run on boxes that change
    every sweep."""
    say("case: chain-move NPT ensemble GPU vs CPU oracle")
    nch, press, temper, edge = 400, 0.5, 1.1, 0.3
    cp_, cmx, cang = 0.1, 0.5, 0.2
    top = W.chain_topology(nch)
    tp = po.Topo(top)
    box, pos, d, q, t = W.chain_lattice(nch, 0.15, 23)
    st, _, _, _ = po.mc_run_chains(tp, po.State(box, pos, d, q, t), 800, temper, 0.3, ROT_HALF, cp_, cmx, cang,
                                   seed=5, ptype=1, press=press, edge_mx=edge)
    vchain = nch * W.chain_volume()
    n = st.n
    cur, drift, _, _ = po.mc_run_chains(tp, st, 200, temper, 0.3, ROT_HALF, cp_, cmx, cang, seed=7, ptype=1,
                                        press=press, edge_mx=edge)
    assert abs(drift) < 1e-7 * max(1.0, n)
    ec = []
    for blk in range(24):
        cur, _, s, ef = po.mc_run_chains(tp, cur, 25, temper, 0.3, ROT_HALF, cp_, cmx, cang, seed=300 + blk, ptype=1,
                                         press=press, edge_mx=edge)
        ec.append((vchain / np.prod(cur.box), ef / n))
    ec = np.array(ec)
    g = sc_amd.GpuSystem(0)
    g.load_topology(top)
    g.set_config(st.box, st.pos, st.dir, st.pd0, st.types)
    # an extent limit of 4.8 sigma keeps 4 x the extent below the compressed
    # box (~21 sigma); wider molecules do not move (same ensemble, slower)
    g.set_chain_moves(cp_, cmx, cang, 4.8)
    say("box %.3f, extent %.3f" % (st.box[0], g.chain_extent()))
    obs = []
    tot = {}
    for sw in range(800):
        s = g.sweep(1, temper, 0.3, ROT_HALF, r=1, seed=11, counter=sw)
        g.volume_move(1, press, temper, edge, seed=12, counter=sw)
        if sw >= 200:
            for k, v in s.items():
                tot[k] = tot.get(k, 0) + v
            if (sw - 200) % 25 == 24:
                b = g.get_config()[0]
                obs.append((vchain / np.prod(b), g.allToAll() / n))
    obs = np.array(obs)
    say("gpu chain moves: displace %d/%d, rotate %d/%d, rejected %d, cells skipped %d; period %d, LDS stage %s" % (
        tot["chain_disp_acc"], tot["chain_disp_try"], tot["chain_rot_acc"], tot["chain_rot_try"],
        tot["chain_rejected"], tot["chain_cells_skipped"], *g.chain_pass_info()))
    assert tot["chain_disp_acc"] > 0 and tot["chain_rot_acc"] > 0 and tot["chain_cells_skipped"] == 0
    for c, name in enumerate(("packing fraction", "E/N")):
        mg, sg = _blocked(obs[:, c])
        mc, scc = _blocked(ec[:, c])
        say("%s: gpu %.5f +- %.5f, cpu %.5f +- %.5f" % (name, mg, sg, mc, scc))
        assert abs(mg - mc) <= 3.0 * np.hypot(sg, scc) + 0.01 * abs(mc), name
    g.close()
