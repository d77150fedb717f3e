"""Whole-molecule moves in the FP64 oracle: mol2others against the compiled
reference's values (tests/golden/chain_golden.npz), and the geometry of the
chain moves (MoveCreator::chainDisplace / chainRotate / clusterRotate,
movecreator.cpp:1075-1305) through invariants a wrong centre, axis or
rotation would break."""
import os

import numpy as np
import pytest

import pyoracle as po
from sc_amd import workloads as W

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
SETS = ("c5", "angles", "bondd", "bondh", "sp")


def load():
    return np.load(os.path.join(GOLD, "chain_golden.npz"), allow_pickle=False)


@pytest.mark.parametrize("name", SETS)
def test_mol2others_matches_reference(name):
    z = load()
    tp = po.Topo(str(z[name + "_top"]))
    st = po.State.from_real(z[name + "_box"], z[name + "_pos"], z[name + "_dir"], z[name + "_pd0"],
                            z[name + "_types"])
    first, size = z[name + "_mfirst"], z[name + "_msize"]
    ref = z[name + "_emol"]
    np.testing.assert_allclose(z[name + "_emol_trial"], ref, rtol=1e-12, atol=1e-12)
    e = po.mol2others(tp, st, first, size)
    np.testing.assert_allclose(e, ref, rtol=1e-11, atol=1e-11)
    assert np.count_nonzero(ref) >= 3
    # identity with the pinned per-particle energies: a molecule's energy with
    # the rest = its members' oneToAll minus the pairs inside it (each counted
    # from both ends)
    n = st.n
    pe = np.zeros((n, n))
    pe[z[name + "_i"], z[name + "_j"]] = z[name + "_e"]
    one = z[name + "_eone"]
    for f, m, want in zip(first, size, ref):
        idx = np.arange(f, f + m)
        inner = pe[np.ix_(idx, idx)].sum()
        assert one[idx].sum() - inner == pytest.approx(want, rel=1e-9, abs=1e-9)


def chain_system(nch, eta, seed):
    top = W.chain_topology(nch)
    box, pos, d, q, t = W.chain_lattice(nch, eta, seed)
    return po.Topo(top), po.State(box, pos, d, q, t)


def unwrap_offsets(st, first, size, vol):
    """bead offsets from the molecule's volume-weighted centre (real units)"""
    p = st.pos[first:first + size]
    d = p - p[0]
    d -= np.rint(d)
    c = (d * vol[:, None]).sum(0) / vol.sum()
    return (d - c) * st.box, p[0] + c


def kabsch(a, b):
    """rotation R minimising |R a_k - b_k| (rows are vectors)"""
    h = a.T @ b
    u, s, vt = np.linalg.svd(h)
    dd = np.sign(np.linalg.det(vt.T @ u.T))
    return vt.T @ np.diag([1.0, 1.0, dd]) @ u.T


def test_chain_moves_are_rigid_and_keep_the_energy_bookkeeping():
    """mol2others leaves out the pairs inside the molecule, which a rigid move
    keeps, as long as no molecule reaches its own periodic image (box edge
    17.1 sigma > chain length 6.5 + maxcut)"""
    tp, st = chain_system(120, 0.15, 3)
    assert st.box[0] > 6.5 + tp.maxcut
    vol = np.array([tp.ia(t, t).volume for t in st.types[:3]])
    st2, drift, stats, ef = po.mc_run_chains(tp, st, 6, 1.1, 0.0, 0.0, 1.0, 0.6, 0.25, seed=7)
    assert stats[7] > 0 and stats[9] > 0
    assert 0 < stats[6] < stats[7] and 0 < stats[8] < stats[9]
    assert stats[1] == 0 and stats[3] == 0          # chainprob 1: no particle moves
    assert abs(drift) < 1e-9 * max(1.0, abs(ef))
    moved = 0
    for c in range(120):
        f = 3 * c
        o0, c0 = unwrap_offsets(st, f, 3, vol)
        o1, c1 = unwrap_offsets(st2, f, 3, vol)
        # distances inside the molecule are kept
        for a in range(3):
            for b in range(a):
                assert np.linalg.norm(o1[a] - o1[b]) == pytest.approx(np.linalg.norm(o0[a] - o0[b]), abs=1e-9)
        # one rotation maps the offsets, axes and patch directions together
        a0 = np.vstack([o0, st.dir[f:f + 3], st.pd0[f:f + 3]])
        a1 = np.vstack([o1, st2.dir[f:f + 3], st2.pd0[f:f + 3]])
        R = kabsch(a0, a1)
        assert np.abs(a0 @ R.T - a1).max() < 1e-9
        if np.abs(a0 - a1).max() > 1e-6 or np.abs(c0 - c1).max() > 1e-9:
            moved += 1
    assert moved > 60


def test_chain_rotation_keeps_the_centre_across_the_box_edge():
    """Displacements of length 0 and rotations only: the volume-weighted
    centre (minimum image from the first particle) stays put, also for a
    molecule split by the periodic boundary."""
    tp, st = chain_system(8, 0.02, 5)
    vol = np.array([tp.ia(t, t).volume for t in st.types[:3]])
    # centre the first molecule on the box corner: it straddles the edges
    d = st.pos[0:3] - st.pos[0]
    d -= np.rint(d)
    c = (d * vol[:, None]).sum(0) / vol.sum()
    st.pos[0:3] = np.mod(d - c, 1.0)
    assert np.ptp(st.pos[0:3, 2]) > 0.5
    st2, drift, stats, ef = po.mc_run_chains(tp, st, 20, 1.1, 0.0, 0.0, 1.0, 0.0, 0.6, seed=11)
    assert stats[8] > 10
    for c in range(8):
        _, c0 = unwrap_offsets(st, 3 * c, 3, vol)
        _, c1 = unwrap_offsets(st2, 3 * c, 3, vol)
        d = c1 - c0
        d -= np.rint(d)
        assert np.abs(d * st.box).max() < 1e-9


def test_chain_displacement_has_the_exact_length():
    """A lone molecule (nothing to interact with, every move accepted), chain
    moves only, rotations of angle 0: after a sweep with exactly one
    displacement the centre moved by exactly chain_mx."""
    tp, st = chain_system(1, 0.002, 2)
    vol = np.array([tp.ia(t, t).volume for t in st.types[:3]])
    seen = 0
    for k in range(24):
        st2, drift, stats, ef = po.mc_run_chains(tp, st, 1, 1.1, 0.0, 0.0, 1.0, 0.7, 0.0, seed=100 + k)
        assert stats[6] == stats[7] and stats[8] == stats[9] and stats[7] + stats[9] == 3
        _, c0 = unwrap_offsets(st, 0, 3, vol)
        _, c1 = unwrap_offsets(st2, 0, 3, vol)
        d = c1 - c0
        d -= np.rint(d)
        dist = np.linalg.norm(d * st.box)
        if stats[7] == 0:
            assert dist < 1e-12
        if stats[7] == 1:
            assert dist == pytest.approx(0.7, abs=1e-9)
            seen += 1
    assert seen >= 3
