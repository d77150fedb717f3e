"""Bonded chains in the FP64 oracle: the reference's golden vectors for every
bond and angle term (tests/golden/chain_golden.npz, written by the compiled
reference) and closed forms of each term on hand-built geometries."""
import os

import numpy as np
import pytest

import pyoracle as po

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
SETS = ("c5", "angles", "bondd", "bondh", "sp")


def load():
    return np.load(os.path.join(GOLD, "chain_golden.npz"), allow_pickle=False)


def golden_state(z, name):
    tp = po.Topo(str(z[name + "_top"]))
    st = po.State.from_real(z[name + "_box"], z[name + "_pos"], z[name + "_dir"], z[name + "_pd0"],
                            z[name + "_types"])
    return tp, st


@pytest.mark.parametrize("name", SETS)
def test_chain_pair_energies_match_reference(name):
    """p2p over every ordered pair (PairE with part1's connectivity list)."""
    z = load()
    tp, st = golden_state(z, name)
    assert tp.has_chains
    pi, pj, ref = z[name + "_i"], z[name + "_j"], z[name + "_e"]
    e = po.pair_energies(tp, st, pi, pj)
    np.testing.assert_allclose(e, ref, rtol=1e-12, atol=1e-12)
    # the bond terms carry weight in these fixtures: bonded pairs exist and
    # their bond part is not lost in the non-bonded energy
    b = po.bond_energies(tp, st, pi, pj)
    bonded = b != 0.0
    assert bonded.sum() >= 20
    assert np.max(np.abs(b[bonded])) > 0.5
    assert np.max(np.abs(ref[bonded])) < 1e5
    # the golden vectors are symmetric: E(i,j) = E(j,i)
    n = st.n
    m = np.zeros((n, n))
    m[pi, pj] = ref
    np.testing.assert_allclose(m, m.T, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("name", SETS)
def test_chain_totals_match_reference(name):
    z = load()
    tp, st = golden_state(z, name)
    e = po.energy_all(tp, st, use_cells=False)
    assert e == pytest.approx(float(z[name + "_eall"]), rel=1e-12, abs=1e-10)
    # cell grid: non-bonded pairs through the stencil, bonds through the chain
    ec = po.energy_all(tp, st, use_cells=True)
    assert ec == pytest.approx(e, rel=1e-12, abs=1e-10)
    one = po.energy_one_each(tp, st, np.arange(st.n))
    np.testing.assert_allclose(one, z[name + "_eone"], rtol=1e-11, atol=1e-10)


# ---------------------------------------------------------------- closed forms
PSC = "T 5 PSC 1.0 1.0 1.0 0.2 80.0 5.0 3.0 0.0"
SPA = "H 4 SPA 1.0 1.2 1.0 0.2"


def chain_top(types_lines, mol_lines, particles, count=1):
    t = "[Types]\n" + "\n".join(types_lines) + "\n[Molecules]\nM: {\n" + "".join(l + "\n" for l in mol_lines)
    t += "".join("particles: %d\n" % p for p in particles) + "}\n[System]\nM %d\n" % count
    return po.Topo(t)


def state(box, pos, dirs, types, pd0=None):
    pos = np.asarray(pos, float)
    dirs = np.asarray(dirs, float)
    if pd0 is None:
        pd0 = np.array([np.cross(d, [0.3, 0.5, 0.8]) for d in dirs])
    return po.State.from_real(np.full(3, float(box)), pos, dirs, pd0, np.asarray(types, np.int32))


def test_harmonic_sc_end_to_end():
    """bead i's -dir end bonds to bead i+1's +dir end: E = k/2 (l - l0)^2."""
    tp = chain_top([PSC], ["bond1: 100 1.0"], [5, 5])
    rng = np.random.default_rng(1)
    for _ in range(20):
        d0 = rng.normal(size=3); d0 /= np.linalg.norm(d0)
        d1 = rng.normal(size=3); d1 /= np.linalg.norm(d1)
        c0 = np.array([10.0, 10.0, 10.0])
        c1 = c0 + rng.normal(size=3) * 2.5
        st = state(20.0, [c0, c1], [d0, d1], [5, 5])
        ell = np.linalg.norm((c0 - 1.5 * d0) - (c1 + 1.5 * d1))
        e = po.bond_energies(tp, st, [0, 1], [1, 0])
        np.testing.assert_allclose(e, 50.0 * (ell - 1.0) ** 2, rtol=1e-12)


def test_harmonic_sc_sphere_head_and_far_partner():
    """A sphere bonds at its centre (sp = 0 for bond1); a bonded pair beyond
    sqmaxcut still interacts (PairE's early exit needs an empty list)."""
    tp = chain_top([SPA, PSC], ["bond1: 100 1.0"], [4, 5])
    c0 = np.array([5.0, 5.0, 5.0])
    d1 = np.array([0.0, 0.6, 0.8])
    for ell in (1.2, 0.7, 9.0):
        c1 = c0 + np.array([ell, 0, 0]) - 1.5 * d1       # tail's +end at distance ell along x
        st = state(30.0, [c0, c1], [[0, 0, 1.0], d1], [4, 5])
        want = 50.0 * (ell - 1.0) ** 2
        b = po.bond_energies(tp, st, [0, 1], [1, 0])
        np.testing.assert_allclose(b, want, rtol=1e-12)
        if np.linalg.norm(c1 - c0) ** 2 > tp.sqmaxcut:
            e = po.pair_energies(tp, st, [0, 1], [1, 0])
            np.testing.assert_allclose(e, want, rtol=1e-12)


def test_bond2_and_harmonic_sp_use_centre_distance():
    tp = chain_top([PSC], ["bond1: 100 1.0", "bond2: 7 2.5"], [5, 5, 5])
    pos = [[5, 5, 5], [5, 5, 1.5], [5, 5, -2.0]]
    st = state(30.0, pos, [[0, 0, 1.0]] * 3, [5, 5, 5])
    b = po.bond_energies(tp, st, [0, 2], [2, 0])
    np.testing.assert_allclose(b, 3.5 * (7.0 - 2.5) ** 2, rtol=1e-12)
    tps = chain_top(["N 7 SPN 1.0 0.95", "S 8 SPA 1.0 1.0 1.12246205 1.6"], ["bond1: 90 1.0", "bond2: 10 2.2"],
                    [7, 8, 8])
    pos = np.array([[5, 5, 5], [5.9, 5.3, 5], [6.2, 6.4, 5.5]])
    st = state(30.0, pos, [[0, 0, 1.0]] * 3, [7, 8, 8])
    d01, d02 = np.linalg.norm(pos[1] - pos[0]), np.linalg.norm(pos[2] - pos[0])
    b = po.bond_energies(tps, st, [0, 0, 1], [1, 2, 0])
    np.testing.assert_allclose(b, [45.0 * (d01 - 1.0) ** 2, 5.0 * (d02 - 2.2) ** 2, 45.0 * (d01 - 1.0) ** 2],
                               rtol=1e-12)


def test_angle1_between_segments():
    """angle1 on two bonded segments: the angle between their axes."""
    tp = chain_top([PSC], ["bond1: 0 1.0", "angle1: 12 40"], [5, 5])
    for th in (0.3, 1.0, 2.2):
        d1 = np.array([np.sin(th), 0, np.cos(th)])
        st = state(30.0, [[5, 5, 5], [5, 5, 1]], [[0, 0, 1.0], d1], [5, 5])
        b = po.bond_energies(tp, st, [0, 1], [1, 0])
        np.testing.assert_allclose(b, 6.0 * (th - np.radians(40)) ** 2, rtol=1e-10)


def test_angle1_sphere_uses_bond_direction():
    """A sphere's axis in angle1 is the unit vector from the sphere centre to
    the segment's bonded end (paire.h:298-316)."""
    tp = chain_top([SPA, PSC], ["bond1: 0 1.0", "angle1: 12 40"], [4, 5])
    c0 = np.array([5.0, 5.0, 5.0])
    d1 = np.array([0.0, 0.0, 1.0])
    for off in ([0.9, 0, 0.2], [0.3, 0.2, 1.0], [0.1, -0.4, -0.8]):
        c1 = c0 - np.array(off) - 1.5 * d1     # tail's +end at c0 - off
        st = state(30.0, [c0, c1], [[1.0, 0, 0], d1], [4, 5])
        v1 = -np.array(off) / np.linalg.norm(off)   # the bonded end minus the sphere centre
        th = np.arccos(v1 @ d1)
        b = po.bond_energies(tp, st, [0, 1], [1, 0])
        np.testing.assert_allclose(b, 6.0 * (th - np.radians(40)) ** 2, rtol=1e-10)


def test_angle2_between_patches():
    """dir_a = z, dir_b = x, patch_a = x, patch_b = cos(phi) z + sin(phi) y:
    the projected patch angle is pi - phi (paire.h:335-350)."""
    tp = chain_top([PSC], ["bond1: 0 1.0", "angle2: 8 30"], [5, 5])
    for phi in (0.4, 1.3, 2.5):
        pb = np.array([0.0, np.sin(phi), np.cos(phi)])
        st = state(30.0, [[5, 5, 5], [5, 5, 1]], [[0, 0, 1.0], [1.0, 0, 0]], [5, 5],
                   pd0=[[1.0, 0, 0], pb])
        b = po.bond_energies(tp, st, [0, 1], [1, 0])
        np.testing.assert_allclose(b, 4.0 * (np.pi - phi - np.radians(30)) ** 2, rtol=1e-10)


def test_bondd_and_bondh_extend_the_ends():
    """bondd: bead i+1's bonded (+dir) end is pushed out by bonddeq; bondh: both
    ends by bondheq and a sphere bonds at bondheq along its axis; eq = 0."""
    tp = chain_top([PSC], ["bondd: 40 0.5"], [5, 5])
    c0, c1 = np.array([5.0, 5, 5]), np.array([5.0, 5.4, 1.0])
    d0, d1 = np.array([0, 0, 1.0]), np.array([0, 0.6, 0.8])
    st = state(30.0, [c0, c1], [d0, d1], [5, 5])
    ell = np.linalg.norm((c0 - 1.5 * d0) - (c1 + 2.0 * d1))
    np.testing.assert_allclose(po.bond_energies(tp, st, [0, 1], [1, 0]), 20.0 * ell ** 2, rtol=1e-12)
    tp = chain_top([SPA, PSC], ["bondh: 40 0.3"], [4, 5])
    dh = np.array([0.0, 1.0, 0.0])
    st = state(30.0, [c0, c1], [dh, d1], [4, 5])
    ell = np.linalg.norm((c0 + 0.3 * dh) - (c1 + 1.8 * d1))
    np.testing.assert_allclose(po.bond_energies(tp, st, [0, 1], [1, 0]), 20.0 * ell ** 2, rtol=1e-12)


def test_atomic_and_unbonded_pairs_have_no_bond_term():
    z = load()
    tp, st = golden_state(z, "angles")   # chains K (5 beads) then atomic A
    n = st.n
    i, j = np.nonzero(~np.eye(n, dtype=bool))
    b = po.bond_energies(tp, st, i, j)
    chain = tp.chain
    same = chain[i, 0] == chain[j, 0]
    mol_i = np.cumsum(chain[:, 1] == 0) - 1
    near = (mol_i[i] == mol_i[j]) & (chain[i, 0] == 0) & (np.abs(i - j) <= 2)
    assert np.all(b[~near] == 0.0)
    assert np.all(b[near] != 0.0)
    assert same.any()
