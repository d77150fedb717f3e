"""The oracle's sequential Metropolis sampler (orc_mc_run), pinned on its own.

Every ensemble test of the GPU sweep is refereed by orc_mc_run, so it is
checked here against things that do not come from this build:

  * its energy bookkeeping: the returned drift E_end - E_start - sum(dE) is
    zero to 1e-9 (updater.cpp:122-150 prints the same quantity);
  * the Boltzmann distribution itself: two patchy spherocylinders in a small
    periodic box, where <E>, <P2(u1.u2)> and the bound fraction follow from
    a quadrature over uniformly drawn relative positions and orientations
    weighted by exp(-E/T), with E from the oracle's pair energy (pinned to the
    compiled reference's golden vectors).  The sampler must agree within 3
    combined standard errors, and planted mistakes (a Metropolis test at 2T,
    rotations or translations without a reverse move) must fail the check;
  * the compiled reference's own sampler (scOOP's main, built from source
    under oracle/_ref), run on the same 500-particle system: equal mean E/N
    and acceptance rates (skipped where the reference sources are absent).

Reference: MoveCreator::particleMove / partDisplace / partRotate / moveTry,
scOOP/mc/movecreator.cpp:11-33, 947-1028, movecreator.h:175-187.
"""
import os
import subprocess
import tempfile

import numpy as np
import pytest

import pyoracle as po
from sc_amd import workloads as W

ROT_HALF = 20.0 / 180 * np.pi / 4

# two sticky PSC rods: EPS 3, flat well to 1.5 sigma, switch 0.5 sigma, patch
# 120 deg, L = 2; box 9 sigma > 2 maxcut (4.4 sigma)
TWO_TOP = """[Types]
A 1 PSC 3.0 1.0 1.5 0.5 120.0 10.0 2.0 0.0
[Molecules]
A: {
particles: 1
}
[System]
A 2
"""
TWO_BOX = 9.0
TWO_T = 1.0


def _unit(rng, m):
    v = rng.normal(size=(m, 3))
    return v / np.linalg.norm(v, axis=1)[:, None]


def _perp(rng, d):
    q = _unit(rng, len(d))
    q -= d * (q * d).sum(1)[:, None]
    return q / np.linalg.norm(q, axis=1)[:, None]


def _pair_obs(tp, st):
    """pair energy, P2 of the axes and the bound flag of pairs (0,1), (2,3), ..."""
    e = po.pair_energies(tp, st, np.arange(0, st.n, 2), np.arange(1, st.n, 2))
    c = (st.dir[0::2] * st.dir[1::2]).sum(1) / (
        np.linalg.norm(st.dir[0::2], axis=1) * np.linalg.norm(st.dir[1::2], axis=1))
    return np.stack([e, 1.5 * c * c - 0.5, (e < -1.5).astype(float)], axis=1)


@pytest.fixture(scope="module")
def two_quadrature():
    """Boltzmann averages of the two-rod system by direct quadrature: 2M
    relative configurations (second centre uniform in the box, both axes
    uniform on the sphere, patch directions uniform about them), weights
    exp(-E/T).  Returns [(mean, standard error)] of (E, P2, bound)."""
    tp = po.Topo(TWO_TOP)
    rng = np.random.default_rng(2024)
    obs = []
    for _ in range(4):
        m = 500_000
        pos = np.zeros((2 * m, 3))
        pos[1::2] = rng.random((m, 3))
        d = _unit(rng, 2 * m)
        st = po.State(np.full(3, TWO_BOX), pos, d, _perp(rng, d), np.ones(2 * m, np.int32))
        obs.append(_pair_obs(tp, st))
    x = np.concatenate(obs)
    w = np.exp(-x[:, 0] / TWO_T)
    out = []
    for k in range(3):
        mu = (w * x[:, k]).sum() / w.sum()
        se = np.sqrt(((w * (x[:, k] - mu)) ** 2).sum()) / w.sum()   # delta method
        out.append((mu, se))
    return out


def _two_mc(nsweeps, fault=0, chunk=20, nb=50):
    """Block means and errors of (E, P2, bound) from orc_mc_run on the two
    rods, sampled every `chunk` sweeps."""
    tp = po.Topo(TWO_TOP)
    rng = np.random.default_rng(7)
    d = _unit(rng, 2)
    st = po.State(np.full(3, TWO_BOX), np.array([[0.0, 0.0, 0.0], [0.5, 0.5, 0.5]]), d, _perp(rng, d),
                  np.ones(2, np.int32))
    po.set_sampler_fault(fault)
    try:
        series = []
        for k in range(nsweeps // chunk):
            st, drift, _, _ = po.mc_run(tp, st, chunk, TWO_T, 0.5, 0.3, seed=1000 + k, energies=(k % 100 == 0))
            if k % 100 == 0:
                assert abs(drift) <= 1e-9
            series.append(_pair_obs(tp, st)[0])
    finally:
        po.set_sampler_fault(0)
    s = np.array(series)
    s = s[: len(s) // nb * nb].reshape(nb, -1, 3).mean(1)
    return s.mean(0), s.std(0, ddof=1) / np.sqrt(nb)


def test_two_particle_boltzmann_matches_quadrature(two_quadrature):
    m, e = _two_mc(1_000_000)
    for k, name in enumerate(("E", "P2", "bound fraction")):
        qm, qs = two_quadrature[k]
        z = (m[k] - qm) / np.hypot(e[k], qs)
        print("%s: sampler %.5f +- %.5f, quadrature %.5f +- %.5f, z %.2f" % (name, m[k], e[k], qm, qs, z))
        assert abs(z) <= 3.0, name


@pytest.mark.parametrize("fault", [1, 2, 3])
def test_boltzmann_check_fails_for_planted_mistakes(two_quadrature, fault):
    """The same check with a mistake planted in the sampler: the Metropolis
    test at 2T (1), rotations about the body axis dir x patchdir in one sense
    (2), translations along +dir only (3).  Each must miss the quadrature
    by far more than 3 standard errors."""
    m, e = _two_mc(200_000, fault=fault)
    z = [(m[k] - two_quadrature[k][0]) / np.hypot(e[k], two_quadrature[k][1]) for k in range(3)]
    print("fault %d: z = %s" % (fault, np.round(z, 1).tolist()))
    assert max(abs(v) for v in z) > 10.0


def _system(name, n, eta, seed):
    import sc_amd
    w = W.workload(name, n_override=n)
    box = w["box"] * (w["eta"] / eta) ** (1.0 / 3.0)
    h = sc_amd.GpuSystem(-1)
    h.load_topology(w["top"])
    pos, d, q = h.generate_rsa(w["types"], box, seed=seed)
    return w, box, pos, d, q


@pytest.mark.parametrize("name,n,eta", [("c1_psc_1000", 500, 0.2), ("c2_cpsc_32768", 400, 0.2),
                                        ("c4_tchpsc_65536", 300, 0.15)])
def test_mc_run_energy_drift_and_acceptance(name, n, eta):
    """E_end - E_start - sum of accepted dE = 0 to 1e-9, and both move kinds
    are tried about equally often and accepted sometimes but not always."""
    w, box, pos, d, q = _system(name, n, eta, seed=5)
    tp, st = po.Topo(w["top"]), po.State(box, pos, d, q, w["types"])
    _, drift, stats, ef = po.mc_run(tp, st, 60, 1.0, 0.3, ROT_HALF, seed=3)
    ta, tt, ra, rt = stats
    print("%s: drift %.3e, E/N %.4f, translations %d/%d, rotations %d/%d" % (name, drift, ef / n, ta, tt, ra, rt))
    assert abs(drift) <= 1e-9
    assert tt + rt == 60 * n
    assert abs(tt - rt) < 6 * np.sqrt(60 * n)
    assert 0 < ta < tt and 0 < ra < rt


@pytest.mark.skipif(not po.ref_available(), reason="compiled reference (oracle/_ref/scref) not built")
def test_mc_run_matches_compiled_reference_sampler():
    """500 PSC at eta 0.2, T* 1, 0.3 sigma translations, <= 10 deg rotations:
    mean E/N over sweeps 200-1200 (every 25 sweeps) of the oracle's sampler
    and of the compiled reference's own program (energy.dat, updater.cpp:
    307-317) agree within 3 combined block errors + 1%; so do the acceptance
    rates of both move kinds (the reference prints whole percent)."""
    w, box, pos, d, q = _system("c1_psc_1000", 500, 0.2, seed=31)
    n = 500
    tp, st = po.Topo(w["top"]), po.State(box, pos, d, q, w["types"])
    series, tot = [], np.zeros(4)
    for blk in range(48):
        st, drift, stats, ef = po.mc_run(tp, st, 25, 1.0, 0.3, ROT_HALF, seed=500 + blk)
        assert abs(drift) <= 1e-9
        tot += stats
        series.append(ef / n)
    with tempfile.TemporaryDirectory() as dd:
        opts = W.options_text(nsweeps=1200, transmx=0.15, rotmx=20.0, temper=1.0, seed=4321)
        open(os.path.join(dd, "options"), "w").write(opts.replace("report = 0", "report = 25"))
        open(os.path.join(dd, "top.init"), "w").write(w["top"])
        open(os.path.join(dd, "config.init"), "w").write(W.config_text(box, pos, d, q))
        r = subprocess.run([po.REF_BIN, "run", os.path.join(dd, "wall.txt")], cwd=dd, capture_output=True,
                           text=True, timeout=600)
        assert r.returncode == 0, r.stderr[-2000:]
        rows = [ln.split(";") for ln in open(os.path.join(dd, "energy.dat")) if ";" in ln]
        ref = np.array([float(v[1].split()[0]) for v in rows]) / n
        acc = {}
        for ln in r.stdout.splitlines():
            if ln.startswith("Single particle translation:") or ln.startswith("Single particle rotation:"):
                f = ln.split(":")[1].split()
                acc.setdefault(ln.split(":")[0], float(f[0]))
    assert len(ref) == 48

    def blocked(x, nb=8):
        b = np.asarray(x[8:]).reshape(nb, -1).mean(1)
        return b.mean(), b.std(ddof=1) / np.sqrt(nb)

    mo, so = blocked(series)
    mr, sr = blocked(ref)
    print("E/N oracle %.5f +- %.5f, reference %.5f +- %.5f" % (mo, so, mr, sr))
    assert abs(mo - mr) <= 3.0 * np.hypot(so, sr) + 0.01 * abs(mr)
    at, ar = 100.0 * tot[0] / tot[1], 100.0 * tot[2] / tot[3]
    print("acceptance %% oracle %.2f / %.2f, reference %s" % (at, ar, acc))
    assert abs(at - acc["Single particle translation"]) <= 1.5
    assert abs(ar - acc["Single particle rotation"]) <= 1.5
