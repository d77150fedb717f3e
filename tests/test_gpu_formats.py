"""The run's report files in the reference's own formats, against Python
restatements of its printf formats fed with the oracle's values:

  * movie frame (Updater::dumpMovie, mc/updater.h:82-93) and config.last
    (main.cpp:304-310): the header, the box line and one Conf::draw line per
    particle (Conf.h:413-421);
  * stat.dat / energy.dat lines (mc/updater.cpp:322-323) with allToAll and
    Updater::alignmentOrder (:466-482);
  * cluster_stat.dat / cluster.dat frames of ClusterSampler::writeCluster
    (mc/clust.cpp:7-49, clust.h:100-176).

Numbers the device computes in FP32 pair terms (allToAll) are compared with
1e-5 relative tolerance; FP64 quantities (alignment order, cluster energies)
to their printed digits; everything else character for character.
"""
import re

import numpy as np
import pytest

import sc_amd
from sc_amd import workloads as W
import pyoracle as po

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def dev_state(g, types):
    b, p, d, q = g.get_config()
    d = d / np.linalg.norm(d, axis=1)[:, None]
    return po.State(b, p, d, q, types)


def draw_lines(st, mol):
    out = []
    for i in range(st.n):
        r = st.box * (st.pos[i] - np.rint(st.pos[i]))
        out.append("%15.8e %15.8e %15.8e   %15.8e %15.8e %15.8e   %15.8e %15.8e %15.8e %d %d" % (
            tuple(r) + tuple(st.dir_raw[i]) + tuple(st.pd0[i]) + (0, mol[i])))
    return out


def chain_state(nch=300, eta=0.25, seed=13):
    g = sc_amd.GpuSystem(0)
    top = W.chain_topology(nch)
    box, pos, d, q, t = W.chain_lattice(nch, eta, seed)
    g.load_topology(top)
    g.set_config(box, pos, d, q, t)
    g.sweep(40, 0.7, 0.3, 20.0 / 180 * np.pi / 4, r=2, seed=3, counter=0)
    return g, po.Topo(top), t


def test_movie_frame_and_config_last():
    say("case: movie frame / config.last")
    g, tp, t = chain_state()
    b, p, d, q = g.get_config()
    st = po.State(b, p, d, q, t)
    st.dir_raw = d
    mol = np.zeros(len(t), np.int32)   # one molecule type
    frame = g.movie_frame(1234).splitlines()
    assert frame[0] == "%d" % len(t)
    assert frame[1] == "sweep 1234; box %.10f %.10f %.10f" % tuple(b)
    want = draw_lines(st, mol)
    assert frame[2:] == want
    last = g.write_config().splitlines()
    assert last[0] == "%15.8e %15.8e %15.8e" % tuple(b)
    assert last[1:] == want
    g.close()


def test_stat_and_energy_lines():
    say("case: stat.dat / energy.dat lines")
    g, tp, t = chain_state()
    st = dev_state(g, t)
    assert g.stat_line(77) == " 77; %.10f\n" % float(np.prod(st.box))
    line = g.energy_line(77)
    m = re.fullmatch(r" 77; (-?\d+\.\d{10})  (-?\d+\.\d{6}) \n", line)
    assert m, repr(line)
    e, al = float(m.group(1)), float(m.group(2))
    eo = po.energy_all(tp, st, use_cells=True)
    # Updater::alignmentOrder restated: pairs i < j with |r_ij|^2 < 1.5^2
    u = np.round(np.mod(st.pos, 1.0) * 4294967296.0).astype(np.uint64) % 4294967296
    pr = po.neighbor_pairs(u.astype(np.uint32), st.box, 1.5)
    d = st.pos[pr[:, 0]] - st.pos[pr[:, 1]]
    d = st.box * (d - np.rint(d))
    keep = np.einsum("ij,ij->i", d, d) < 2.25
    al_o = float(np.einsum("ij,ij->i", st.dir[pr[keep, 0]], st.dir[pr[keep, 1]]).sum())
    say("energy line %r; oracle E %.6f alignment %.6f" % (line, eo, al_o))
    assert abs(e - eo) <= 1e-5 * max(1.0, abs(eo))
    assert line.endswith("%f \n" % al_o) or abs(al - al_o) <= 2e-6
    assert abs(g.alignment_order() - al_o) <= 1e-9 * max(1.0, abs(al_o))
    g.close()


def cluster_texts(tp, st, sweep):
    """writeCluster's two frames from the oracle's labels and FP64 pair energies"""
    lab = po.cluster_labels(tp, st)
    roots = []
    members = {}
    for i, r in enumerate(lab):
        if r not in members:
            members[r] = []
            roots.append(r)
        members[r].append(i)
    sizes = [len(members[r]) for r in roots]
    maxc = max(sizes)
    stat = "Sweep: %d | Maximal size: %d\n" % (sweep, maxc)
    for k in range(maxc):
        stat += "%5d\t%5d\n" % (k + 1, sizes.count(k + 1))
    energies = []
    for r in roots:
        m = np.array(members[r], np.int32)
        if len(m) < 2:
            energies.append(0.0)
            continue
        i, j = np.nonzero(np.triu(np.ones((len(m), len(m)), bool), 1))
        energies.append(float(po.pair_energies(tp, st, m[i], m[j]).sum()))
    cl = "Sweep: %d | Number of clusters: %d\n" % (sweep, len(roots))
    for c, r in enumerate(roots):
        cl += "%3d(%f):" % (c + 1, energies[c]) + "".join("%5d" % (p + 1) for p in members[r]) + "\n"
    return stat, cl, energies


def _split_energies(text):
    """cluster.dat frame -> (text with energies blanked, energies)"""
    es = [float(x) for x in re.findall(r"\((-?\d+\.\d+)\):", text)]
    return re.sub(r"\((-?\d+\.\d+)\):", "(E):", text), es


@pytest.mark.parametrize("system", ["chains", "psc"])
def test_cluster_files(system):
    say("case: cluster files", system)
    if system == "chains":
        g, tp, t = chain_state(120, 0.25, 5)
    else:
        w = W.workload("c1_psc_1000", n_override=600)
        g = sc_amd.GpuSystem(0)
        box = w["box"] * (0.2 / 0.3) ** (1.0 / 3.0)
        g.load_topology(w["top"])
        pos, d, q = g.generate_rsa(w["types"], box, seed=8)
        g.set_config(box, pos, d, q, w["types"])
        g.sweep(60, 0.6, 0.3, 20.0 / 180 * np.pi / 4, r=2, seed=8, counter=0)
        tp, t = po.Topo(w["top"]), w["types"]
    st = dev_state(g, t)
    stat, cl = g.cluster_files(500)
    stat_o, cl_o, eo = cluster_texts(tp, st, 500)
    assert stat == stat_o
    a, ea = _split_energies(cl)
    b, eb = _split_energies(cl_o)
    assert a == b
    np.testing.assert_allclose(ea, eb, rtol=0, atol=1e-5)
    say("%d clusters, largest %s" % (len(ea), stat.splitlines()[0]))
    assert len(ea) < st.n
    g.close()
