import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: long CPU test")


@pytest.fixture(scope="session")
def golden_dir():
    return os.path.join(ROOT, "tests", "golden")


@pytest.fixture(scope="session")
def ensemble_system():
    """500 PSC at eta 0.2, T*=1 and the CPU oracle's sequential Metropolis
    reference: E/N, nematic S2 and mean cluster size at the end of each of 40
    blocks of 25 sweeps after 200 sweeps.  Returns (workload, (pos, dir, pd0)
    of the start, [40 x 3] observables, rot_half)."""
    import numpy as np
    import sc_amd
    from sc_amd import workloads as W
    import pyoracle as po
    w = W.workload("c1_psc_1000", n_override=500)
    g = sc_amd.GpuSystem(0)
    g.load_topology(w["top"])
    pos, d, q = g.generate_rsa(w["types"], w["box"], seed=31)
    g.set_config(w["box"], pos, d, q, w["types"])
    b, p, dd, qq = g.get_config()
    g.close()
    tp, st = po.Topo(w["top"]), po.State(b, p, dd, qq, w["types"])
    rot_half = 20.0 / 180 * np.pi / 4
    stc, _, _, _ = po.mc_run(tp, st, 200, 1.0, 0.3, rot_half, seed=17)
    ec = []
    cur = stc
    for blk in range(40):
        cur, _, _, ef = po.mc_run(tp, cur, 25, 1.0, 0.3, rot_half, seed=100 + blk)
        s2 = po.nematic_s2(cur.dir)
        ncl = len(np.unique(po.cluster_labels(tp, cur)))
        ec.append((ef / st.n, s2, st.n / ncl))
    return w, (pos, d, q), np.array(ec), rot_half
