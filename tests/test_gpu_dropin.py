"""The reference's own program with the GPU backend at its plug point.

oracle/_ref/scdropin is scOOP 4.0 (robertvacha/SC) compiled from its
unmodified sources with one line changed, the typedef that selects the
energy backend (mc/totalenergycalculator.h:1227), to
`typedef GpuTotalE<Sim, Conf> TotalEnergyCalculator;`
(include/sc_gpu_totale.hpp), and linked to sc_amd/lib/libscamd.so: the
reference's Updater / MoveCreator drive every oneToAll, oneToAllTrial,
update(int) and allToAll of a plain NVT run through the C ABI.  The run's own
energy bookkeeping (updater.cpp:122-150, the drift it prints every tenth of
the run and at the end) must stay within 1e-8 |E| (every call prices a pair
term with the same FP32 rounding, so only FP64 sums differ), its ending
energy (the allToAll after the last sweep, updater.cpp:345-352) must match
the FP64 oracle on the configuration it writes (config.last), and
its acceptance rates must be those of the oracle's sampler on the same
system.  Skipped when the binary was not built (make -C oracle dropin needs
the reference sources).
"""
import os
import re
import subprocess
import tempfile

import numpy as np
import pytest

import sc_amd
from sc_amd import workloads as W
import pyoracle as po

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "oracle", "_ref", "scdropin")


@pytest.mark.skipif(not os.path.exists(BIN), reason="drop-in build of the reference absent (make -C oracle dropin)")
def test_reference_program_runs_on_the_gpu_backend():
    print("case: scOOP main() with GpuTotalE, 500 PSC eta 0.2, 200 sweeps", flush=True)
    w = W.workload("c1_psc_1000", n_override=500)
    h = sc_amd.GpuSystem(-1)
    h.load_topology(w["top"])
    pos, d, q = h.generate_rsa(w["types"], w["box"], seed=31)
    with tempfile.TemporaryDirectory() as dd:
        open(os.path.join(dd, "options"), "w").write(
            W.options_text(nsweeps=200, transmx=0.15, rotmx=20.0, temper=1.0, seed=4321))
        open(os.path.join(dd, "top.init"), "w").write(w["top"])
        open(os.path.join(dd, "config.init"), "w").write(W.config_text(w["box"], pos, d, q))
        r = subprocess.run([BIN], cwd=dd, capture_output=True, text=True, timeout=300)
        out = r.stdout
        print(out[-1500:], flush=True)
        print(r.stderr[-1500:], flush=True)
        assert r.returncode == 0
        assert "GpuTotalE" not in r.stderr   # no failed or uncovered call
        drifts = [float(x) for x in re.findall(r"particles: \d+ drift: (\S+)", out)]
        energies = [float(x) for x in re.findall(r"InteractionEnergy: (\S+?),", out)]
        assert len(drifts) == 10 and len(energies) == 10
        # the tenth report is printed at the start of the last sweep; the
        # ending energy is the allToAll after it
        e_last = float(re.search(r"Ending energy: (\S+)", out).group(1))
        drifts.append(float(re.search(r"Energy drift: (\S+)", out).group(1)))
        print("drifts %s; energies %.6f .. %.6f, ending %.6f" % (["%.2e" % x for x in drifts], energies[0],
                                                                energies[-1], e_last), flush=True)
        assert max(abs(x) for x in drifts) <= 1e-8 * max(1.0, abs(e_last))
        acc = {m.group(1): float(m.group(2)) for m in re.finditer(r"Single particle (translation|rotation):\s+(\d+)",
                                                                  out)}
        # the configuration the run wrote, in the reference's config.last format
        lines = open(os.path.join(dd, "config.last")).read().split("\n")
        box = np.array([float(x) for x in lines[0].split()])
        rows = np.array([[float(x) for x in ln.split()[:9]] for ln in lines[1:] if ln.strip()])
    st = po.State.from_real(box, rows[:, 0:3], rows[:, 3:6], rows[:, 6:9], w["types"])
    tp = po.Topo(w["top"])
    eo = po.energy_all(tp, st, use_cells=False)
    print("ending energy %.6f, oracle on config.last %.6f; acceptance %s" % (e_last, eo, acc), flush=True)
    # the run prints 6 significant digits (cout's default precision)
    ulp6 = 10.0 ** (np.floor(np.log10(abs(eo))) - 5) if eo != 0 else 1e-6
    assert abs(e_last - eo) <= 1e-5 * abs(eo) + 0.5 * ulp6
    # the oracle's sampler on the same start: translation ~51%, rotation ~84%
    _, _, stats, _ = po.mc_run(tp, po.State(w["box"], pos, d, q, w["types"]), 200, 1.0, 0.3,
                               20.0 / 180 * np.pi / 4, seed=3, energies=False)
    at, ar = 100.0 * stats[0] / stats[1], 100.0 * stats[2] / stats[3]
    print("oracle acceptance %.1f / %.1f" % (at, ar), flush=True)
    assert abs(acc["translation"] - at) <= 2.0 and abs(acc["rotation"] - ar) <= 2.0
