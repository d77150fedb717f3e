"""Seeded starting states at the benchmark packing fractions.

Random sequential addition jams below the target packing fraction for long
rods (L=4 at eta 0.25, L=3 at eta 0.3), so the state is built in two steps:
RSA at a packing fraction it reaches quickly, then a gradual isotropic
compression of the box (positions are box fractions, as in SC) with
checkerboard MC sweeps between the steps to relax the overlaps the
compression creates.
"""
import numpy as np


def initial_state(g, w, seed=1, eta_rsa=0.18, steps=12, sweeps_per_step=4, relax_sweeps=20, temper=1.0,
                  transmx=0.3, rot_half=20.0 / 180 * np.pi / 4, log=None):
    """Load workload `w` (sc_amd.workloads.workload) into GpuSystem `g`.
    The compression sweeps visit every colour (visit factor 1, left set)."""
    g.load_topology(w["top"])
    g.set_visit_factor(1)
    box_t = np.asarray(w["box"], np.float64)
    eta = w["eta"]
    if eta <= eta_rsa:
        pos, d, q = g.generate_rsa(w["types"], box_t, seed)
        g.set_config(box_t, pos, d, q, w["types"])
        return
    box0 = box_t * (eta / eta_rsa) ** (1.0 / 3.0)
    pos, d, q = g.generate_rsa(w["types"], box0, seed)
    g.set_config(box0, pos, d, q, w["types"])
    counter = 10 ** 9
    for k in range(1, steps + 1):
        b = box0 * (box_t / box0) ** (k / steps)
        g.update_box(b)
        g.sweep(sweeps_per_step, temper, transmx, rot_half, r=2, seed=seed + 17, counter=counter)
        counter += sweeps_per_step
        if log:
            log("compress %d/%d box %.4f E %.3f" % (k, steps, b[0], g.allToAll()))
    g.sweep(relax_sweeps, temper, transmx, rot_half, r=2, seed=seed + 17, counter=counter)
