"""Synthetic inputs in SC's own file formats for the benchmark systems.

Each workload is a top.init text, a particle count and a packing fraction;
the box edge follows from eta = N v / V with the geometric volume of a
spherocylinder v = pi sigma^3/6 + pi sigma^2 L/4 (NOT Ia_param::volume, which
holds half the cylinder, scOOP/mc/inicializer.cpp:842).

Parameter mapping of the patch model: "attraction depth 1 kT, range 1 sigma,
switch 0.2 sigma" is taken literally as EPSILON=1, ATTRACT_DIST=1.0,
ATTRACT_SWITCH=0.2 (rcut = 1.2 sigma); the patch "half-angle 40 deg" is
PATCH_ANGLE=80 (the reference reads the full opening angle,
inicializer.cpp:794) with PATCH_SWITCH=5.
"""
import numpy as np

TYPES = {
    # NAME NUMBER GEOTYPE EPS SIGMA ATTRACT_DIST ATTRACT_SWITCH PATCH_ANGLE PATCH_SWITCH L PARALLEL [...]
    "psc": "A 1 PSC 1.0 1.0 1.0 0.2 80.0 5.0 3.0 0.0",
    "cpsc": "B 2 CPSC 1.0 1.0 1.0 0.2 90.0 5.0 3.0 0.0",
    # second patch rotated 90 deg, chirality 10 deg, length 4
    "tchpsc": "C 3 TCHPSC 1.0 1.0 1.0 0.2 80.0 5.0 4.0 0.0 90.0 80.0 5.0 10.0",
}


def top_init(type_lines, counts):
    """top.init text for atomic molecules: type_lines [(text)], counts [n]."""
    out = ["[Types]"]
    out += type_lines
    out.append("[Molecules]")
    names = []
    for tl in type_lines:
        f = tl.split()
        names.append(f[0])
        out.append("%s: {\nparticles: %s\n}" % (f[0], f[1]))
    out.append("[System]")
    for nm, c in zip(names, counts):
        out.append("%s %d" % (nm, c))
    return "\n".join(out) + "\n"


def particle_volume(sigma, length):
    return np.pi * sigma ** 3 / 6.0 + np.pi * sigma ** 2 * length / 4.0


def box_for(n_by_type, vols, eta):
    v = sum(n * vv for n, vv in zip(n_by_type, vols))
    return (v / eta) ** (1.0 / 3.0)


WORKLOADS = {
    # name: (type keys, counts, packing fraction, length)
    "c1_psc_1000": (["psc"], [1000], 0.20),
    "c2_cpsc_32768": (["cpsc"], [32768], 0.30),
    "c4_tchpsc_65536": (["tchpsc"], [65536], 0.25),
    "psc_65536": (["psc"], [65536], 0.25),
}


def workload(name, n_override=None):
    keys, counts, eta = WORKLOADS[name]
    if n_override is not None:
        counts = [n_override]
    lines = [TYPES[k] for k in keys]
    vols = []
    for l in lines:
        f = l.split()
        vols.append(particle_volume(float(f[4]), float(f[9]) if len(f) > 9 else 0.0))
    L = box_for(counts, vols, eta)
    types = np.concatenate([np.full(c, int(l.split()[1]), np.int32) for l, c in zip(lines, counts)])
    return {"top": top_init(lines, counts), "box": np.array([L, L, L]), "types": types, "eta": eta, "name": name}


# Self-assembling chains: one SPA head (sigma 1.2) bonded to two PSC tails
# (L = 3) by harmonic bonds k = 100 kT/sigma^2, l0 = 1 sigma (bond1: the
# head centre to the first tail's +dir end, each tail's -dir end to the next
# tail's +dir end, HarmonicSc, paire.h:239-280).
CHAIN_TYPES = ["H 1 SPA 1.0 1.2 1.0 0.2", "T 2 PSC 1.0 1.0 1.0 0.2 80.0 5.0 3.0 0.0"]
CHAIN_MOLECULE = "CH: {\nbond1: 100 1.0\nparticles: 1\nparticles: 2\nparticles: 2\n}"


def chain_topology(nchains):
    return "\n".join(["[Types]"] + CHAIN_TYPES + ["[Molecules]", CHAIN_MOLECULE, "[System]", "CH %d" % nchains]) + "\n"


def chain_volume():
    """geometric volume of one head + two tails"""
    return particle_volume(1.2, 0.0) + 2 * particle_volume(1.0, 3.0)


def chain_lattice(nchains, eta, seed):
    """Straight chains (head, tail, tail along +-z, bonds near their rest
    length) on a lattice of a cubic box at packing fraction eta, with random
    sense, patch directions, in-plane jitter and small tilts.  Returns (box, box-fraction
    positions, dirs, patch dirs, types) in [System] order."""
    rng = np.random.default_rng(seed)
    L = (nchains * chain_volume() / eta) ** (1.0 / 3.0)
    span = 9.1                                # head top (+0.6) to the far tail's cap (-8.5)
    nz = max(1, int(L // (span + 0.2)))
    nxy = int(np.ceil(np.sqrt(nchains / nz)))
    a_xy, a_z = L / nxy, L / nz
    if a_xy < 1.25:
        raise ValueError("packing fraction too high for the chain lattice")
    sites = [(ix, iy, iz) for iz in range(nz) for iy in range(nxy) for ix in range(nxy)]
    pick = rng.permutation(len(sites))[:nchains]
    pos, dirs, pd = [], [], []
    for k in pick:
        ix, iy, iz = sites[k]
        u = 1.0 if rng.random() < 0.5 else -1.0
        jit = (a_xy - 1.22) / 2
        x = (ix + 0.5) * a_xy + rng.uniform(-jit, jit)
        y = (iy + 0.5) * a_xy + rng.uniform(-jit, jit)
        zb = iz * a_z + (a_z - span) / 2
        c0 = zb + (8.5 if u > 0 else 0.6)     # head centre
        for off in (0.0, -2.5, -6.5):           # head, tail 1, tail 2 centres along u
            # small lateral jitter and tilt: no bead sits exactly on another's axis
            jx, jy = rng.normal(0.0, 0.03, 2)
            pos.append([x + jx, y + jy, c0 + u * off])
            d = np.array([rng.normal(0.0, 0.03), rng.normal(0.0, 0.03), u])
            d /= np.linalg.norm(d)
            dirs.append(d)
            ph = rng.uniform(0, 2 * np.pi)
            q = np.array([np.cos(ph), np.sin(ph), 0.0])
            q -= d * (q @ d)
            pd.append(q / np.linalg.norm(q))
    pos = np.array(pos) / L
    types = np.tile(np.array([1, 2, 2], np.int32), nchains)
    return np.full(3, L), pos - np.floor(pos), np.array(dirs), np.array(pd), types


OPTIONS = """ptype = 1
press = 0
paralpress = 0
shave = 0
nequil = 0
adjust = 0
nsweeps = {nsweeps}
paramfrq = 0
report = 0
nrepchange = 0
nGrandCanon = 0
nClustMove = 0
movie = 0
chainprob = 0
transmx = {transmx}
rotmx = {rotmx}
edge_mx = 0.0
chainmmx = 0.0
chainrmx = 0.0
temper = {temper}
paraltemper = {temper}
wlm = 0
wlmtype = 0
switchprob = 0.00
pairlist_update = 10
seed = {seed}
write_cluster = 0
"""


def options_text(nsweeps=1000, transmx=0.15, rotmx=20.0, temper=1.0, seed=12345):
    return OPTIONS.format(nsweeps=nsweeps, transmx=transmx, rotmx=rotmx, temper=temper, seed=seed)


def config_text(box, pos_frac, dir, pd0):
    """config.init text; positions written in real units centred on the box."""
    lines = ["%.17g %.17g %.17g" % tuple(box)]
    p = np.asarray(pos_frac).reshape(-1, 3)
    real = (p - np.rint(p)) * np.asarray(box)
    for r, d, q in zip(real, np.asarray(dir).reshape(-1, 3), np.asarray(pd0).reshape(-1, 3)):
        lines.append("%.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g 0" % (tuple(r) + tuple(d) + tuple(q)))
    return "\n".join(lines) + "\n"
