"""Writes the golden vectors under tests/golden/ from the compiled reference.

Runs in the build container only: it needs oracle/_ref/scref, built by
`make -C oracle ref` from the unmodified robertvacha/SC sources.  The fixtures
are data (inputs and the reference's outputs); no reference source is copied.

    python tests/golden/make_golden.py

Sets:
  topo_golden.json    ia_params of a 12-type topology (every geotype)
  pair_golden.npz     p2p(i,j) for a dense mixed configuration (all pairs
                      inside sqmaxcut) and for an Interactions-test-style grid
                      (particle 0 at the centre, the other type on a grid,
                      Tests/Interactions_tests/main.cpp) for every type pair
  total_golden.npz    allToAll and oneToAll(i) of overlap-free PSC, CPSC and
                      TCHPSC systems at the benchmark packing fractions
  derived_golden.npz  patch sides / chiral axes / second patch after
                      Particle::init for the mixed configuration
  mindist_golden.npz  EPatch::minDistSegments on random and degenerate inputs
  chain_golden.npz    p2p(i,j) over all ordered pairs, allToAll and oneToAll(i)
                      of chain systems covering every bond and angle term
                      (bond1, bond2, bondd, bondh, angle1, angle2; spheres at
                      either end, PSC and CPSC segments), and mol2others /
                      mol2othersTrial of each of their molecules

`python tests/golden/make_golden.py chains` rewrites chain_golden.npz only.
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, ROOT)
import pyoracle as po  # noqa: E402

MIXED_TYPES = """[Types]
A 1 PSC 1.0 1.0 1.8 0.2 80.0 5.0 3.0 0.3
B 2 CPSC 1.5 1.1 1.8 0.3 90.0 5.0 3.0 0.0
C 3 CHPSC 1.0 1.0 1.6 0.2 80.0 5.0 3.0 0.2 10.0
D 4 CHCPSC 1.0 1.0 1.8 0.2 90.0 5.0 3.0 0.0 -10.0
E 5 TPSC 1.0 1.0 1.8 0.2 80.0 5.0 3.0 0.0 90.0 80.0 5.0
F 6 TCPSC 1.2 1.0 1.8 0.25 90.0 5.0 3.0 -0.2 120.0 60.0 10.0
G 7 TCHPSC 1.0 1.0 1.0 0.2 80.0 5.0 3.0 0.0 90.0 80.0 5.0 10.0
H 8 TCHCPSC 1.0 0.9 1.8 0.2 90.0 5.0 3.0 0.0 90.0 80.0 5.0 25.0
I 9 SPN 1.0 1.2
J 10 SPA 1.0 1.2 1.6 0.3
K 11 SCN 1.0 1.0 3.0
L 12 SCA 1.0 1.0 1.8 0.2 3.0
"""
NAMES = "ABCDEFGHIJKL"


def mixed_top(counts):
    t = MIXED_TYPES + "[Molecules]\n"
    for k, nm in enumerate(NAMES):
        t += "%s: {\nparticles: %d\n}\n" % (nm, k + 1)
    t += "[System]\n"
    for nm, c in zip(NAMES, counts):
        t += "%s %d\n" % (nm, c)
    return t


def two_type_top(ta, tb, nb):
    lines = [l for l in MIXED_TYPES.splitlines()[1:]]
    la, lb = lines[ta - 1], lines[tb - 1]
    if ta == tb:
        t = "[Types]\n%s\n[Molecules]\nX: {\nparticles: %d\n}\n[System]\nX %d\n" % (la, ta, 1 + nb)
    else:
        t = "[Types]\n%s\n%s\n[Molecules]\nX: {\nparticles: %d\n}\nY: {\nparticles: %d\n}\n[System]\nX 1\nY %d\n" % (
            la, lb, ta, tb, nb)
    return t


def fib_dirs(n):
    """Fibonacci sphere as in Tests/Interactions_tests/main.cpp:137-154"""
    off = 2.0 / n
    inc = np.pi * (3.0 - np.sqrt(5.0))
    out = []
    for i in range(n):
        z = ((i * off) - 1) + off / 2
        r = np.sqrt(1 - z * z)
        out.append((np.cos(i * inc) * r, np.sin(i * inc) * r, z))
    return np.array(out)


def perp_unit(d, ang):
    ref = np.array([1.0, 0, 0]) if abs(d[0]) < 0.9 else np.array([0, 1.0, 0])
    e1 = ref - d * np.dot(ref, d)
    e1 /= np.linalg.norm(e1)
    e2 = np.cross(d, e1)
    return np.cos(ang) * e1 + np.sin(ang) * e2


def pairs_query(top, box, pos, d, q, pi, pj):
    txt = po.ref_query("pairs", top, box, pos, d, q, "\n".join("%d %d" % (a, b) for a, b in zip(pi, pj)) + "\n")
    return np.array([float(x) for x in txt.split()])


def make_mixed(rng):
    per = 22
    counts = [per] * 12
    top = mixed_top(counts)
    n = per * 12
    types = np.repeat(np.arange(1, 13), per).astype(np.int32)
    box = np.array([8.3, 8.9, 9.4])
    pos = rng.random((n, 3)) * box - box / 2
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    q = rng.normal(size=(n, 3))
    tp = po.Topo(top)
    st = po.State.from_real(box, pos, d, q, types)
    pi, pj = [], []
    for i in range(n):
        for j in range(n):
            if i == j:
                continue
            dd = st.pos[i] - st.pos[j]
            dd = box * (dd - np.rint(dd))
            if dd @ dd <= tp.sqmaxcut:
                pi.append(i)
                pj.append(j)
    pi, pj = np.array(pi, np.int32), np.array(pj, np.int32)
    e = pairs_query(top, box, pos, d, q, pi, pj)
    return top, box, pos, d, q, types, pi, pj, e


def make_grid():
    """particle 0 (type a) at the box centre, dir (1,0,0), patch (0,1,0);
    type-b particles on the offsets {-4,-1,2}^3 \\ 0 of a 10^3 box with
    Fibonacci directions and rotated patches (Interactions_tests layout)."""
    offs = [-4.0, -1.0, 2.0]
    sites = [(x, y, z) for x in offs for y in offs for z in offs if (x, y, z) != (0, 0, 0)]
    dirs = np.vstack([fib_dirs(4), [[1, 0, 0], [0, 0, 1]]])
    out = []
    for ta in range(1, 13):
        for tb in range(1, 13):
            pos, dd, qq = [[5.0, 5.0, 5.0]], [[1.0, 0, 0]], [[0, 1.0, 0]]
            for s in sites:
                for di, dv in enumerate(dirs):
                    ang = 0.7 * di
                    pos.append([5.0 + s[0], 5.0 + s[1], 5.0 + s[2]])
                    dd.append(dv)
                    qq.append(perp_unit(dv, ang))
            pos, dd, qq = np.array(pos), np.array(dd), np.array(qq)
            nb = len(pos) - 1
            top = two_type_top(ta, tb, nb)
            box = np.array([10.0, 10.0, 10.0])
            pj = np.arange(1, nb + 1, dtype=np.int32)
            pi = np.zeros(nb, np.int32)
            e = pairs_query(top, box, pos, dd, qq, pi, pj)
            out.append((ta, tb, top, pos, dd, qq, e))
    return out


def make_totals():
    sys.path.insert(0, ROOT)
    import sc_amd
    from sc_amd import workloads as W
    sets = []
    for key, n, eta, seed in (("psc", 400, 0.20, 11), ("cpsc", 400, 0.30, 12), ("tchpsc", 400, 0.25, 13)):
        lines = [W.TYPES[key]]
        f = lines[0].split()
        v = W.particle_volume(float(f[4]), float(f[9]))
        L = W.box_for([n], [v], eta)
        box = np.array([L, L, L])
        top = W.top_init(lines, [n])
        types = np.full(n, int(f[1]), np.int32)
        h = sc_amd.GpuSystem(-1)
        h.load_topology(top)
        pos_f, d, q = h.generate_rsa(types, box, seed)
        h.close()
        real = (pos_f - np.rint(pos_f)) * box
        txt = po.ref_query("total", top, box, real, d, q)
        vals = np.array([float(x) for x in txt.split()])
        sets.append((key, top, box, real, d, q, types, vals[0], vals[1:]))
    return sets


def make_mindist(rng):
    cases = []
    for _ in range(400):
        a = rng.normal(size=3); a /= np.linalg.norm(a)
        b = rng.normal(size=3); b /= np.linalg.norm(b)
        h1, h2 = rng.uniform(0.0, 2.5, 2)
        r = rng.normal(size=3) * 2.0
        cases.append(np.concatenate([a, b, [h1, h2], r]))
    for _ in range(150):   # (anti)parallel and near-parallel
        a = rng.normal(size=3); a /= np.linalg.norm(a)
        b = a * (1 if rng.random() < 0.5 else -1) + rng.normal(size=3) * 10 ** rng.uniform(-12, -3)
        b /= np.linalg.norm(b)
        h1, h2 = rng.uniform(0.2, 2.5, 2)
        r = a * rng.uniform(-5, 5) + rng.normal(size=3) * rng.uniform(0, 2)
        cases.append(np.concatenate([a, b, [h1, h2], r]))
    for _ in range(50):    # zero length pieces
        a = rng.normal(size=3); a /= np.linalg.norm(a)
        b = rng.normal(size=3); b /= np.linalg.norm(b)
        cases.append(np.concatenate([a, b, [0.0, rng.uniform(0, 2)], rng.normal(size=3)]))
    return np.array(cases)


def run_mindist(cases):
    import subprocess
    import tempfile
    with tempfile.TemporaryDirectory() as d:
        inp = os.path.join(d, "in.txt")
        with open(inp, "w") as f:
            f.write("\n".join(" ".join("%.17g" % v for v in c) for c in cases) + "\n")
        out = os.path.join(d, "out.txt")
        subprocess.run([po.REF_BIN, "mindist", out, inp], check=True, cwd=d, stdout=subprocess.DEVNULL)
        return np.loadtxt(out).reshape(-1, 3)


CHAIN_TYPES = """[Types]
H 4 SPA 1.0 1.2 1.0 0.2
T 5 PSC 1.0 1.0 1.0 0.2 80.0 5.0 3.0 0.0
C 6 CPSC 1.0 1.0 1.2 0.2 90.0 5.0 3.0 0.0
"""
SP_TYPES = """[Types]
N 7 SPN 1.0 0.95
S 8 SPA 1.0 1.0 1.12246205 1.6
"""
# name: (types text, molecules text, [(molecule, count)], box edge)
CHAIN_SETS = {
    "c5": (CHAIN_TYPES, "CH: {\nbond1: 100 1\nparticles: 4\nparticles: 5\nparticles: 5\n}\n",
           [("CH", 16)], 12.0),
    "angles": (CHAIN_TYPES, "K: {\nbond1: 80 0.8\nbond2: 5 3.0\nangle1: 20 30\nangle2: 15 60\nparticles: 4\n"
                            "particles: 5\nparticles: 6\nparticles: 6\nparticles: 4\n}\nA: {\nparticles: 4\n}\n",
               [("K", 8), ("A", 6)], 12.0),
    "bondd": (CHAIN_TYPES, "D: {\nbondd: 50 0.5\nparticles: 5\nparticles: 5\nparticles: 6\n}\n", [("D", 12)], 11.0),
    "bondh": (CHAIN_TYPES, "B: {\nbondh: 50 0.3\nparticles: 4\nparticles: 5\nparticles: 5\nparticles: 4\n}\n",
              [("B", 10)], 11.0),
    "sp": (SP_TYPES, "W: {\nbond1: 90 1.0\nbond2: 10 2.2\nparticles: 7\nparticles: 8\nparticles: 8\n}\n",
           [("W", 20)], 9.0),
}


def chain_top(name):
    types, mols, system, _ = CHAIN_SETS[name]
    return types + "[Molecules]\n" + mols + "[System]\n" + "".join("%s %d\n" % s for s in system)


def _seg_dist(a, b):
    """closest distance of two beads (pos, dir, half length)"""
    r = b[0] - a[0]
    v = po.min_dist_segments(a[1], b[1], a[2], b[2], -r)
    return np.linalg.norm(v)


def make_chain_config(name, rng):
    """Chains grown bead by bead: each bond vector between the attachment
    points (HarmonicSc, paire.h:239-280: bead i's -dir end to bead i+1's +dir
    end, spheres at sp along dir; mod = bondd/bondh extension) has a random
    direction and a length within +-35 % of equilibrium; beads keep a segment
    distance >= 0.6 sigma from their bonded neighbour, >= 0.8 sigma from the
    rest of their chain and >= 0.85 sigma from earlier chains (retries), so
    every energy is finite and moderate and the bond terms are not lost in
    the rounding of a large repulsion."""
    tp = po.Topo(chain_top(name))
    box = np.full(3, CHAIN_SETS[name][3])
    types = tp.types
    n = len(types)
    half = {t: tp.ia(t, t).half_len[0] for t in set(types.tolist())}
    sph = {t: tp.ia(t, t).geotype[0] >= 30 for t in set(types.tolist())}
    pos, dirs = np.zeros((n, 3)), np.zeros((n, 3))
    i = 0
    while i < n:
        mt, mp = tp.chain[i]
        m = tp.t.mol[mt]
        size = m.size
        for attempt in range(200):
            beads = []
            for k in range(size):
                t = types[i + k]
                d = rng.normal(size=3)
                d /= np.linalg.norm(d)
                if k == 0:
                    c = rng.random(3) * box
                else:
                    pc, pd, pt = beads[-1]
                    # previous bead's attachment towards its next partner
                    if sph[pt]:
                        sp = m.bondheq if m.bondhc >= 0 else 0.0
                        a0 = pc + pd * sp
                    else:
                        mod1 = m.bondheq if m.bondhc >= 0 else 0.0
                        a0 = pc - pd * (half[pt] + mod1)
                    eq = m.bond1eq if m.bond1c >= 0 else (0.3 if m.bonddc >= 0 or m.bondhc >= 0 else 1.0)
                    u = rng.normal(size=3)
                    u /= np.linalg.norm(u)
                    a1 = a0 + u * eq * rng.uniform(0.65, 1.35) if eq > 0 else a0 + u * rng.uniform(0.0, 0.4)
                    if not sph[t]:   # body runs from the +dir end along -dir: away from the bond
                        d = -u + 0.6 * d
                        d /= np.linalg.norm(d)
                    if sph[t]:
                        sp = m.bondheq if m.bondhc >= 0 else 0.0
                        c = a1 - d * sp
                    else:
                        mod0 = m.bonddeq if m.bonddc >= 0 else (m.bondheq if m.bondhc >= 0 else 0.0)
                        c = a1 - d * (half[t] + mod0)
                beads.append((c, d, t))
            ok = True
            for k, (c, d, t) in enumerate(beads):
                for k2 in range(k):
                    c2, d2, t2 = beads[k2]
                    if _seg_dist((c2, d2, half[t2]), (c, d, half[t])) < (0.6 if k2 == k - 1 else 0.8):
                        ok = False
                for j in range(i):
                    r = c - pos[j]
                    r -= box * np.rint(r / box)
                    if np.linalg.norm(r) > 6.0:
                        continue
                    if _seg_dist((pos[j], dirs[j], half[types[j]]), (pos[j] + r, d, half[t])) < 0.85:
                        ok = False
                        break
                if not ok:
                    break
            if ok:
                break
        if not ok:
            raise RuntimeError("could not place chain at particle %d" % i)
        for k, (c, d, t) in enumerate(beads):
            pos[i + k] = c - box * np.floor(c / box)
            dirs[i + k] = d
        i += size
    q = rng.normal(size=(n, 3))
    return tp, box, pos - box / 2, dirs, q, types


def make_chains(rng):
    kw = {}
    for name in CHAIN_SETS:
        tp, box, pos, d, q, types = make_chain_config(name, rng)
        top = chain_top(name)
        n = len(types)
        pi, pj = np.nonzero(~np.eye(n, dtype=bool))
        pi, pj = pi.astype(np.int32), pj.astype(np.int32)
        e = pairs_query(top, box, pos, d, q, pi, pj)
        txt = po.ref_query("total", top, box, pos, d, q)
        vals = np.array([float(x) for x in txt.split()])
        mol = np.array([[float(x) for x in l.split()] for l in
                        po.ref_query("molecules", top, box, pos, d, q).strip().splitlines()])
        kw.update({name + "_top": top, name + "_box": box, name + "_pos": pos, name + "_dir": d, name + "_pd0": q,
                   name + "_types": types, name + "_i": pi, name + "_j": pj, name + "_e": e,
                   name + "_eall": vals[0], name + "_eone": vals[1:],
                   name + "_mfirst": mol[:, 0].astype(np.int32), name + "_msize": mol[:, 1].astype(np.int32),
                   name + "_emol": mol[:, 2], name + "_emol_trial": mol[:, 3]})
        print("chains %s: %d particles, %d pairs, E_all %.6f" % (name, n, len(e), vals[0]))
    np.savez_compressed(os.path.join(HERE, "chain_golden.npz"), **kw)


def main():
    if not po.ref_available():
        sys.exit("oracle/_ref/scref missing: run `make -C oracle ref` first")
    if sys.argv[1:] == ["chains"]:
        make_chains(np.random.default_rng(20241101))
        return
    rng = np.random.default_rng(20240601)
    # topology dump
    top = mixed_top([1] * 12)
    box = np.array([10.0, 10, 10])
    txt = po.ref_query("topo", top, box, np.zeros((12, 3)) + 5.0, np.tile([1.0, 0, 0], (12, 1)),
                       np.tile([0, 1.0, 0], (12, 1)))
    rows = {}
    extra = {}
    for line in txt.strip().splitlines():
        f = line.split()
        if f[0] == "sqmaxcut":
            extra = {"sqmaxcut": float(f[1]), "maxcut": float(f[3])}
            continue
        rows["%s,%s" % (f[0], f[1])] = [float(x) for x in f[2:]]
    with open(os.path.join(HERE, "topo_golden.json"), "w") as f:
        json.dump({"top": top, "rows": rows, **extra,
                   "columns": "geotype0 geotype1 sigma epsilon A B pdis pswitch pswitchINV rcut rcutSq rcutwca "
                              "rcutwcaSq len0 len1 half_len0 half_len1 parallel volume pcangl0-3 pcanglsw0-3 "
                              "pcoshalfi0-3 psinhalfi0-3 csecpatchrot0-1 ssecpatchrot0-1 chiral_cos0-1 "
                              "chiral_sin0-1 exclude"}, f, indent=0)
    # pairs
    mtop, mbox, mpos, md, mq, mtypes, pi, pj, e = make_mixed(rng)
    grid = make_grid()
    g_ta = np.array([g[0] for g in grid], np.int32)
    g_tb = np.array([g[1] for g in grid], np.int32)
    np.savez_compressed(os.path.join(HERE, "pair_golden.npz"),
                        mixed_top=mtop, mixed_box=mbox, mixed_pos=mpos, mixed_dir=md, mixed_pd0=mq,
                        mixed_types=mtypes, mixed_i=pi, mixed_j=pj, mixed_e=e,
                        grid_ta=g_ta, grid_tb=g_tb, grid_pos=grid[0][3], grid_dir=grid[0][4], grid_pd0=grid[0][5],
                        grid_e=np.array([g[6] for g in grid]),
                        grid_tops=np.array([g[2] for g in grid]))
    # derived vectors of the mixed configuration
    dtxt = po.ref_query("derived", mtop, mbox, mpos, md, mq)
    der = np.array([[float(x) for x in l.split()] for l in dtxt.strip().splitlines()])
    np.savez_compressed(os.path.join(HERE, "derived_golden.npz"), derived=der)
    # totals
    tot = make_totals()
    kw = {}
    for key, top_t, box_t, real, d, q, types, e_all, e_one in tot:
        kw.update({key + "_top": top_t, key + "_box": box_t, key + "_pos": real, key + "_dir": d, key + "_pd0": q,
                   key + "_types": types, key + "_eall": e_all, key + "_eone": e_one})
    np.savez_compressed(os.path.join(HERE, "total_golden.npz"), **kw)
    # minDistSegments
    cases = make_mindist(rng)
    res = run_mindist(cases)
    np.savez_compressed(os.path.join(HERE, "mindist_golden.npz"), cases=cases, result=res)
    print("pairs mixed %d (nonzero %d), grid %d; totals %s" % (
        len(e), np.count_nonzero(e), sum(len(g[6]) for g in grid), [(t[0], t[7]) for t in tot]))
    make_chains(np.random.default_rng(20241101))


if __name__ == "__main__":
    main()
