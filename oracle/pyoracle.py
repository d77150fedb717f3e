"""ctypes binding of the FP64 CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker / timed CPU baseline.  The product
package (sc_amd) never imports this module.

Also holds the reference-harness runner (oracle/_ref/scref, built from the
reference sources by oracle/Makefile) used to write golden vectors.
"""
import ctypes
import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_BIN = os.path.join(HERE, "_ref", "scref")

MAXT = 40
GEOTYPES = {"SCN": 10, "SCA": 11, "PSC": 12, "CPSC": 13, "CHPSC": 14, "CHCPSC": 15, "TPSC": 16,
            "TCPSC": 17, "TCHPSC": 18, "TCHCPSC": 19, "SP": 30, "SPN": 30, "SPA": 31}


class _IA(ctypes.Structure):
    _fields_ = [("geotype", ctypes.c_int * 2),
                ("sigma", ctypes.c_double), ("epsilon", ctypes.c_double), ("A", ctypes.c_double),
                ("B", ctypes.c_double), ("pdis", ctypes.c_double), ("pswitch", ctypes.c_double),
                ("pswitchINV", ctypes.c_double),
                ("pangl", ctypes.c_double * 4), ("panglsw", ctypes.c_double * 4),
                ("pcangl", ctypes.c_double * 4), ("pcanglsw", ctypes.c_double * 4),
                ("rcut", ctypes.c_double), ("rcutSq", ctypes.c_double), ("rcutwca", ctypes.c_double),
                ("rcutwcaSq", ctypes.c_double),
                ("pcoshalfi", ctypes.c_double * 4), ("psinhalfi", ctypes.c_double * 4),
                ("csecpatchrot", ctypes.c_double * 2), ("ssecpatchrot", ctypes.c_double * 2),
                ("len", ctypes.c_double * 2), ("half_len", ctypes.c_double * 2),
                ("chiral_cos", ctypes.c_double * 2), ("chiral_sin", ctypes.c_double * 2),
                ("parallel", ctypes.c_double), ("volume", ctypes.c_double), ("exclude", ctypes.c_int)]


class _Vec(ctypes.Structure):
    _fields_ = [("x", ctypes.c_double), ("y", ctypes.c_double), ("z", ctypes.c_double)]


MAXMT = 100
MOL_KEYS = ("bond1c", "bond1eq", "bond2c", "bond2eq", "bonddc", "bonddeq", "bondhc", "bondheq",
            "angle1c", "angle1eq", "angle2c", "angle2eq")
DEGTORAD = 0.017453292519943295769236907683   # macros.h:67


class _Mol(ctypes.Structure):
    _fields_ = [("size", ctypes.c_int)] + [(k, ctypes.c_double) for k in MOL_KEYS]


class _Topo(ctypes.Structure):
    _fields_ = [("ia", (_IA * MAXT) * MAXT), ("sqmaxcut", ctypes.c_double), ("maxcut", ctypes.c_double),
                ("mol", _Mol * MAXMT), ("nmol", ctypes.c_int), ("nchain", ctypes.c_int),
                ("chain", ctypes.POINTER(ctypes.c_int))]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle not built: run `make -C oracle oracle`")
        L = ctypes.CDLL(LIB_PATH)
        D = ctypes.POINTER(ctypes.c_double)
        I = ctypes.POINTER(ctypes.c_int)
        U = ctypes.POINTER(ctypes.c_uint32)
        T = ctypes.POINTER(_Topo)
        L.orc_topo_clear.argtypes = [T]
        L.orc_topo_add_type.argtypes = [T, ctypes.c_int, ctypes.c_int, D, ctypes.c_int]
        L.orc_topo_finalize.argtypes = [T, ctypes.c_void_p]
        L.orc_soa_pair_energies.argtypes = [T, D, ctypes.c_int, D, D, D, ctypes.c_void_p, I, ctypes.c_int, I, I, D]
        L.orc_topo_set_mol.argtypes = [T, ctypes.c_int, ctypes.c_int, D]
        L.orc_topo_set_chain.argtypes = [T, ctypes.c_int, ctypes.c_void_p]
        L.orc_soa_bond_energies.argtypes = [T, D, ctypes.c_int, D, D, D, ctypes.c_void_p, I, ctypes.c_int, I, I, D]
        L.orc_soa_energy_all.argtypes = [T, D, ctypes.c_int, D, D, D, ctypes.c_void_p, I, ctypes.c_int]
        L.orc_soa_energy_all.restype = ctypes.c_double
        L.orc_soa_energy_one_each.argtypes = [T, D, ctypes.c_int, D, D, D, ctypes.c_void_p, I, ctypes.c_int, I, D]
        L.orc_soa_initconfig.argtypes = [ctypes.c_int, D, D, D, D]
        L.orc_soa_derived.argtypes = [T, ctypes.c_int, D, D, ctypes.c_void_p, I, D]
        L.orc_cell_list.argtypes = [ctypes.c_int, U, U, U, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                    ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                    ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32),
                                    ctypes.POINTER(ctypes.c_int32)]
        L.orc_neighbor_pairs.argtypes = [ctypes.c_int, U, U, U, D, ctypes.c_double, ctypes.c_void_p, ctypes.c_long]
        L.orc_neighbor_pairs.restype = ctypes.c_long
        L.orc_min_dist_segments.argtypes = [_Vec, _Vec, ctypes.c_double, ctypes.c_double, _Vec]
        L.orc_min_dist_segments.restype = _Vec
        L.orc_mc_run.argtypes = [T, D, ctypes.c_int, D, D, D, ctypes.c_void_p, I, ctypes.c_long, ctypes.c_double,
                                 ctypes.c_double, ctypes.c_double, ctypes.c_uint64, ctypes.POINTER(ctypes.c_long), D]
        L.orc_mc_run.restype = ctypes.c_double
        L.orc_mc_run_npt.argtypes = [T, D, ctypes.c_int, D, D, D, ctypes.c_void_p, I, ctypes.c_long, ctypes.c_double,
                                     ctypes.c_double, ctypes.c_double, ctypes.c_uint64, ctypes.POINTER(ctypes.c_long), D,
                                     ctypes.c_int, ctypes.c_double, ctypes.c_double]
        L.orc_mc_run_npt.restype = ctypes.c_double
        L.orc_mc_run_chains.argtypes = [T, D, ctypes.c_int, D, D, D, ctypes.c_void_p, I, ctypes.c_long,
                                        ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                        ctypes.c_double, ctypes.c_double, ctypes.c_uint64,
                                        ctypes.POINTER(ctypes.c_long), D, ctypes.c_int, ctypes.c_double,
                                        ctypes.c_double]
        L.orc_mc_run_chains.restype = ctypes.c_double
        L.orc_soa_mol2others.argtypes = [T, D, ctypes.c_int, D, D, D, ctypes.c_void_p, I, ctypes.c_int, I, I, D]
        L.orc_nematic_s2.argtypes = [ctypes.c_int, D]
        L.orc_nematic_s2.restype = ctypes.c_double
        L.orc_set_sampler_fault.argtypes = [ctypes.c_int]
        L.orc_set_sampler_fault.restype = None
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def _up(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))


def _opt(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def parse_types(top_text):
    """[Types] lines of a top.init (inicializer.cpp:453-585, :687-691):
    returns [(type_number, geotype_code, [params...])]."""
    out, section = [], None
    for raw in top_text.splitlines():
        line = raw.split("#", 1)[0].strip()
        if not line:
            continue
        if line.startswith("["):
            section = line[1:].split("]", 1)[0].strip().upper()
            continue
        if section == "TYPES":
            f = line.split()
            out.append((int(f[1]), GEOTYPES[f[2]], [float(x) for x in f[3:]]))
    return out


def parse_exclusions(top_text):
    ex = np.zeros((MAXT, MAXT), np.uint8)
    section = None
    for raw in top_text.splitlines():
        line = raw.split("#", 1)[0].strip()
        if not line:
            continue
        if line.startswith("["):
            section = line[1:].split("]", 1)[0].strip().upper()
            continue
        if section == "EXCLUDE":
            v = [int(x) for x in line.split()]
            for a, b in zip(v[0::2], v[1::2]):
                ex[a, b] = ex[b, a] = 1
    return ex


def parse_molecules(top_text):
    """[Molecules] and [System] (Inicializer::fillMol, inicializer.cpp:929-1108;
    particle order of setParticlesParamss, :409-450).  Returns
    ([(name, [types], {param: value})], [(name, count)]); angles in radians,
    absent terms -1 (MoleculeParams defaults, moleculeparams.h:42-43)."""
    mols, system, section, cur = [], [], None, None
    for raw in top_text.splitlines():
        line = raw.split("#", 1)[0].strip()
        if not line:
            continue
        if line.startswith("["):
            section = line[1:].split("]", 1)[0].strip().upper()
            continue
        if section == "MOLECULES":
            rest = line
            if cur is None:
                name, _, rest = rest.partition(":")
                cur = (name.strip(), [], {k: -1.0 for k in MOL_KEYS})
                mols.append(cur)
                rest = rest.partition("{")[2]
            closes = "}" in rest
            rest = rest.split("}", 1)[0].strip()
            if rest:
                cmd, _, val = rest.partition(":")
                cmd, v = cmd.strip().upper(), val.split()
                if cmd == "PARTICLES":
                    cur[1].append(int(v[0]))
                elif cmd in ("BOND1", "BOND2", "BONDD", "BONDH", "ANGLE1", "ANGLE2"):
                    k = cmd.lower()
                    c, eq = float(v[0]), float(v[1])
                    if k.startswith("angle"):
                        eq *= DEGTORAD
                    cur[2][k + "c"], cur[2][k + "eq"] = c, eq
            if closes:
                cur = None
        elif section == "SYSTEM":
            f = line.split()
            system.append((f[0], int(f[1])))
    return mols, system


class Topo:
    """Oracle topology built from top.init text, with the chain layout of its
    [System] section (molecule type and position of every particle)."""

    def __init__(self, top_text):
        self.t = _Topo()
        L = lib()
        L.orc_topo_clear(ctypes.byref(self.t))
        for num, g, params in parse_types(top_text):
            arr = np.ascontiguousarray(params, np.float64)
            if not L.orc_topo_add_type(ctypes.byref(self.t), num, g, _dp(arr), len(params)):
                raise ValueError("TOPOLOGY ERROR: wrong number of parameters for type %d" % num)
        ex = parse_exclusions(top_text)
        L.orc_topo_finalize(ctypes.byref(self.t), ex.ctypes.data_as(ctypes.c_void_p))
        mols, system = parse_molecules(top_text)
        names = [m[0] for m in mols]
        for k, (name, types, prm) in enumerate(mols):
            arr = np.array([prm[key] for key in MOL_KEYS], np.float64)
            L.orc_topo_set_mol(ctypes.byref(self.t), k, max(1, len(types)), _dp(arr))
        chain, ptypes = [], []
        for name, count in system:
            k = names.index(name)
            for _ in range(count):
                for pos, ty in enumerate(mols[k][1]):
                    chain.append((k, pos))
                    ptypes.append(ty)
        self.types = np.array(ptypes, np.int32)
        self.chain = np.ascontiguousarray(np.array(chain, np.int32).reshape(-1, 2))
        self.has_chains = any(len(m[1]) > 1 for m in mols)
        if self.has_chains:
            L.orc_topo_set_chain(ctypes.byref(self.t), len(self.chain), self.chain.ctypes.data_as(ctypes.c_void_p))

    @property
    def ref(self):
        return ctypes.byref(self.t)

    def ia(self, i, j):
        return self.t.ia[i][j]

    @property
    def maxcut(self):
        return self.t.maxcut

    @property
    def sqmaxcut(self):
        return self.t.sqmaxcut


class State:
    """SoA FP64 particle state in box fractions (as Particle::pos)."""

    def __init__(self, box, pos, dir, pd0, types, pd1=None):
        self.box = np.ascontiguousarray(box, np.float64)
        self.pos = np.ascontiguousarray(pos, np.float64).reshape(-1, 3)
        self.dir = np.ascontiguousarray(dir, np.float64).reshape(-1, 3)
        self.pd0 = np.ascontiguousarray(pd0, np.float64).reshape(-1, 3)
        self.pd1 = None if pd1 is None else np.ascontiguousarray(pd1, np.float64).reshape(-1, 3)
        self.types = np.ascontiguousarray(types, np.int32)

    @property
    def n(self):
        return len(self.types)

    @classmethod
    def from_real(cls, box, pos_real, dir, pd0, types):
        """Apply Inicializer::initConfig's transformations to real-unit input."""
        box = np.ascontiguousarray(box, np.float64)
        pos = np.array(pos_real, np.float64).reshape(-1, 3).copy()
        d = np.array(dir, np.float64).reshape(-1, 3).copy()
        q = np.array(pd0, np.float64).reshape(-1, 3).copy()
        lib().orc_soa_initconfig(len(pos), _dp(box), _dp(pos), _dp(d), _dp(q))
        return cls(box, pos, d, q, types)


def pair_energies(topo, st, pi, pj):
    pi = np.ascontiguousarray(pi, np.int32)
    pj = np.ascontiguousarray(pj, np.int32)
    out = np.zeros(len(pi), np.float64)
    lib().orc_soa_pair_energies(topo.ref, _dp(st.box), st.n, _dp(st.pos), _dp(st.dir), _dp(st.pd0), _opt(st.pd1),
                                _ip(st.types), len(pi), _ip(pi), _ip(pj), _dp(out))
    return out


def bond_energies(topo, st, pi, pj):
    """bond + angle part of the pair energies (i as part1)"""
    pi = np.ascontiguousarray(pi, np.int32)
    pj = np.ascontiguousarray(pj, np.int32)
    out = np.zeros(len(pi), np.float64)
    lib().orc_soa_bond_energies(topo.ref, _dp(st.box), st.n, _dp(st.pos), _dp(st.dir), _dp(st.pd0), _opt(st.pd1),
                                _ip(st.types), len(pi), _ip(pi), _ip(pj), _dp(out))
    return out


def energy_all(topo, st, use_cells=None):
    if use_cells is None:
        use_cells = st.n > 3000
    return lib().orc_soa_energy_all(topo.ref, _dp(st.box), st.n, _dp(st.pos), _dp(st.dir), _dp(st.pd0),
                                    _opt(st.pd1), _ip(st.types), int(bool(use_cells)))


def energy_one_each(topo, st, targets):
    t = np.ascontiguousarray(targets, np.int32)
    out = np.zeros(len(t), np.float64)
    lib().orc_soa_energy_one_each(topo.ref, _dp(st.box), st.n, _dp(st.pos), _dp(st.dir), _dp(st.pd0),
                                  _opt(st.pd1), _ip(st.types), len(t), _ip(t), _dp(out))
    return out


def derived(topo, st):
    out = np.zeros((st.n, 30), np.float64)
    lib().orc_soa_derived(topo.ref, st.n, _dp(st.dir), _dp(st.pd0), _opt(st.pd1), _ip(st.types), _dp(out))
    return out


def min_dist_segments(a, b, h1, h2, r):
    res = lib().orc_min_dist_segments(_Vec(*map(float, a)), _Vec(*map(float, b)), float(h1), float(h2),
                                      _Vec(*map(float, r)))
    return np.array([res.x, res.y, res.z])


def cell_list(u, nc, shift=(0, 0, 0)):
    u = np.ascontiguousarray(u, np.uint32).reshape(-1, 3)
    n = len(u)
    ux, uy, uz = [np.ascontiguousarray(u[:, k]) for k in range(3)]
    ncell = int(nc[0] * nc[1] * nc[2])
    cell_of = np.zeros(n, np.int32)
    start = np.zeros(ncell + 1, np.int32)
    srt = np.zeros(n, np.int32)
    i32 = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
    lib().orc_cell_list(n, _up(ux), _up(uy), _up(uz), int(nc[0]), int(nc[1]), int(nc[2]),
                        int(shift[0]) & 0xFFFFFFFF, int(shift[1]) & 0xFFFFFFFF, int(shift[2]) & 0xFFFFFFFF,
                        i32(cell_of), i32(start), i32(srt))
    return cell_of, start, srt


def neighbor_pairs(u, box, rc):
    u = np.ascontiguousarray(u, np.uint32).reshape(-1, 3)
    ux, uy, uz = [np.ascontiguousarray(u[:, k]) for k in range(3)]
    box = np.ascontiguousarray(box, np.float64)
    L = lib()
    cnt = L.orc_neighbor_pairs(len(u), _up(ux), _up(uy), _up(uz), _dp(box), rc, None, 0)
    out = np.zeros((max(cnt, 1), 2), np.int32)
    L.orc_neighbor_pairs(len(u), _up(ux), _up(uy), _up(uz), _dp(box), rc, out.ctypes.data_as(ctypes.c_void_p), cnt)
    return out[:cnt]


def mc_run(topo, st, nsweeps, temper, transmx, rot_half, seed, energies=True):
    """Sequential Metropolis on a copy of st; returns (new_state, drift, stats, e_final).
    energies=False skips the whole-system energies used for the drift check."""
    s = State(st.box, st.pos.copy(), st.dir.copy(), st.pd0.copy(), st.types.copy(),
              None if st.pd1 is None else st.pd1.copy())
    stats = (ctypes.c_long * 4)()
    ef = ctypes.c_double(0.0)
    drift = lib().orc_mc_run(topo.ref, _dp(s.box), s.n, _dp(s.pos), _dp(s.dir), _dp(s.pd0), _opt(s.pd1),
                             _ip(s.types), int(nsweeps), float(temper), float(transmx), float(rot_half),
                             int(seed) & 0xFFFFFFFFFFFFFFFF, stats, ctypes.byref(ef) if energies else None)
    return s, drift, list(stats), ef.value


def set_sampler_fault(kind):
    """Plant a mistake in the sequential sampler (tests of the sampler's own
    checks only): 0 none, 1 Metropolis test at 2T, 2 rotation leaving the
    patch direction unrotated, 3 translations along +dir only."""
    lib().orc_set_sampler_fault(int(kind))


def mc_run_npt(topo, st, nsweeps, temper, transmx, rot_half, seed, ptype, press, edge_mx):
    """mc_run plus one volume move per sweep; returns (new_state, drift,
    stats [trans acc, try, rot acc, try, volume acc, try], e_final)."""
    s = State(st.box.copy(), st.pos.copy(), st.dir.copy(), st.pd0.copy(), st.types.copy(),
              None if st.pd1 is None else st.pd1.copy())
    stats = (ctypes.c_long * 6)()
    ef = ctypes.c_double(0.0)
    drift = lib().orc_mc_run_npt(topo.ref, _dp(s.box), s.n, _dp(s.pos), _dp(s.dir), _dp(s.pd0), _opt(s.pd1),
                                 _ip(s.types), int(nsweeps), float(temper), float(transmx), float(rot_half),
                                 int(seed) & 0xFFFFFFFFFFFFFFFF, stats, ctypes.byref(ef), int(ptype), float(press),
                                 float(edge_mx))
    return s, drift, list(stats), ef.value


def mc_run_chains(topo, st, nsweeps, temper, transmx, rot_half, chainprob, chain_mx, chain_angle, seed,
                  ptype=-1, press=0.0, edge_mx=0.0):
    """mc_run with whole-molecule moves (each step a chain move with
    probability chainprob) and, with ptype >= 0, one volume move per sweep;
    returns (new_state, drift, stats [trans acc, try, rot acc, try, volume
    acc, try, chain displace acc, try, chain rotate acc, try], e_final)."""
    s = State(st.box.copy(), st.pos.copy(), st.dir.copy(), st.pd0.copy(), st.types.copy(),
              None if st.pd1 is None else st.pd1.copy())
    stats = (ctypes.c_long * 10)()
    ef = ctypes.c_double(0.0)
    drift = lib().orc_mc_run_chains(topo.ref, _dp(s.box), s.n, _dp(s.pos), _dp(s.dir), _dp(s.pd0), _opt(s.pd1),
                                    _ip(s.types), int(nsweeps), float(temper), float(transmx), float(rot_half),
                                    float(chainprob), float(chain_mx), float(chain_angle),
                                    int(seed) & 0xFFFFFFFFFFFFFFFF, stats, ctypes.byref(ef), int(ptype),
                                    float(press), float(edge_mx))
    return s, drift, list(stats), ef.value


def mol2others(topo, st, first, size):
    """Non-bonded energy of each molecule [first, first + size) with every
    particle outside it (TotalE mol2others / mol2othersTrial)."""
    first = np.ascontiguousarray(first, np.int32)
    size = np.ascontiguousarray(size, np.int32)
    out = np.zeros(len(first))
    lib().orc_soa_mol2others(topo.ref, _dp(st.box), st.n, _dp(st.pos), _dp(st.dir), _dp(st.pd0), _opt(st.pd1),
                             _ip(st.types), len(first), _ip(first), _ip(size), _dp(out))
    return out


def nematic_s2(dirs):
    d = np.ascontiguousarray(dirs, np.float64).reshape(-1, 3)
    return lib().orc_nematic_s2(len(d), _dp(d))


def cluster_labels(topo, st, threshold=-0.10):
    """Cluster labels (ClusterSampler::sameCluster, clust.cpp:51-99: the same
    molecule, or pair energy <= threshold; genClusterList, clust.cpp:105-163:
    each particle ends labelled with the smallest index of its component).
    Pairs within maxcut (beyond it the non-bonded energy is 0), FP64 oracle
    energies, union by smallest root."""
    u = np.round(np.mod(st.pos, 1.0) * 4294967296.0).astype(np.uint64) % 4294967296
    pr = neighbor_pairs(u.astype(np.uint32), st.box, topo.maxcut)
    e = pair_energies(topo, st, pr[:, 0], pr[:, 1]) if len(pr) else np.zeros(0)
    parent = np.arange(st.n)

    def root(a):
        while parent[a] != a:
            parent[a] = parent[parent[a]]
            a = parent[a]
        return a

    links = [(a, b) for (a, b), ee in zip(pr, e) if ee <= threshold]
    if topo.has_chains:
        sizes = np.array([topo.t.mol[k].size for k in range(max(1, topo.t.nmol))])
        c = topo.chain
        links += [(i, i + 1) for i in range(st.n - 1) if c[i, 1] + 1 < sizes[c[i, 0]]]
    for a, b in links:
        ra, rb = root(a), root(b)
        if ra != rb:
            parent[max(ra, rb)] = min(ra, rb)
    return np.array([root(a) for a in range(st.n)], np.int32)


def cluster_labels_reference_loop(energy_of_pair, n, threshold=-0.10):
    """Literal restatement of genClusterList's relabelling (clust.cpp:105-163)
    over all pairs i < j, for small n: restart the sweep after every change."""
    lab = list(range(n))
    change = True
    while change:
        change = False
        for i in range(n):
            for j in range(i, n):
                if i != j and energy_of_pair(i, j) <= threshold:
                    if lab[i] < lab[j]:
                        lab[j] = lab[i]
                        change = True
                        break
                    if lab[j] < lab[i]:
                        lab[i] = lab[j]
                        change = True
                        break
            if change:
                break
    return np.array(lab, np.int32)


# ------------------------------------------------------------------ reference harness

OPTIONS_TEMPLATE = """ptype = 1
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
pairlist_update = 0
seed = {seed}
write_cluster = 0
"""


def config_text(box, pos_real, dir, pd0):
    lines = ["%.17g %.17g %.17g" % tuple(box)]
    for p, d, q in zip(np.asarray(pos_real).reshape(-1, 3), np.asarray(dir).reshape(-1, 3),
                       np.asarray(pd0).reshape(-1, 3)):
        lines.append("%.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g 0" % (tuple(p) + tuple(d) + tuple(q)))
    return "\n".join(lines) + "\n"


def ref_available():
    return os.path.exists(REF_BIN)


def ref_query(mode, top_text, box, pos_real, dir, pd0, extra=None, options=None):
    """Run the compiled reference harness on one configuration; returns the
    output file text.  `extra` is the text of the mode's input file."""
    with tempfile.TemporaryDirectory() as d:
        with open(os.path.join(d, "options"), "w") as f:
            f.write(options or OPTIONS_TEMPLATE.format(nsweeps=0, transmx=0.0, rotmx=0.0, temper=1.0, seed=1))
        with open(os.path.join(d, "top.init"), "w") as f:
            f.write(top_text)
        with open(os.path.join(d, "config.init"), "w") as f:
            f.write(config_text(box, pos_real, dir, pd0))
        args = [REF_BIN, mode, os.path.join(d, "out.txt")]
        if extra is not None:
            with open(os.path.join(d, "in.txt"), "w") as f:
                f.write(extra)
            args.append(os.path.join(d, "in.txt"))
        r = subprocess.run(args, cwd=d, stdout=subprocess.DEVNULL, stderr=subprocess.PIPE, timeout=600)
        if r.returncode != 0:
            raise RuntimeError("reference harness failed: %s" % r.stderr.decode()[-2000:])
        with open(os.path.join(d, "out.txt")) as f:
            return f.read()
