"""sc_amd: MI355X (gfx950) Monte Carlo hot path for patchy spherocylinders.

Python front-end over the C ABI of sc_amd/lib/libscamd.so (include/sc_gpu.h).
The energy / move methods mirror robertvacha/SC's TotalEnergyCalculator
(scOOP/mc/totalenergycalculator.h:135-522) so host code reads like the
reference's MoveCreator.  Every compute call runs the HIP kernels; there is no
CPU fallback: without the built library or a GPU the calls raise.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libscamd.so")

SC_ERRORS = {-1: "SC_EINVAL", -2: "SC_EHIP", -3: "SC_ESTATE", -4: "SC_ECAPACITY", -5: "SC_ENODEV"}


class SCError(RuntimeError):
    pass


class SweepStats(ctypes.Structure):
    _fields_ = [("trans_acc", ctypes.c_uint64), ("trans_try", ctypes.c_uint64), ("rot_acc", ctypes.c_uint64),
                ("rot_try", ctypes.c_uint64), ("left_cell", ctypes.c_uint64), ("stage1_acc", ctypes.c_uint64),
                ("stage1_try", ctypes.c_uint64), ("stage2_acc", ctypes.c_uint64), ("stage2_try", ctypes.c_uint64),
                ("energy_change", ctypes.c_double), ("bond_rejected", ctypes.c_uint64),
                ("chain_disp_acc", ctypes.c_uint64), ("chain_disp_try", ctypes.c_uint64),
                ("chain_rot_acc", ctypes.c_uint64), ("chain_rot_try", ctypes.c_uint64),
                ("chain_rejected", ctypes.c_uint64), ("chain_cells_skipped", ctypes.c_uint64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


# exported symbol -> (restype, argtypes)
_P = ctypes.c_void_p
_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int32)
_U = ctypes.POINTER(ctypes.c_uint32)
_U8 = ctypes.POINTER(ctypes.c_uint8)
_F = ctypes.POINTER(ctypes.c_float)
SIGNATURES = {
    "sc_create": (ctypes.c_int, [ctypes.POINTER(_P), ctypes.c_int]),
    "sc_destroy": (None, [_P]),
    "sc_last_error": (ctypes.c_char_p, [_P]),
    "sc_load_topology": (ctypes.c_int, [_P, ctypes.c_char_p]),
    "sc_load_options": (ctypes.c_int, [_P, ctypes.c_char_p]),
    "sc_get_option": (ctypes.c_int, [_P, ctypes.c_char_p, _D]),
    "sc_load_config": (ctypes.c_int, [_P, ctypes.c_char_p]),
    "sc_write_config": (ctypes.c_long, [_P, ctypes.c_char_p, ctypes.c_long]),
    "sc_set_block_waves": (ctypes.c_int, [_P, ctypes.c_int]),
    "sc_set_speculation": (ctypes.c_int, [_P, ctypes.c_int]),
    "sc_set_network_min_entries": (ctypes.c_int, [_P, ctypes.c_int]),
    "sc_set_home_cache": (ctypes.c_int, [_P, ctypes.c_int]),
    "sc_set_visit_factor": (ctypes.c_int, [_P, ctypes.c_int]),
    "sc_set_max_active_cells": (ctypes.c_int, [_P, ctypes.c_int]),
    "sc_write_movie_frame": (ctypes.c_long, [_P, ctypes.c_long, ctypes.c_char_p, ctypes.c_long]),
    "sc_write_stat_line": (ctypes.c_long, [_P, ctypes.c_long, ctypes.c_char_p, ctypes.c_long]),
    "sc_write_energy_line": (ctypes.c_long, [_P, ctypes.c_long, ctypes.c_char_p, ctypes.c_long]),
    "sc_alignment_order": (ctypes.c_int, [_P, _D]),
    "sc_write_clusters": (ctypes.c_int, [_P, ctypes.c_double, ctypes.c_long, ctypes.c_char_p, ctypes.c_long,
                                         ctypes.c_char_p, ctypes.c_long, ctypes.POINTER(ctypes.c_long),
                                         ctypes.POINTER(ctypes.c_long)]),
    "sc_num_particles": (ctypes.c_int, [_P]),
    "sc_particle_types": (ctypes.c_int, [_P, _I, ctypes.c_int]),
    "sc_maxcut": (ctypes.c_int, [_P, _D]),
    "sc_pair_params": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_int, _D]),
    "sc_set_config": (ctypes.c_int, [_P, ctypes.c_int, _D, _D, _D, _D, _I]),
    "sc_get_config": (ctypes.c_int, [_P, _D, _D, _D, _D]),
    "sc_get_fixed_positions": (ctypes.c_int, [_P, _U]),
    "sc_build_cells": (ctypes.c_int, [_P, ctypes.c_int, _U, _I, _I, _I, _I]),
    "sc_neighbor_pairs": (ctypes.c_int, [_P, ctypes.c_double, _I, ctypes.c_long, ctypes.POINTER(ctypes.c_long)]),
    "sc_pair_energies": (ctypes.c_int, [_P, ctypes.c_int, _I, _I, _D, ctypes.c_int]),
    "sc_pair_energies_noalloc": (ctypes.c_int, [_P, ctypes.c_int, _I, _I, _D, ctypes.c_int]),
    "sc_energy_all": (ctypes.c_int, [_P, _D]),
    "sc_energy_one": (ctypes.c_int, [_P, ctypes.c_int, _D]),
    "sc_energy_one_trial": (ctypes.c_int, [_P, ctypes.c_int, _D, _D, _D, _D]),
    "sc_commit_particle": (ctypes.c_int, [_P, ctypes.c_int, _D, _D, _D]),
    "sc_energy_mol": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_int, _D]),
    "sc_energy_mol_trial": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_int, _D, _D, _D, _D]),
    "sc_commit_mol": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_int, _D, _D, _D]),
    "sc_set_chain_moves": (ctypes.c_int, [_P, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double]),
    "sc_set_chain_visit_factor": (ctypes.c_int, [_P, ctypes.c_int]),
    "sc_chain_visit_info": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_int32)]),
    "sc_get_chain_extent": (ctypes.c_int, [_P, _D]),
    "sc_chain_pass_info": (ctypes.c_int, [_P, _I, _I]),
    "sc_energy_per_particle": (ctypes.c_int, [_P, _D]),
    "sc_energy_all_trial": (ctypes.c_int, [_P, _D, _D]),
    "sc_commit_box": (ctypes.c_int, [_P, _D]),
    "sc_overlap_flags": (ctypes.c_int, [_P, ctypes.c_int, _I, _I, _U8, _D]),
    "sc_overlap_count": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_long)]),
    "sc_sweep": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                ctypes.c_uint64, ctypes.c_uint64, ctypes.POINTER(SweepStats)]),
    "sc_carried_energy": (ctypes.c_int, [_P, _D, ctypes.POINTER(ctypes.c_int)]),
    "sc_volume_move": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                      ctypes.c_uint64, ctypes.c_uint64, ctypes.POINTER(ctypes.c_int), _D]),
    "sc_generate_rsa": (ctypes.c_int, [_P, ctypes.c_int, _I, _D, ctypes.c_uint64, ctypes.c_long, _D, _D, _D]),
    "sc_last_timing": (ctypes.c_int, [_P, _D, ctypes.POINTER(ctypes.c_long), _D]),
    "sc_set_profiling": (ctypes.c_int, [_P, ctypes.c_int]),
    "sc_copy_probe": (ctypes.c_int, [_P, ctypes.c_long, ctypes.c_int]),
    "sc_set_trial_mode": (ctypes.c_int, [_P, ctypes.c_int]),
    "sc_mlp_init": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_uint64]),
    "sc_mlp_set_weights": (ctypes.c_int, [_P, _F, ctypes.c_int]),
    "sc_mlp_get_weights": (ctypes.c_int, [_P, _F, ctypes.c_int]),
    "sc_mlp_generate": (ctypes.c_int, [_P, ctypes.c_long, ctypes.c_uint64, ctypes.c_double, ctypes.c_double]),
    "sc_mlp_training_data": (ctypes.c_int, [_P, ctypes.c_long, ctypes.c_long, _F, _F, _F]),
    "sc_mlp_train": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_uint64,
                                    ctypes.c_double, _D, _D]),
    "sc_mlp_gradients": (ctypes.c_int, [_P, ctypes.c_int, _F, _F, _F, ctypes.c_int, _D]),
    "sc_mlp_nparam": (ctypes.c_int, []),
    "sc_mlp_forward": (ctypes.c_int, [_P, ctypes.c_int, _F, _F, ctypes.c_int]),
    "sc_mlp_pair_features": (ctypes.c_int, [_P, ctypes.c_int, _I, _I, _F, _F]),
    "sc_set_surrogate": (ctypes.c_int, [_P, ctypes.c_int]),
    "sc_cluster_labels": (ctypes.c_int, [_P, ctypes.c_double, _I, ctypes.POINTER(ctypes.c_int)]),
    "sc_batch_set": (ctypes.c_int, [_P, ctypes.c_int, _I, _D]),
    "sc_batch_sweep": (ctypes.c_int, [_P, ctypes.c_int, _D, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                      ctypes.c_uint64, ctypes.c_uint64, ctypes.POINTER(SweepStats)]),
    "sc_replica_seed": (ctypes.c_uint64, [ctypes.c_uint64, ctypes.c_int]),
    "sc_batch_energies": (ctypes.c_int, [_P, _D]),
    "sc_batch_exchange": (ctypes.c_int, [_P, _D, _D, ctypes.c_int, ctypes.c_uint64, ctypes.c_uint64, _I]),
    "sc_batch_get_config": (ctypes.c_int, [_P, ctypes.c_int, _D, _D, _D]),
    "sc_autotune": (ctypes.c_int, [_P, ctypes.c_int, _D, ctypes.c_double, ctypes.c_double, ctypes.c_uint64, _D,
                                   ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]),
    "sc_nematic_order": (ctypes.c_int, [_P, _D, _D]),
}

MLP_NPARAM = 9217   # W1[12][64] b1 W2[64][64] b2 W3[64][64] b3 W4[64] b4
NOALLOC_MAX_PAIRS = 65536   # SC_NOALLOC_MAX_PAIRS of include/sc_gpu.h
MLP_NF = 12

_lib = None


def lib():
    """Load the HIP library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise SCError("sc_amd native library missing (%s): run __graft_entry__.build()" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        if L.sc_mlp_nparam() != MLP_NPARAM:
            raise SCError("library built with %d surrogate weights, binding expects %d" % (L.sc_mlp_nparam(), MLP_NPARAM))
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(_D)


def _i(a):
    return a.ctypes.data_as(_I)


def _f(a):
    return a.ctypes.data_as(_F)


def _vec3(x):
    return np.ascontiguousarray(x, np.float64).reshape(3)


class GpuSystem:
    """One simulated system on one GPU: the drop-in counterpart of SC's
    TotalEnergyCalculator plus the particle store it reads (Conf::pvec)."""

    def __init__(self, device=0):
        L = lib()
        h = _P()
        rc = L.sc_create(ctypes.byref(h), int(device))
        if rc != 0:
            raise SCError("sc_create failed (%s): no usable HIP device %d" % (SC_ERRORS.get(rc, rc), device))
        self._h = h
        self._L = L
        self.n = 0

    def close(self):
        if getattr(self, "_h", None):
            self._L.sc_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _ck(self, rc, what):
        if rc != 0:
            msg = self._L.sc_last_error(self._h)
            raise SCError("%s: %s (%s)" % (what, msg.decode() if msg else "", SC_ERRORS.get(rc, rc)))

    # ---- SC text inputs
    def load_topology(self, text):
        self._ck(self._L.sc_load_topology(self._h, text.encode()), "load_topology")
        self.n = self._L.sc_num_particles(self._h)

    def load_options(self, text):
        self._ck(self._L.sc_load_options(self._h, text.encode()), "load_options")

    def option(self, key):
        v = ctypes.c_double()
        self._ck(self._L.sc_get_option(self._h, key.encode(), ctypes.byref(v)), "option " + key)
        return v.value

    def load_config(self, text):
        self._ck(self._L.sc_load_config(self._h, text.encode()), "load_config")

    def write_config(self):
        n = self._L.sc_write_config(self._h, None, 0)
        if n < 0:
            self._ck(int(n), "write_config")
        buf = ctypes.create_string_buffer(n + 1)
        self._L.sc_write_config(self._h, buf, n + 1)
        return buf.value.decode()

    def _text(self, fn, what, *args):
        n = fn(self._h, *args, None, 0)
        if n < 0:
            self._ck(int(n), what)
        buf = ctypes.create_string_buffer(n + 1)
        fn(self._h, *args, buf, n + 1)
        return buf.value.decode()

    def movie_frame(self, sweep):
        """one frame of the movie file (Updater::dumpMovie)"""
        return self._text(self._L.sc_write_movie_frame, "write_movie_frame", int(sweep))

    def stat_line(self, sweep):
        """one line of stat.dat: sweep; volume"""
        return self._text(self._L.sc_write_stat_line, "write_stat_line", int(sweep))

    def energy_line(self, sweep):
        """one line of energy.dat: sweep; allToAll  alignmentOrder"""
        return self._text(self._L.sc_write_energy_line, "write_energy_line", int(sweep))

    def alignment_order(self):
        v = ctypes.c_double()
        self._ck(self._L.sc_alignment_order(self._h, ctypes.byref(v)), "alignment_order")
        return v.value

    def cluster_files(self, sweep, threshold=-0.10):
        """(cluster_stat.dat frame, cluster.dat frame) of ClusterSampler::writeCluster"""
        ls, lc = ctypes.c_long(), ctypes.c_long()
        self._ck(self._L.sc_write_clusters(self._h, float(threshold), int(sweep), None, 0, None, 0, ctypes.byref(ls),
                                           ctypes.byref(lc)), "write_clusters")
        bs, bc = ctypes.create_string_buffer(ls.value + 1), ctypes.create_string_buffer(lc.value + 1)
        self._ck(self._L.sc_write_clusters(self._h, float(threshold), int(sweep), bs, ls.value + 1, bc, lc.value + 1,
                                           ctypes.byref(ls), ctypes.byref(lc)), "write_clusters")
        return bs.value.decode(), bc.value.decode()

    def particle_types(self):
        n = self._L.sc_num_particles(self._h)
        out = np.zeros(max(n, 1), np.int32)
        self._L.sc_particle_types(self._h, _i(out), n)
        return out[:n]

    @property
    def maxcut(self):
        v = ctypes.c_double()
        self._ck(self._L.sc_maxcut(self._h, ctypes.byref(v)), "maxcut")
        return v.value

    def pair_params(self, ti, tj):
        out = np.zeros(44)
        self._ck(self._L.sc_pair_params(self._h, ti, tj, _d(out)), "pair_params")
        return out

    # ---- raw state
    def set_config(self, box, pos, dir, pd0, types=None):
        pos = np.ascontiguousarray(pos, np.float64).reshape(-1, 3)
        n = len(pos)
        box = _vec3(box)
        d = np.ascontiguousarray(dir, np.float64).reshape(-1, 3)
        q = np.ascontiguousarray(pd0, np.float64).reshape(-1, 3)
        t = None if types is None else np.ascontiguousarray(types, np.int32)
        self._ck(self._L.sc_set_config(self._h, n, _d(box), _d(pos), _d(d), _d(q), None if t is None else _i(t)),
                 "set_config")
        self.n = n

    def get_config(self):
        n = self.n
        box = np.zeros(3)
        pos, d, q = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros((n, 3))
        self._ck(self._L.sc_get_config(self._h, _d(box), _d(pos), _d(d), _d(q)), "get_config")
        return box, pos, d, q

    def fixed_positions(self):
        u = np.zeros((self.n, 3), np.uint32)
        self._ck(self._L.sc_get_fixed_positions(self._h, u.ctypes.data_as(_U)), "fixed_positions")
        return u

    def generate_rsa(self, types, box, seed, max_attempts=None):
        t = np.ascontiguousarray(types, np.int32)
        n = len(t)
        box = _vec3(box)
        pos, d, q = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros((n, 3))
        ma = max_attempts if max_attempts is not None else 2000 * max(n, 1)
        self._ck(self._L.sc_generate_rsa(self._h, n, _i(t), _d(box), int(seed), int(ma), _d(pos), _d(d), _d(q)),
                 "generate_rsa")
        return pos, d, q

    # ---- cells
    def build_cells(self, r=2, shift=(0, 0, 0)):
        n = self.n
        sh = np.ascontiguousarray(shift, np.uint32)
        nc = np.zeros(3, np.int32)
        cell_of = np.zeros(max(n, 1), np.int32)
        srt = np.zeros(max(n, 1), np.int32)
        # first call for the grid size
        self._ck(self._L.sc_build_cells(self._h, r, sh.ctypes.data_as(_U), None, None, None, _i(nc)), "build_cells")
        start = np.zeros(int(np.prod(nc)) + 1, np.int32)
        self._ck(self._L.sc_build_cells(self._h, r, sh.ctypes.data_as(_U), _i(cell_of), _i(start), _i(srt), _i(nc)),
                 "build_cells")
        return cell_of[:n], start, srt[:n], tuple(int(x) for x in nc)

    def neighbor_pairs(self, rc):
        cnt = ctypes.c_long()
        self._ck(self._L.sc_neighbor_pairs(self._h, rc, None, 0, ctypes.byref(cnt)), "neighbor_pairs")
        out = np.zeros((max(cnt.value, 1), 2), np.int32)
        self._ck(self._L.sc_neighbor_pairs(self._h, rc, _i(out), cnt.value, ctypes.byref(cnt)), "neighbor_pairs")
        return out[:cnt.value]

    # ---- energies (TotalEnergyCalculator method set)
    def p2p(self, i, j, fp64=False):
        """pair energies with i as part1 (TotalE::p2p)"""
        i = np.ascontiguousarray(np.atleast_1d(i), np.int32)
        j = np.ascontiguousarray(np.atleast_1d(j), np.int32)
        out = np.zeros(len(i))
        self._ck(self._L.sc_pair_energies(self._h, len(i), _i(i), _i(j), _d(out), int(bool(fp64))), "pair_energies")
        return out

    def p2p_noalloc(self, i, j, fp64=False):
        """p2p through the entry point that allocates nothing (at most SC_NOALLOC_MAX_PAIRS pairs)"""
        i = np.ascontiguousarray(np.atleast_1d(i), np.int32)
        j = np.ascontiguousarray(np.atleast_1d(j), np.int32)
        out = np.zeros(len(i))
        self._ck(self._L.sc_pair_energies_noalloc(self._h, len(i), _i(i), _i(j), _d(out), int(bool(fp64))),
                 "pair_energies_noalloc")
        return out

    def allToAll(self):
        e = ctypes.c_double()
        self._ck(self._L.sc_energy_all(self._h, ctypes.byref(e)), "energy_all")
        return e.value

    def allToAllTrial(self, box_new):
        e = ctypes.c_double()
        self._ck(self._L.sc_energy_all_trial(self._h, _d(_vec3(box_new)), ctypes.byref(e)), "energy_all_trial")
        return e.value

    def oneToAll(self, target):
        e = ctypes.c_double()
        self._ck(self._L.sc_energy_one(self._h, int(target), ctypes.byref(e)), "energy_one")
        return e.value

    def oneToAllTrial(self, target, pos, dir, pd0):
        e = ctypes.c_double()
        self._ck(self._L.sc_energy_one_trial(self._h, int(target), _d(_vec3(pos)), _d(_vec3(dir)), _d(_vec3(pd0)),
                                             ctypes.byref(e)), "energy_one_trial")
        return e.value

    def update_particle(self, target, pos, dir, pd0):
        self._ck(self._L.sc_commit_particle(self._h, int(target), _d(_vec3(pos)), _d(_vec3(dir)), _d(_vec3(pd0))),
                 "commit_particle")

    def mol2others(self, first, size):
        """TotalEMatrix::mol2others: molecule [first, first + size) with the rest"""
        e = ctypes.c_double()
        self._ck(self._L.sc_energy_mol(self._h, int(first), int(size), ctypes.byref(e)), "energy_mol")
        return e.value

    def mol2othersTrial(self, first, size, pos, dir, pd0):
        """the same with the molecule's particles at trial states (size x 3 each)"""
        pos, dir, pd0 = [np.ascontiguousarray(a, np.float64).reshape(int(size), 3) for a in (pos, dir, pd0)]
        e = ctypes.c_double()
        self._ck(self._L.sc_energy_mol_trial(self._h, int(first), int(size), _d(pos), _d(dir), _d(pd0),
                                             ctypes.byref(e)), "energy_mol_trial")
        return e.value

    def update_mol(self, first, size, pos, dir, pd0):
        pos, dir, pd0 = [np.ascontiguousarray(a, np.float64).reshape(int(size), 3) for a in (pos, dir, pd0)]
        self._ck(self._L.sc_commit_mol(self._h, int(first), int(size), _d(pos), _d(dir), _d(pd0)), "commit_mol")

    def set_chain_moves(self, chainprob, chain_mx, chain_angle, extent=0.0):
        """whole-molecule moves in sweep(): chainprob per step, displacement
        length chain_mx (doubled chainmmx), maximum quaternion half-angle
        chain_angle (radians); extent 0 = 1.2 x the widest molecule now"""
        self._ck(self._L.sc_set_chain_moves(self._h, float(chainprob), float(chain_mx), float(chain_angle),
                                            float(extent)), "set_chain_moves")

    def set_chain_visit_factor(self, v):
        """the chain-move passes visit 1/v' of the coarse colours with v' times the trials each (v' the largest
        divisor of their colour count not above v; default 1)"""
        self._ck(self._L.sc_set_chain_visit_factor(self._h, int(v)), "set_chain_visit_factor")

    def chain_visit_info(self):
        """the chain visit factor v' of the last sweep's chain passes (0 before any)"""
        v = ctypes.c_int32()
        self._ck(self._L.sc_chain_visit_info(self._h, ctypes.byref(v)), "chain_visit_info")
        return v.value

    def chain_pass_info(self):
        """(colour period, staged) of the last sweep's chain moves"""
        p, st = ctypes.c_int32(), ctypes.c_int32()
        self._ck(self._L.sc_chain_pass_info(self._h, ctypes.byref(p), ctypes.byref(st)), "chain_pass_info")
        return p.value, bool(st.value)

    def chain_extent(self):
        e = ctypes.c_double()
        self._ck(self._L.sc_get_chain_extent(self._h, ctypes.byref(e)), "get_chain_extent")
        return e.value

    def update_box(self, box_new):
        self._ck(self._L.sc_commit_box(self._h, _d(_vec3(box_new))), "commit_box")

    def energy_per_particle(self):
        out = np.zeros(self.n)
        self._ck(self._L.sc_energy_per_particle(self._h, _d(out)), "energy_per_particle")
        return out

    # ---- overlaps
    def overlap_flags(self, i, j):
        i = np.ascontiguousarray(i, np.int32)
        j = np.ascontiguousarray(j, np.int32)
        out = np.zeros(len(i), np.uint8)
        dm = np.zeros(len(i))
        self._ck(self._L.sc_overlap_flags(self._h, len(i), _i(i), _i(j), out.ctypes.data_as(_U8), _d(dm)),
                 "overlap_flags")
        return out, dm

    def overlap_count(self):
        c = ctypes.c_long()
        self._ck(self._L.sc_overlap_count(self._h, ctypes.byref(c)), "overlap_count")
        return c.value

    # ---- Monte Carlo
    def sweep(self, nsweeps, temper, transmx, rot_half, r=2, seed=1, counter=0):
        st = SweepStats()
        self._ck(self._L.sc_sweep(self._h, int(nsweeps), float(temper), float(transmx), float(rot_half), int(r),
                                  int(seed) & 0xFFFFFFFFFFFFFFFF, int(counter) & 0xFFFFFFFFFFFFFFFF,
                                  ctypes.byref(st)), "sweep")
        return st.as_dict()

    def volume_move(self, ptype, press, temper, edge_mx, seed, counter):
        acc = ctypes.c_int()
        box = np.zeros(3)
        self._ck(self._L.sc_volume_move(self._h, int(ptype), float(press), float(temper), float(edge_mx),
                                        int(seed) & 0xFFFFFFFFFFFFFFFF, int(counter) & 0xFFFFFFFFFFFFFFFF,
                                        ctypes.byref(acc), _d(box)), "volume_move")
        return bool(acc.value), box

    def carried_energy(self):
        """(E carried for the volume moves, volume moves since its last full evaluation)"""
        e = ctypes.c_double()
        age = ctypes.c_int()
        self._ck(self._L.sc_carried_energy(self._h, ctypes.byref(e), ctypes.byref(age)), "carried_energy")
        return e.value, age.value

    # ---- learned surrogate + delayed acceptance
    def mlp_init(self, ptype, seed=1):
        self._ck(self._L.sc_mlp_init(self._h, int(ptype), int(seed) & 0xFFFFFFFFFFFFFFFF), "mlp_init")

    def mlp_set_weights(self, w):
        w = np.ascontiguousarray(w, np.float32).reshape(MLP_NPARAM)
        self._ck(self._L.sc_mlp_set_weights(self._h, _f(w), MLP_NPARAM), "mlp_set_weights")

    def mlp_get_weights(self):
        w = np.zeros(MLP_NPARAM, np.float32)
        self._ck(self._L.sc_mlp_get_weights(self._h, _f(w), MLP_NPARAM), "mlp_get_weights")
        return w

    def mlp_generate(self, nsamples, seed=1, dmax=2.5, near_frac=0.7):
        self._ck(self._L.sc_mlp_generate(self._h, int(nsamples), int(seed) & 0xFFFFFFFFFFFFFFFF, float(dmax),
                                         float(near_frac)), "mlp_generate")

    def mlp_training_data(self, first, count):
        x = np.zeros((count, MLP_NF), np.float32)
        y = np.zeros(count, np.float32)
        e = np.zeros(count, np.float32)
        self._ck(self._L.sc_mlp_training_data(self._h, int(first), int(count), _f(x), _f(y), _f(e)),
                 "mlp_training_data")
        return x, y, e

    def mlp_train(self, epochs, batch=4096, lr=2e-3, lr_final=5e-5, seed=1, val_frac=0.05):
        tl, vr = ctypes.c_double(), ctypes.c_double()
        self._ck(self._L.sc_mlp_train(self._h, int(epochs), int(batch), float(lr), float(lr_final),
                                      int(seed) & 0xFFFFFFFFFFFFFFFF, float(val_frac), ctypes.byref(tl),
                                      ctypes.byref(vr)), "mlp_train")
        return tl.value, vr.value

    def mlp_gradients(self, x, y):
        x = np.ascontiguousarray(x, np.float32).reshape(-1, MLP_NF)
        y = np.ascontiguousarray(y, np.float32).reshape(-1)
        g = np.zeros(MLP_NPARAM, np.float32)
        loss = ctypes.c_double()
        if len(x) != len(y):
            raise ValueError("mlp_gradients: %d feature rows, %d targets" % (len(x), len(y)))
        self._ck(self._L.sc_mlp_gradients(self._h, len(y), _f(x), _f(y), _f(g), MLP_NPARAM, ctypes.byref(loss)),
                 "mlp_gradients")
        return g, loss.value

    def mlp_forward(self, x, bf16=True):
        x = np.ascontiguousarray(x, np.float32).reshape(-1, MLP_NF)
        out = np.zeros(len(x), np.float32)
        self._ck(self._L.sc_mlp_forward(self._h, len(x), _f(x), _f(out), int(bool(bf16))), "mlp_forward")
        return out

    def mlp_pair_features(self, i, j):
        i = np.ascontiguousarray(i, np.int32)
        j = np.ascontiguousarray(j, np.int32)
        f = np.zeros((len(i), MLP_NF), np.float32)
        e = np.zeros(len(i), np.float32)
        self._ck(self._L.sc_mlp_pair_features(self._h, len(i), _i(i), _i(j), _f(f), _f(e)), "mlp_pair_features")
        return f, e

    DA_OFF, DA_NET, DA_REP = 0, 1, 2

    def set_surrogate(self, mode=1):
        """0 (or False): exact sweep; 1 (or True): delayed acceptance with the
        network screen; 2: delayed acceptance screening only the exact cheap
        terms."""
        self._ck(self._L.sc_set_surrogate(self._h, int(mode)), "set_surrogate")

    # ---- replica batch (parallel tempering)
    def batch_set(self, n_per_rep, boxes):
        ns = np.ascontiguousarray(n_per_rep, np.int32)
        b = np.ascontiguousarray(boxes, np.float64).reshape(-1, 3)
        if len(b) != len(ns):
            raise ValueError("one box per replica")
        self._ck(self._L.sc_batch_set(self._h, len(ns), _i(ns), _d(b)), "batch_set")
        self.nrep = len(ns)
        self.rep_n = ns.copy()

    def batch_sweep(self, nsweeps, temper, transmx, rot_half, r=0, seed=1, counter=0):
        t = np.ascontiguousarray(temper, np.float64)
        st = (SweepStats * self.nrep)()
        self._ck(self._L.sc_batch_sweep(self._h, int(nsweeps), _d(t), float(transmx), float(rot_half), int(r),
                                        int(seed) & 0xFFFFFFFFFFFFFFFF, int(counter) & 0xFFFFFFFFFFFFFFFF, st),
                 "batch_sweep")
        return [x.as_dict() for x in st]

    def batch_energies(self):
        e = np.zeros(self.nrep)
        self._ck(self._L.sc_batch_energies(self._h, _d(e)), "batch_energies")
        return e

    def batch_exchange(self, temper, energy, parity, seed=1, counter=0):
        """returns (new temperatures, accepted flags of the nrep-1 ladder pairs)"""
        t = np.array(temper, np.float64)
        e = np.ascontiguousarray(energy, np.float64)
        acc = np.zeros(max(self.nrep - 1, 1), np.int32)
        self._ck(self._L.sc_batch_exchange(self._h, _d(t), _d(e), int(parity), int(seed) & 0xFFFFFFFFFFFFFFFF,
                                           int(counter) & 0xFFFFFFFFFFFFFFFF, _i(acc)), "batch_exchange")
        return t, acc[: self.nrep - 1]

    def batch_get_config(self, r):
        m = int(self.rep_n[r])
        pos, d, q = np.zeros((m, 3)), np.zeros((m, 3)), np.zeros((m, 3))
        self._ck(self._L.sc_batch_get_config(self._h, int(r), _d(pos), _d(d), _d(q)), "batch_get_config")
        return pos, d, q

    @staticmethod
    def replica_seed(seed, r):
        return lib().sc_replica_seed(int(seed) & 0xFFFFFFFFFFFFFFFF, int(r))

    def autotune(self, sweeps, temper, transmx, rot_half, seed=1):
        """times the four launch shapes; returns (r, trial mode, rates[4])"""
        t = np.ascontiguousarray(np.atleast_1d(temper), np.float64)
        rates = np.zeros(4)
        rb, mb = ctypes.c_int(), ctypes.c_int()
        self._ck(self._L.sc_autotune(self._h, int(sweeps), _d(t), float(transmx), float(rot_half),
                                     int(seed) & 0xFFFFFFFFFFFFFFFF, _d(rates), ctypes.byref(rb), ctypes.byref(mb)),
                 "autotune")
        return rb.value, mb.value, rates

    # ---- observables
    def cluster_labels(self, threshold=-0.10):
        """(labels, number of clusters); labels[i] = smallest index of i's cluster"""
        labels = np.zeros(self.n, np.int32)
        nc = ctypes.c_int()
        self._ck(self._L.sc_cluster_labels(self._h, float(threshold), _i(labels), ctypes.byref(nc)), "cluster_labels")
        return labels, nc.value

    def nematic_order(self):
        """(S2, director)"""
        s2 = ctypes.c_double()
        d = np.zeros(3)
        self._ck(self._L.sc_nematic_order(self._h, ctypes.byref(s2), _d(d)), "nematic_order")
        return s2.value, d

    def set_block_waves(self, nw):
        """waves per trial of an active cell in the sweep (1, 2, 4; 0 = four waves with a trial each)"""
        self._ck(self._L.sc_set_block_waves(self._h, int(nw)), "set_block_waves")

    def set_network_min_entries(self, m=-1):
        """the network screen runs for trials with >= m (pair, state) entries in range (< 0: automatic; 0: all)"""
        self._ck(self._L.sc_set_network_min_entries(self._h, int(m)), "set_network_min_entries")

    def set_speculation(self, on=True):
        """speculative trials, one per wave, in sweep() (default on; block waves 0)"""
        self._ck(self._L.sc_set_speculation(self._h, int(bool(on))), "set_speculation")

    def set_home_cache(self, cap):
        """cells of up to `cap` particles (<= 64) cache their derived vectors in the sweep; 0 disables it"""
        self._ck(self._L.sc_set_home_cache(self._h, int(cap)), "set_home_cache")

    def set_visit_factor(self, v):
        """a sweep visits 1/v of the colours, each with v times the trials (default 1)"""
        self._ck(self._L.sc_set_visit_factor(self._h, int(v)), "set_visit_factor")

    def set_max_active_cells(self, n):
        """at most n active cells per colour in sweep grids (0: no limit, < 0: two per CU)"""
        self._ck(self._L.sc_set_max_active_cells(self._h, int(n)), "set_max_active_cells")

    def set_trial_mode(self, mode):
        self._ck(self._L.sc_set_trial_mode(self._h, int(mode)), "set_trial_mode")

    def copy_probe(self, nbytes, reps=3):
        self._ck(self._L.sc_copy_probe(self._h, int(nbytes), int(reps)), "copy_probe")

    def set_profiling(self, on=True):
        self._ck(self._L.sc_set_profiling(self._h, int(bool(on))), "set_profiling")

    def last_timing(self):
        ms, kms = ctypes.c_double(), ctypes.c_double()
        nl = ctypes.c_long()
        self._ck(self._L.sc_last_timing(self._h, ctypes.byref(ms), ctypes.byref(nl), ctypes.byref(kms)),
                 "last_timing")
        return ms.value, nl.value, kms.value
