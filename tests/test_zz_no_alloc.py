"""sc_pair_energies_noalloc, the entry point that allocates nothing, against
sc_pair_energies through the C ABI: the same kernel on the same inputs, so
the energies must agree bit for bit; two calls in a row, the second smaller,
in FP32 and FP64 device modes; more pairs than the fixed staging holds is an
error, not a fault."""
import ctypes

import numpy as np
import pytest

import sc_amd
from sc_amd import workloads as W

pytestmark = pytest.mark.gpu


def test_noalloc_pair_energies_equal_the_allocating_entry_point():
    print("case: sc_pair_energies_noalloc vs sc_pair_energies, 4000 PSC", flush=True)
    w = W.workload("c1_psc_1000", n_override=4000)
    g = sc_amd.GpuSystem(0)
    try:
        g.load_topology(w["top"])
        pos, d, q = g.generate_rsa(w["types"], w["box"], seed=41)
        g.set_config(w["box"], pos, d, q, w["types"])
        # pairs within the interaction cutoff (both orders: i is part1)
        pairs = g.neighbor_pairs(g.maxcut)
        pairs = np.concatenate([pairs, pairs[:, ::-1]])
        rng = np.random.default_rng(3)
        for fp64 in (False, True):
            for m in (60000, 1234):   # the second call smaller than the first
                pk = pairs[rng.choice(len(pairs), m, replace=False)]
                i, j = np.ascontiguousarray(pk[:, 0]), np.ascontiguousarray(pk[:, 1])
                a = g.p2p(i, j, fp64=fp64)
                b = g.p2p_noalloc(i, j, fp64=fp64)
                print("fp64=%d m=%d: %d non-zero energies, max |diff| %.3e" % (
                    fp64, m, np.count_nonzero(a), np.max(np.abs(a - b))), flush=True)
                assert np.all(np.isfinite(a)) and np.count_nonzero(a) > 0
                np.testing.assert_array_equal(a, b)
        # capacity: refused with SC_ECAPACITY and a message, nothing launched
        m = sc_amd.NOALLOC_MAX_PAIRS + 1
        i = np.zeros(m, np.int32)
        with pytest.raises(sc_amd.SCError, match="exceed"):
            g.p2p_noalloc(i, i)
        assert len(g.p2p_noalloc(np.zeros(0, np.int32), np.zeros(0, np.int32))) == 0
        # the context still works after the refused call
        np.testing.assert_array_equal(g.p2p_noalloc([0, 1], [2, 3]), g.p2p([0, 1], [2, 3]))
    finally:
        g.close()
