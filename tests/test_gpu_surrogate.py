"""The learned surrogate and the delayed-acceptance sweep on the GPU, through
the C ABI, against the FP64 oracles.

Tolerances:
  * gradients of the FP32 training kernels vs the FP64 oracle: per weight
    block, max |difference| <= 2e-4 of the block's largest gradient (FP32
    sums over the batch, accumulated with atomics);
  * FP32 inference vs the FP64 oracle: 1e-5 relative to the output scale;
  * BF16 MFMA inference (the sweep's path) vs the FP64 oracle: RMS error
    <= 2% of the RMS output, max error <= 10% of the largest output;
  * features of configuration pairs vs the FP64 restatement: 2e-5 absolute;
    their exact attraction vs the FP64 pair oracle: 1e-4 relative / 1e-5
    absolute for >= 99% of pairs;
  * delayed-acceptance ensemble: mean E/N equal to the CPU oracle's within 3
    combined standard errors + 1%.
"""
import time

import numpy as np
import pytest

import sc_amd
from sc_amd import prepare
from sc_amd import surrogate
from sc_amd import workloads as W
import mlp_oracle as mo
import pyoracle as po
from test_gpu_parity import ensemble_check

pytestmark = pytest.mark.gpu
TCH = 3
PSC = 1
ROT_HALF = 20.0 / 180 * np.pi / 4


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def tch():
    w = W.workload("c4_tchpsc_65536", n_override=2000)
    g = sc_amd.GpuSystem(0)
    g.load_topology(w["top"])
    g.mlp_init(TCH, seed=11)
    g.mlp_generate(1 << 16, seed=12, dmax=2.5, near_frac=0.7)
    yield g, w
    g.close()


BLOCKS = [("W1", mo.OFF_W1, mo.OFF_B1), ("b1", mo.OFF_B1, mo.OFF_W2), ("W2", mo.OFF_W2, mo.OFF_B2),
          ("b2", mo.OFF_B2, mo.OFF_W3), ("W3", mo.OFF_W3, mo.OFF_B3), ("b3", mo.OFF_B3, mo.OFF_W4),
          ("W4", mo.OFF_W4, mo.OFF_B4), ("b4", mo.OFF_B4, mo.NPARAM)]


def _check_grads(g, x, y, label):
    w = g.mlp_get_weights().astype(np.float64)
    gg, loss = g.mlp_gradients(x, y)
    lo, gr = mo.loss_and_grad(w, x, y)
    say("%s: loss gpu %.8e oracle %.8e" % (label, loss, lo))
    assert loss == pytest.approx(lo, rel=1e-5)
    for name, a, b in BLOCKS:
        scale = np.max(np.abs(gr[a:b]))
        err = np.max(np.abs(gg[a:b] - gr[a:b]))
        say("  %s: max|g| %.3e max err %.3e" % (name, scale, err))
        assert err <= 2e-4 * scale + 1e-9, name


def test_mlp_gradients_match_oracle(tch):
    say("case: surrogate gradients vs FP64 oracle, TCHPSC training pairs")
    g, _ = tch
    x, y, _ = g.mlp_training_data(0, 2048)
    assert np.all(np.isfinite(x)) and np.all(np.isfinite(y))
    assert np.count_nonzero(y) > 200
    _check_grads(g, x, y, "initial weights")
    # and after some training (weights away from their initial distribution)
    g.mlp_train(1, batch=1024, lr=2e-3, lr_final=1e-3, seed=3, val_frac=0.1)
    x2, y2, _ = g.mlp_training_data(4096, 1500)   # a batch size that is no multiple of 256
    _check_grads(g, x2, y2, "trained weights")


def test_mlp_forward_fp32_and_bf16_mfma_vs_oracle(tch):
    say("case: surrogate inference FP32 and BF16 MFMA vs FP64 oracle")
    g, _ = tch
    rng = np.random.default_rng(9)
    w = rng.normal(0, 1, mo.NPARAM).astype(np.float32)
    w[mo.OFF_W1:mo.OFF_B1] *= np.sqrt(2.0 / mo.NF)
    w[mo.OFF_W2:mo.OFF_B3] *= np.sqrt(2.0 / mo.NH)
    w[mo.OFF_B1:mo.OFF_W2] *= 0.1
    w[mo.OFF_B2:mo.OFF_W3] *= 0.1
    w[mo.OFF_B3:mo.OFF_W4] *= 0.1
    w[mo.OFF_W4:] *= 0.3
    w0 = g.mlp_get_weights()
    try:
        g.mlp_set_weights(w)
        x, _, _ = g.mlp_training_data(0, 1000)   # 1000 rows: the last tile is partial
        ref = mo.forward(w.astype(np.float64), x)
        o32 = g.mlp_forward(x, bf16=False)
        o16 = g.mlp_forward(x, bf16=True)
    finally:
        g.mlp_set_weights(w0)
    scale = np.max(np.abs(ref))
    e32 = np.max(np.abs(o32 - ref))
    rms16 = np.sqrt(np.mean((o16 - ref) ** 2))
    say("fp32 max err %.3e, bf16 rms err %.3e (rms out %.3e), max err %.3e (max out %.3e)" % (
        e32, rms16, np.sqrt(np.mean(ref ** 2)), np.max(np.abs(o16 - ref)), scale))
    assert e32 <= 1e-5 * max(1.0, scale)
    assert rms16 <= 0.02 * np.sqrt(np.mean(ref ** 2))
    assert np.max(np.abs(o16 - ref)) <= 0.1 * scale
    # a row's output does not depend on its tile mates (tiles of 32 rows)
    g.mlp_set_weights(w)
    try:
        a = g.mlp_forward(x[:64], bf16=True)
        b = g.mlp_forward(x[16:80], bf16=True)
    finally:
        g.mlp_set_weights(w0)
    np.testing.assert_array_equal(a[16:64], b[:48])


def _wca(d2, eps, A, B, rcwsq):
    i6 = 1.0 / d2 ** 3
    return np.where(d2 <= rcwsq, A * i6 * i6 - B * i6 + eps, 0.0)


def test_pair_features_and_exact_attraction(tch):
    say("case: surrogate features + exact attraction of configuration pairs vs FP64 oracles")
    g, w = tch
    w2 = W.workload("c4_tchpsc_65536", n_override=2000)
    box = w2["box"] * (0.25 / 0.15) ** (1.0 / 3.0)   # RSA reaches eta 0.15 quickly
    pos, d, q = g.generate_rsa(w2["types"], box, seed=44)
    g.set_config(box, pos, d, q, w2["types"])
    b, p, dd, qq = g.get_config()
    tp, st = po.Topo(w2["top"]), po.State(b, p, dd, qq, w2["types"])
    pr = g.neighbor_pairs(7.0)
    i, j = pr[:, 0].copy(), pr[:, 1].copy()
    f, ea = g.mlp_pair_features(i, j)
    e_full = po.pair_energies(tp, st, i, j)
    u = g.fixed_positions().astype(np.int64)
    du = (u[i] - u[j] + 2 ** 31) % 2 ** 32 - 2 ** 31
    r = du * (b / 2.0 ** 32)
    ia = tp.ia(TCH, TCH)
    h = ia.half_len[0]
    dseg = np.array([np.linalg.norm(po.min_dist_segments(dd[a], dd[c], h, h, r[k])) for k, (a, c) in
                     enumerate(zip(i, j))])
    ref_attr = e_full - _wca(dseg ** 2, ia.epsilon, ia.A, ia.B, ia.rcutwcaSq)
    err = np.abs(ea - ref_attr)
    ok = (err <= 1e-4 * np.abs(ref_attr)) | (err <= 1e-5)
    nz = np.count_nonzero(ref_attr)
    say("%d pairs, %d attracting, %d outside tolerance, max err %.3e" % (len(i), nz, (~ok).sum(), err.max()))
    assert nz > 100
    assert ok.mean() >= 0.99
    lsc = ia.half_len[0] + ia.half_len[1] + ia.rcut
    fr = mo.features(dd[i], qq[i], dd[j], qq[j], r, dseg, lsc, ia.rcut, ia.pdis, ia.pswitch)
    ferr = np.max(np.abs(f - fr), axis=0)
    say("feature max errors: " + " ".join("%.1e" % v for v in ferr))
    assert np.all(ferr <= 2e-5)


def test_training_reduces_error(tch):
    """4M device-generated exact pairs, 10 epochs of the device trainer: the
    held-out error of the BF16 MFMA inference, overall and for the pairs inside
    the attraction range (the only ones the sweep asks the network about)."""
    say("case: surrogate training on the GPU, BF16 MFMA held-out error")
    g, _ = tch
    ns = 4 << 20
    t = surrogate.train(g, TCH, nsamples=ns, epochs=10, seed=21, log=say)
    x, y, _ = g.mlp_training_data(ns - 65536, 65536)
    out = g.mlp_forward(x, bf16=True)
    inside = x[:, 9] <= 1.0
    rmse = np.sqrt(np.mean((out - y) ** 2))
    rmse_in = np.sqrt(np.mean((out[inside] - y[inside]) ** 2))
    g.mlp_init(TCH, seed=22)
    out0 = g.mlp_forward(x, bf16=True)
    rmse0 = np.sqrt(np.mean((out0 - y) ** 2))
    say("held-out rmse: untrained %.4f, trained %.4f (inside rcut %.4f, std there %.4f), target std %.4f" % (
        rmse0, rmse, rmse_in, np.std(y[inside]), np.std(y)))
    assert rmse < 0.35 * np.std(y)
    assert rmse_in < 0.35 * np.std(y[inside])
    assert rmse < 0.35 * rmse0


def _blocked(x, nb=8):
    b = x[: len(x) // nb * nb].reshape(nb, -1).mean(1)
    return b.mean(), b.std(ddof=1) / np.sqrt(nb)


@pytest.mark.parametrize("mode", [1, 2])
def test_da_sweep_ensemble_matches_cpu_oracle(ensemble_system, mode):
    """Delayed acceptance samples the exact ensemble: E/N, S2 and mean cluster
    size of 500 PSC at eta 0.2, T*=1 with the network screen (mode 1) and
    with the cheap-terms screen alone (mode 2) equal the sequential CPU
    oracle's within 3 combined standard errors + 1%."""
    say("case: delayed-acceptance ensemble vs CPU oracle, 500 PSC, mode", mode)
    w, (pos, d, q), ec, rot_half = ensemble_system
    g = sc_amd.GpuSystem(0)
    g.load_topology(w["top"])
    if mode == 1:
        surrogate.train(g, PSC, nsamples=1 << 20, epochs=4, seed=5, log=say)
    g.set_config(w["box"], pos, d, q, w["types"])
    g.set_surrogate(mode)
    g.set_network_min_entries(0)   # the network screens every trial that passes the first screen
    n = len(w["types"])
    tot = ensemble_check(g, n, ec, rot_half, "DA mode %d" % mode)
    g.close()
    acc = tot["trans_acc"] + tot["rot_acc"]
    passed = tot["stage2_acc"] if mode == 1 else tot["stage1_acc"]
    say("screen 1 %.4f, screen 2 %.4f, exact stage %.4f, overall %.4f" % (
        tot["stage1_acc"] / max(1, tot["stage1_try"]), tot["stage2_acc"] / max(1, tot["stage2_try"]),
        acc / max(1, passed), acc / max(1, tot["trans_try"] + tot["rot_try"])))
    assert tot["stage1_try"] > 0 and tot["stage1_acc"] > 0
    assert (tot["stage2_try"] > 0) == (mode == 1)


def test_da_gain_over_exact_sweep():
    """What the screens and the trained network buy, on 16,384 TCHPSC at eta
    0.25 from one state: sweeps/s of the exact sweep, of delayed acceptance
    screening the exact cheap terms only, and with an untrained and a trained
    network screen.  The trained screen passes moves the exact stage then
    accepts more often than the untrained one."""
    say("case: DA gain, 16384 TCHPSC eta 0.25")
    w = W.workload("c4_tchpsc_65536", n_override=16384)
    n = len(w["types"])
    g = sc_amd.GpuSystem(0)
    prepare.initial_state(g, w, seed=7)
    box, pos, d, q = g.get_config()
    g.set_trial_mode(1)
    g.set_network_min_entries(0)   # the network screens every trial that passes the first screen

    def run(label, mode, k=20):
        g.set_config(box, pos, d, q, w["types"])
        g.set_surrogate(mode)
        g.sweep(2, 1.0, 0.3, ROT_HALF, r=2, seed=3, counter=0)
        t0 = time.perf_counter()
        s = g.sweep(k, 1.0, 0.3, ROT_HALF, r=2, seed=3, counter=100)
        dt = time.perf_counter() - t0
        sps = (s["trans_try"] + s["rot_try"]) / float(n) / dt
        acc = s["trans_acc"] + s["rot_acc"]
        passed = s["stage2_acc"] if mode == 1 else (s["stage1_acc"] if mode == 2 else acc)
        fid = acc / max(1, passed)
        say("%s: %.1f sweeps/s, acceptance %.4f, screen 1 %.4f, screen 2 %.4f, exact stage %.4f" % (
            label, sps, acc / max(1, s["trans_try"] + s["rot_try"]), s["stage1_acc"] / max(1, s["stage1_try"]),
            s["stage2_acc"] / max(1, s["stage2_try"]), fid))
        return sps, fid

    exact, _ = run("exact", 0)
    rep, _ = run("cheap-terms screen", 2)
    g.mlp_init(TCH, seed=1)
    raw, fid_raw = run("untrained network screen", 1)
    surrogate.train(g, TCH, nsamples=4_000_000, epochs=10, seed=9, log=say)
    net, fid_net = run("trained network screen", 1)
    g.close()
    say("sweeps/s relative to exact: cheap-terms screen %.3f, untrained %.3f, trained %.3f" % (
        rep / exact, raw / exact, net / exact))
    assert fid_net > fid_raw
