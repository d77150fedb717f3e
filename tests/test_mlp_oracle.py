"""Pins of the FP64 surrogate-network oracle (oracle/mlp_oracle.py): an
independent implementation (torch float64 autograd), finite differences,
geometric invariances and closed forms of the features, and the reference's
attraction profile."""
import numpy as np
import pytest
import torch

import mlp_oracle as mo
import pyoracle as po
from sc_amd import workloads as W


def _weights(seed):
    rng = np.random.default_rng(seed)
    w = np.zeros(mo.NPARAM)
    w[mo.OFF_W1:mo.OFF_B1] = rng.normal(0, np.sqrt(2.0 / mo.NF), mo.NF * mo.NH)
    w[mo.OFF_W2:mo.OFF_B2] = rng.normal(0, np.sqrt(2.0 / mo.NH), mo.NH * mo.NH)
    w[mo.OFF_W3:mo.OFF_B3] = rng.normal(0, np.sqrt(2.0 / mo.NH), mo.NH * mo.NH)
    w[mo.OFF_W4:mo.OFF_B4] = rng.normal(0, 0.2, mo.NH)
    for a, b in ((mo.OFF_B1, mo.OFF_W2), (mo.OFF_B2, mo.OFF_W3), (mo.OFF_B3, mo.OFF_W4)):
        w[a:b] = rng.normal(0, 0.1, b - a)
    w[mo.OFF_B4] = 0.3
    return w


def _torch_loss(w, x, y):
    """The same network as torch.nn layers (float64), built from the flat vector."""
    wt = torch.tensor(w, dtype=torch.float64, requires_grad=True)
    W1, b1, W2, b2, W3, b3, w4, b4 = (wt[mo.OFF_W1:mo.OFF_B1].view(mo.NF, mo.NH), wt[mo.OFF_B1:mo.OFF_W2],
                                       wt[mo.OFF_W2:mo.OFF_B2].view(mo.NH, mo.NH), wt[mo.OFF_B2:mo.OFF_W3],
                                       wt[mo.OFF_W3:mo.OFF_B3].view(mo.NH, mo.NH), wt[mo.OFF_B3:mo.OFF_W4],
                                       wt[mo.OFF_W4:mo.OFF_B4], wt[mo.OFF_B4])
    xt = torch.tensor(x, dtype=torch.float64)
    h = torch.nn.functional.silu(torch.nn.functional.linear(xt, W1.T, b1))
    h = torch.nn.functional.silu(torch.nn.functional.linear(h, W2.T, b2))
    h = torch.nn.functional.silu(torch.nn.functional.linear(h, W3.T, b3))
    out = h @ w4 + b4
    loss = torch.nn.functional.mse_loss(out, torch.tensor(y, dtype=torch.float64))
    loss.backward()
    return out.detach().numpy(), loss.item(), wt.grad.numpy()


def test_forward_loss_grad_match_torch_autograd():
    rng = np.random.default_rng(3)
    w = _weights(4)
    x = rng.normal(size=(300, mo.NF))
    y = rng.normal(size=300)
    out_t, loss_t, g_t = _torch_loss(w, x, y)
    np.testing.assert_allclose(mo.forward(w, x), out_t, rtol=1e-13, atol=1e-13)
    loss, g = mo.loss_and_grad(w, x, y)
    assert loss == pytest.approx(loss_t, rel=1e-13)
    np.testing.assert_allclose(g, g_t, rtol=1e-10, atol=1e-14)


def test_grad_matches_finite_differences():
    rng = np.random.default_rng(5)
    w = _weights(6)
    x = rng.normal(size=(64, mo.NF))
    y = rng.normal(size=64)
    _, g = mo.loss_and_grad(w, x, y)
    # one parameter of every block, plus random ones
    idx = [mo.OFF_W1 + 7, mo.OFF_B1 + 3, mo.OFF_W2 + 100, mo.OFF_B2 + 60, mo.OFF_W3 + 4000, mo.OFF_B3 + 1,
           mo.OFF_W4 + 33, mo.OFF_B4] + list(rng.integers(0, mo.NPARAM, 24))
    h = 1e-6
    for i in idx:
        wp, wm = w.copy(), w.copy()
        wp[i] += h
        wm[i] -= h
        fd = (mo.loss_and_grad(wp, x, y)[0] - mo.loss_and_grad(wm, x, y)[0]) / (2 * h)
        assert fd == pytest.approx(g[i], rel=1e-6, abs=1e-9), i


def _rot(rng):
    q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    return q * np.sign(np.linalg.det(q))


def _unit(v):
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def test_features_invariant_under_common_rotation():
    rng = np.random.default_rng(7)
    n = 200
    ad = _unit(rng.normal(size=(n, 3)))
    ap = _unit(np.cross(ad, rng.normal(size=(n, 3))))
    bd = _unit(rng.normal(size=(n, 3)))
    bp = _unit(np.cross(bd, rng.normal(size=(n, 3))))
    r = rng.normal(size=(n, 3)) * 2
    d = rng.uniform(0, 1.3, n)
    f = mo.features(ad, ap, bd, bp, r, d, 4.2, 1.2, 1.0, 0.2)
    R = _rot(rng)
    g = mo.features(ad @ R.T, ap @ R.T, bd @ R.T, bp @ R.T, r @ R.T, d, 4.2, 1.2, 1.0, 0.2)
    np.testing.assert_allclose(f, g, atol=1e-12)
    # a frame feature changes under a rotation of b alone
    h = mo.features(ad, ap, bd @ R.T, bp @ R.T, r, d, 4.2, 1.2, 1.0, 0.2)
    assert np.max(np.abs(h[:, 3:9] - f[:, 3:9])) > 0.1


def test_features_closed_forms():
    # b straight ahead along a's axis, parallel, patches aligned
    f = mo.features([0, 0, 1], [1, 0, 0], [0, 0, 1], [1, 0, 0], [0, 0, -5.0], 2.0, 5.0, 1.2, 1.0, 0.2)[0]
    np.testing.assert_allclose(f, [1, 0, 0, 1, 0, 0, 0, 1, 0, 2 / 1.2, 1, 0], atol=1e-15)
    # b beside a in the patch direction, antiparallel, patch facing back
    f = mo.features([0, 0, 1], [1, 0, 0], [0, 0, -1], [-1, 0, 0], [-1.1, 0, 0], 1.1, 5.0, 1.2, 1.0, 0.2)[0]
    w = np.cos(0.5 * np.pi * 0.1 / 0.2) ** 2
    np.testing.assert_allclose(f, [0, 1.1 / 5, 0, -1, 0, 0, 0, -1, 0, 1.1 / 1.2, 1.1 / 5, w], atol=1e-15)
    # e3 = dir x patchdir: b along +y sees f[2] = +|v|/lsc
    f = mo.features([0, 0, 1], [1, 0, 0], [0, 1, 0], [0, 0, 1], [0, -2.0, 0], 1.5, 4.0, 1.2, 1.0, 0.2)[0]
    assert f[2] == pytest.approx(0.5) and f[5] == pytest.approx(1.0) and f[6] == pytest.approx(1.0)


def test_switch_weight_matches_reference_attraction_profile():
    """Two parallel PSCs side by side with patches facing each other: beyond
    the WCA range the pair energy is -eps * f0 * w(d) with f0 fixed, so
    E(d1)/E(d2) = w(d1)/w(d2) (EPatch::atrE, paire.cpp:244-301)."""
    top = W.top_init([W.TYPES["psc"]], [2])
    tp = po.Topo(top)
    box = np.array([30.0, 30.0, 30.0])
    ds = np.array([1.125, 1.14, 1.16, 1.18, 1.195])
    es = []
    for d in ds:
        pos = np.array([[15.0, 15.0, 15.0], [15.0 + d, 15.0, 15.0]])
        dirs = np.array([[0.0, 0.0, 1.0], [0.0, 0.0, 1.0]])
        pd0 = np.array([[1.0, 0.0, 0.0], [-1.0, 0.0, 0.0]])
        st = po.State.from_real(box, pos, dirs, pd0, np.array([1, 1], np.int32))
        es.append(po.pair_energies(tp, st, np.array([0], np.int32), np.array([1], np.int32))[0])
    es = np.array(es)
    assert np.all(es < 0)
    w = mo.switch_weight(ds, 1.0, 0.2)
    np.testing.assert_allclose(es / es[0], w / w[0], rtol=1e-9)

