"""FP64 oracle of the pair-energy surrogate network (test infrastructure).

Only the tests, the smoke check and the CPU leg of the benchmark use this
module, as the checker of the HIP training and inference kernels; the product
path never imports it.

The network has no counterpart in robertvacha/SC: the reference evaluates
every trial move with the exact pair functors (scOOP/mc/paire.h:466-827,
scOOP/mc/paire.cpp:244-301).  The features restate, in FP64, the body-frame
description of a pair that the device computes; the target they stand in for
is the reference's patch attraction (SpheroCylinder::operator(),
paire.h:1141-1170), which the FP64 pair oracle (pyoracle) evaluates.

Weight vector (flat, float): W1[12][64] b1[64] W2[64][64] b2[64] W3[64][64]
b3[64] W4[64] b4; layer l computes h = SiLU(x W + b), the output h3 W4 + b4.
Loss: mean squared error over the batch.

Pins (tests/test_mlp_oracle.py):
  forward, loss and gradient  - torch float64 autograd of the same network
                                (an independent implementation), and central
                                finite differences of the loss;
  features                      - invariance under a common rotation of both
                                particles and the offset, closed forms for
                                collinear and side-by-side pairs, and the cos^2
                                switch against the reference's attraction
                                profile (paire.cpp:283-290).
"""
import numpy as np

NF = 12
NH = 64
OFF_W1 = 0
OFF_B1 = OFF_W1 + NF * NH
OFF_W2 = OFF_B1 + NH
OFF_B2 = OFF_W2 + NH * NH
OFF_W3 = OFF_B2 + NH
OFF_B3 = OFF_W3 + NH * NH
OFF_W4 = OFF_B3 + NH
OFF_B4 = OFF_W4 + NH
NPARAM = OFF_B4 + 1


def unpack(w):
    w = np.asarray(w, np.float64)
    assert w.shape == (NPARAM,)
    return (w[OFF_W1:OFF_B1].reshape(NF, NH), w[OFF_B1:OFF_W2], w[OFF_W2:OFF_B2].reshape(NH, NH), w[OFF_B2:OFF_W3],
            w[OFF_W3:OFF_B3].reshape(NH, NH), w[OFF_B3:OFF_W4], w[OFF_W4:OFF_B4], w[OFF_B4])


def silu(z):
    return z / (1.0 + np.exp(-z))


def silu_grad(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(w, x):
    """Network output for feature rows x (B x 12), FP64."""
    W1, b1, W2, b2, W3, b3, w4, b4 = unpack(w)
    x = np.asarray(x, np.float64)
    h1 = silu(x @ W1 + b1)
    h2 = silu(h1 @ W2 + b2)
    h3 = silu(h2 @ W3 + b3)
    return h3 @ w4 + b4


def loss_and_grad(w, x, y):
    """Mean squared error over the batch and its gradient w.r.t. the flat
    weight vector, by reverse-mode differentiation written out layer by layer."""
    W1, b1, W2, b2, W3, b3, w4, b4 = unpack(w)
    x = np.asarray(x, np.float64)
    y = np.asarray(y, np.float64)
    B = len(y)
    z1 = x @ W1 + b1
    h1 = silu(z1)
    z2 = h1 @ W2 + b2
    h2 = silu(z2)
    z3 = h2 @ W3 + b3
    h3 = silu(z3)
    out = h3 @ w4 + b4
    e = out - y
    loss = np.mean(e * e)
    dout = 2.0 * e / B
    g = np.zeros(NPARAM)
    g[OFF_W4:OFF_B4] = h3.T @ dout
    g[OFF_B4] = dout.sum()
    dz3 = np.outer(dout, w4) * silu_grad(z3)
    g[OFF_W3:OFF_B3] = (h2.T @ dz3).ravel()
    g[OFF_B3:OFF_W4] = dz3.sum(0)
    dz2 = (dz3 @ W3.T) * silu_grad(z2)
    g[OFF_W2:OFF_B2] = (h1.T @ dz2).ravel()
    g[OFF_B2:OFF_W3] = dz2.sum(0)
    dz1 = (dz2 @ W2.T) * silu_grad(z1)
    g[OFF_W1:OFF_B1] = (x.T @ dz1).ravel()
    g[OFF_B1:OFF_W2] = dz1.sum(0)
    return loss, g


def switch_weight(d, pdis, pswitch):
    """The radial factor of the patch attraction as a function of the closest
    segment distance: 1 inside pdis, cos^2(pi/2 (d - pdis)/pswitch) on the
    switch, 0 beyond (EPatch::atrE, paire.cpp:283-290)."""
    d = np.asarray(d, np.float64)
    c = np.cos(0.5 * np.pi * (d - pdis) / pswitch)
    return np.where(d < pdis, 1.0, np.where(d > pdis + pswitch, 0.0, c * c))


def features(adir, apd, bdir, bpd, r, d, lsc, rcut, pdis, pswitch):
    """Features of pairs (rows): particle a (moving, part1) with axis adir and
    patch direction apd, neighbour b with bdir, bpd; r = pos_a - pos_b
    (minimum image, Cuboid::image convention, geometry.h:110-128); d = the
    closest segment distance."""
    adir, apd, bdir, bpd, r = (np.atleast_2d(np.asarray(v, np.float64)) for v in (adir, apd, bdir, bpd, r))
    e3 = np.cross(adir, apd)
    v = -r
    f = np.zeros((len(v), NF))
    f[:, 0] = np.sum(v * adir, 1) / lsc
    f[:, 1] = np.sum(v * apd, 1) / lsc
    f[:, 2] = np.sum(v * e3, 1) / lsc
    for q, u in enumerate((bdir, bpd)):
        f[:, 3 + 3 * q] = np.sum(u * adir, 1)
        f[:, 4 + 3 * q] = np.sum(u * apd, 1)
        f[:, 5 + 3 * q] = np.sum(u * e3, 1)
    f[:, 9] = np.asarray(d, np.float64) / rcut
    f[:, 10] = np.linalg.norm(v, axis=1) / lsc
    f[:, 11] = switch_weight(d, pdis, pswitch)
    return f
