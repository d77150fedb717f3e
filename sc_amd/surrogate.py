"""Training of the pair-attraction surrogate on the GPU and the switch to the
delayed-acceptance sweep.

The network (12 body-frame features -> 3 x 64 SiLU -> 1) is trained by the
library's own FP32 forward/backward kernels and Adam (sc_mlp_train) on pair
energies that the exact FP32 potential generates on the device
(sc_mlp_generate: uniform orientations, closest segment distances 0 - dmax).
The sweep runs it in BF16 on MFMA as the first stage of two-stage delayed
acceptance, so the exact energy still decides every move and the sampled
distribution is unchanged whatever the network's error.
"""
import time

import numpy as np


def train(g, ptype, nsamples=10_000_000, epochs=10, batch=4096, lr=2e-3, lr_final=5e-5, dmax=2.5, near_frac=0.7,
          seed=1, val_frac=0.02, log=None):
    """Initialise the surrogate of type `ptype` (reference type number) in
    GpuSystem `g`, generate `nsamples` exact pairs and train on them, all on
    the device.  Returns what was done and the held-out error (BF16 MFMA
    inference, the arithmetic the sweep uses)."""
    t0 = time.perf_counter()
    g.mlp_init(ptype, seed=seed)
    g.mlp_generate(nsamples, seed=seed + 1, dmax=dmax, near_frac=near_frac)
    t1 = time.perf_counter()
    loss, rmse = g.mlp_train(epochs, batch=batch, lr=lr, lr_final=lr_final, seed=seed + 2, val_frac=val_frac)
    t2 = time.perf_counter()
    nval = int(nsamples * val_frac)
    _, yv, _ = g.mlp_training_data(nsamples - nval, nval)
    inside = yv != 0.0
    res = {"samples": int(nsamples), "epochs": int(epochs), "batch": int(batch), "dmax": float(dmax),
           "near_frac": float(near_frac), "train_mse": float(loss), "val_rmse": float(rmse),
           "target_std": float(np.std(yv)),
           "target_std_inside_rcut": float(np.std(yv[inside])) if inside.any() else 0.0,
           "generate_s": t1 - t0, "train_s": t2 - t1}
    # arithmetic of the training steps: forward 2 (12*64 + 64*64 + 64*64 + 64) = 18,048 flop per sample,
    # backward twice that (input and weight gradients); FP32 on the vector units (peak 157.3 TFLOP/s)
    steps = ((nsamples - nval) // batch) * epochs
    flop = 3.0 * 18048.0 * batch * steps
    res["train_steps"] = int(steps)
    res["train_tflops"] = flop / max(t2 - t1, 1e-9) / 1e12
    res["train_roofline"] = {"bound": "fp32", "achieved": res["train_tflops"], "peak": 157.3, "unit": "TFLOP/s",
                             "frac": res["train_tflops"] / 157.3}
    if log:
        log("surrogate: %d samples (%.2fs), %d epochs on the GPU (%.2fs): train mse %.3e, held-out rmse %.4f "
            "(target std %.4f)" % (nsamples, t1 - t0, epochs, t2 - t1, loss, rmse, res["target_std"]))
    return res
