"""Writes cnn.npz: forward output, MSE loss and parameter gradients of the closure network
(periodic 3x3x3 convolutions 9 -> 32 -> 32 -> 32 -> 6, exact GELU) computed by torch on the CPU in
float64 with autograd, on a small seeded case (batch 2, 6^3 grid).  The fixture pins the numpy
oracle's hand-written forward/backward independently of its own formulas.

    python tests/golden/make_cnn_golden.py
"""
import os

import numpy as np
import torch
import torch.nn.functional as F

HERE = os.path.dirname(os.path.abspath(__file__))
CH = (9, 32, 32, 32, 6)


def main():
    rng = np.random.default_rng(20261202)
    B, N = 2, 6
    x = rng.standard_normal((B, N, N, N, CH[0]))
    t = rng.standard_normal((B, N, N, N, CH[-1])) * 0.1
    params = []
    for ci, co in zip(CH[:-1], CH[1:]):
        bound = 1.0 / np.sqrt(ci * 27)
        params.append((rng.uniform(-bound, bound, (co, ci, 3, 3, 3)), rng.uniform(-bound, bound, co)))
    tp = [(torch.tensor(W, requires_grad=True), torch.tensor(b, requires_grad=True)) for W, b in params]
    h = torch.tensor(x).permute(0, 4, 1, 2, 3)  # [B, C, X, Y, Z]
    for l, (W, b) in enumerate(tp):
        h = F.conv3d(F.pad(h, (1, 1, 1, 1, 1, 1), mode="circular"), W, b)
        if l < len(tp) - 1:
            h = F.gelu(h)
    y = h.permute(0, 2, 3, 4, 1)
    loss = torch.mean((y - torch.tensor(t)) ** 2)
    loss.backward()
    flat = lambda ps: np.concatenate([np.concatenate([W.ravel(), b.ravel()]) for W, b in ps])
    grads = np.concatenate([np.concatenate([W.grad.numpy().ravel(), b.grad.numpy().ravel()]) for W, b in tp])
    np.savez_compressed(os.path.join(HERE, "cnn.npz"), x=x, t=t, params=flat(params), y=y.detach().numpy(),
                        loss=np.array(loss.item()), grads=grads)
    print("wrote cnn.npz", y.shape, float(loss))


if __name__ == "__main__":
    main()
