"""CPU oracle of the learned-closure path (TEST INFRASTRUCTURE ONLY: imported by tests/ and by
nothing that ships).  numpy, float64.

The reference solver has no LES or closure; this restates, from their definitions, what the GPU
library computes for the learned subgrid-stress model:

  * the LES filter: Gaussian G(k) = exp(-k^2 Delta^2 / 24), Delta = pi / k_c, followed by the sharp
    cut of the LES grid's 2/3 rule (|k| <= N_les // 3), with FFTW normalisation rescaled by
    (N_les / N)^3 (a band-limited field's FFTW coefficients scale with the point count);
  * the subgrid stress of the filtered equations in rotational form: with u x w = grad(|u|^2/2) -
    div(u u) and the projection P removing gradients, P F(u x w) - P T(ub x wb) = -P div(tau),
    tau_ij = F(u_i u_j) - T(ub_i ub_j), where F is the filter above and T the sharp cut alone; the
    model input is grad(ub) (9 channels, g[.., 3 i + j] = d_j ub_i) and the target the deviatoric
    part of tau in the channel order (xx, yy, zz, xy, xz, yz);
  * the network: four periodic 3x3x3 convolutions 9 -> 32 -> 32 -> 32 -> 6 (cross-correlation,
    as torch.nn.functional.conv3d with circular padding), exact-erf GELU after the first three,
    MSE loss over points and the 6 channels; forward and backward by hand, with optional bfloat16
    rounding at the same points as the GPU kernels (weights, layer inputs, dL/d(pre-activation)).

Arrays are channels-last: a field on an N^3 grid is [..., N, N, N, C] with grid index order
(x, y, z) and z fastest.  Spectral arrays use the reference's AoS layout [N][N][N/2+1][3].
"""
from __future__ import annotations

import math

import numpy as np

CHANNELS = (9, 32, 32, 32, 6)
TAU_ORDER = ((0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2))


# ------------------------------------------------------------------------------------ bfloat16
def bf16(a):
    """round-to-nearest-even to bfloat16, returned as float64"""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    out = r.astype(np.uint32).view(np.float32).astype(np.float64)
    nan = np.isnan(f)
    if nan.any():
        out[nan] = np.nan
    return out


# ------------------------------------------------------------------------------------ spectra
def wavenumbers(N):
    k = np.fft.fftfreq(N, 1.0 / N)
    kz = np.arange(N // 2 + 1, dtype=np.float64)
    return np.meshgrid(k, k, kz, indexing="ij")


def c2r(uh, N):
    """physical real field(s) from FFTW-normalised half-spectrum [N][N][N/2+1][C]"""
    return np.fft.irfftn(uh, s=(N, N, N), axes=(0, 1, 2))


def r2c(u):
    """FFTW-normalised half-spectrum of a real field [N][N][N][C]"""
    return np.fft.rfftn(u, axes=(0, 1, 2))


def sphere_mask(N, kcut):
    kx, ky, kz = wavenumbers(N)
    return (kx * kx + ky * ky + kz * kz) <= kcut * kcut


def downsample(uh, N, Nl, kc, gaussian=True):
    """F(u) on the LES grid: Gaussian (Delta = pi/k_c) times the sharp |k| <= Nl//3 cut, rescaled
    to the Nl^3 FFTW normalisation.  uh: [N][N][N/2+1][C] -> [Nl][Nl][Nl/2+1][C]"""
    C = uh.shape[-1]
    out = np.zeros((Nl, Nl, Nl // 2 + 1, C), dtype=np.complex128)
    kcut = Nl // 3
    kl = np.fft.fftfreq(Nl, 1.0 / Nl).astype(int)
    delta = math.pi / kc
    for a, kx in enumerate(kl):
        for b, ky in enumerate(kl):
            for kz in range(Nl // 2 + 1):
                k2 = kx * kx + ky * ky + kz * kz
                if k2 > kcut * kcut or abs(kx) > N // 2 - 1 or abs(ky) > N // 2 - 1:
                    continue
                g = math.exp(-k2 * delta * delta / 24.0) if gaussian else 1.0
                out[a, b, kz] = g * uh[kx % N, ky % N, kz] * (Nl / N) ** 3
    return out


def curl_free_grad(uh, N):
    """g[..., 3 i + j] = d_j u_i, physical, on the grid of uh"""
    kx, ky, kz = wavenumbers(N)
    ks = (kx, ky, kz)
    g = np.empty((N, N, N, 9))
    for i in range(3):
        for j in range(3):
            g[..., 3 * i + j] = c2r(1j * ks[j] * uh[..., i], N)
    return g


def sgs_sample(uh, N, Nl, kc):
    """(ubar_hat [Nl][Nl][Nl/2+1][3], grad ubar [Nl^3][9], deviatoric tau [Nl^3][6]) from a DNS
    state uh [N][N][N/2+1][3] (FFTW-normalised, 2/3-dealiased)"""
    ub = downsample(uh, N, Nl, kc)
    u = c2r(uh, N)  # physical
    prod = np.stack([u[..., i] * u[..., j] for i, j in TAU_ORDER], axis=-1)
    F = downsample(r2c(prod), N, Nl, kc)
    ubr = c2r(ub, Nl)
    pb = np.stack([ubr[..., i] * ubr[..., j] for i, j in TAU_ORDER], axis=-1)
    # products of |k| <= Nl//3 fields are alias-free on the Nl grid for |k| <= Nl//3
    Q = r2c(pb) * sphere_mask(Nl, Nl // 3)[..., None]
    tau = c2r(F - Q, Nl)
    tr = (tau[..., 0] + tau[..., 1] + tau[..., 2]) / 3.0
    for c in range(3):
        tau[..., c] -= tr
    g = curl_free_grad(ub, Nl)
    return ub, g.reshape(-1, 9), tau.reshape(-1, 6)


def closure_term(tau, Nl):
    """M_hat = -P div(tau) on |k| <= Nl//3 (FFTW-normalised); tau: [Nl^3][6] physical"""
    t = r2c(tau.reshape(Nl, Nl, Nl, 6))
    kx, ky, kz = wavenumbers(Nl)
    ks = (kx, ky, kz)
    full = {}
    for c, (i, j) in enumerate(TAU_ORDER):
        full[(i, j)] = full[(j, i)] = t[..., c]
    M = np.stack([-1j * sum(ks[j] * full[(i, j)] for j in range(3)) for i in range(3)], axis=-1)
    k2 = kx * kx + ky * ky + kz * kz
    kdot = (kx * M[..., 0] + ky * M[..., 1] + kz * M[..., 2]) / np.where(k2 > 0, k2, 1.0)
    for i in range(3):
        M[..., i] -= ks[i] * kdot
    M[k2 == 0] = 0.0
    return M * sphere_mask(Nl, Nl // 3)[..., None]


# ------------------------------------------------------------------------------------ network
def gelu(x):
    from scipy.special import erf

    return 0.5 * x * (1.0 + erf(x / math.sqrt(2.0)))


def gelu_grad(x):
    from scipy.special import erf

    return 0.5 * (1.0 + erf(x / math.sqrt(2.0))) + x * np.exp(-0.5 * x * x) / math.sqrt(2.0 * math.pi)


def conv3(x, W, b):
    """periodic 3x3x3 cross-correlation: x [B,N,N,N,Ci], W [Co,Ci,3,3,3] -> [B,N,N,N,Co]
    out[p] = b + sum_{a,b,c} W[:, :, a, b, c] @ x[p + (a-1, b-1, c-1)]"""
    out = np.zeros(x.shape[:-1] + (W.shape[0],))
    for a in range(3):
        for bb in range(3):
            for c in range(3):
                xs = np.roll(x, shift=(1 - a, 1 - bb, 1 - c), axis=(1, 2, 3))
                out += xs @ W[:, :, a, bb, c].T
    return out + b


def conv3_transpose(d, W):
    """adjoint of conv3 in x: d [B,N,N,N,Co] -> [B,N,N,N,Ci]"""
    out = np.zeros(d.shape[:-1] + (W.shape[1],))
    for a in range(3):
        for bb in range(3):
            for c in range(3):
                ds = np.roll(d, shift=(a - 1, bb - 1, c - 1), axis=(1, 2, 3))
                out += ds @ W[:, :, a, bb, c]
    return out


def conv3_weight_grad(x, d):
    """dL/dW for out = conv3(x): [Co,Ci,3,3,3]"""
    Ci, Co = x.shape[-1], d.shape[-1]
    g = np.zeros((Co, Ci, 3, 3, 3))
    df = d.reshape(-1, Co)
    for a in range(3):
        for bb in range(3):
            for c in range(3):
                xs = np.roll(x, shift=(1 - a, 1 - bb, 1 - c), axis=(1, 2, 3)).reshape(-1, Ci)
                g[:, :, a, bb, c] = df.T @ xs
    return g


def init_params(seed, channels=CHANNELS):
    """[(W, b)] with torch's default conv bound 1/sqrt(fan_in), from numpy's PCG64(seed)"""
    rng = np.random.default_rng(seed)
    params = []
    for ci, co in zip(channels[:-1], channels[1:]):
        bound = 1.0 / math.sqrt(ci * 27)
        params.append((rng.uniform(-bound, bound, (co, ci, 3, 3, 3)), rng.uniform(-bound, bound, co)))
    return params


def flatten(params):
    return np.concatenate([np.concatenate([W.ravel(), b.ravel()]) for W, b in params])


def unflatten(flat, channels=CHANNELS):
    params, o = [], 0
    for ci, co in zip(channels[:-1], channels[1:]):
        n = co * ci * 27
        W = flat[o:o + n].reshape(co, ci, 3, 3, 3)
        o += n
        b = flat[o:o + co]
        o += co
        params.append((W, b))
    return params


def forward(params, x, emulate_bf16=False):
    """x [B,N,N,N,9] -> (y [B,N,N,N,6], cache).  emulate_bf16 rounds the weights and every layer
    input to bfloat16 (accumulation stays exact)"""
    r = bf16 if emulate_bf16 else (lambda v: v)
    act = r(x)
    cache = [act]
    L = len(params)
    for l, (W, b) in enumerate(params):
        pre = conv3(act, r(W), b)
        if l < L - 1:
            cache.append(pre)
            act = r(gelu(pre))
            cache.append(act)
        else:
            return pre, cache


def loss_and_grads(params, x, t, emulate_bf16=False):
    """MSE over every point and output channel; returns (loss, [(dW, db)])"""
    r = bf16 if emulate_bf16 else (lambda v: v)
    y, cache = forward(params, x, emulate_bf16)
    n = y.size
    loss = float(np.sum((y - t) ** 2) / n)
    d = r(2.0 * (y - t) / n)  # dL/d(pre) of the last layer
    grads = [None] * len(params)
    for l in range(len(params) - 1, -1, -1):
        W, _ = params[l]
        act_in = cache[2 * l]
        grads[l] = (conv3_weight_grad(act_in, d), d.reshape(-1, d.shape[-1]).sum(axis=0))
        if l > 0:
            da = conv3_transpose(d, r(W))
            d = r(da * gelu_grad(cache[2 * l - 1]))
    return loss, grads
