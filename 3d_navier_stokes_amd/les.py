"""Learned-closure LES pipeline on one GPU, every step through libnsg.so.

1. Forced DNS (constant-power forcing eps_in on 0 < |k| <= 2.5, CFL-adaptive RK4, FP32-bulk) from
   a k^4 random-phase start.  The spin-up to the statistically steady state runs on a coarser
   grid, is resampled onto the DNS grid and settles there; the run is then sampled every `every`
   steps: each sample filters the DNS state onto the LES grid (Gaussian, cutoff k_c, plus the LES
   grid's 2/3 cut) and returns the network input grad(ub) and the target, the deviatoric subgrid
   stress; the filtered energy spectrum of every sample is kept as the reference.
2. Training of the closure network on those samples (BF16 MFMA forward/backward, Adam), the last
   samples held out.  The model's mean subgrid dissipation -<tau_ij S_ij> is compared with the
   true one on the training samples; their ratio is the gain of a calibrated copy of the model.
   Per shell of |k|, the ratio of the true to the model's subgrid transfer spectrum on the training
   samples is a second, per-shell calibration (applied to the closure term by the library).
3. LES runs on the coarse grid from the last filtered DNS state, same forcing and viscosity: with
   the trained closure, with its calibrated copy, with its per-shell calibration, without a model (the plain coarse DNS, the
   fallback), and without a model at the constant eddy viscosity
   nu_eff = eps_in (2 pi)^3 / (2 Omega_f) that balances the injection with the filtered field's
   enstrophy Omega_f (a volume integral, like every measure here).  Each run's time-averaged E(k)
   is compared with the filtered DNS E(k) on 1 <= k <= k_c.

Usage (one GPU):  python 3d_navier_stokes_amd/les.py --dns 512 --les 64 --out report.json
"""
from __future__ import annotations

import argparse
import importlib.util
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def load_nsg():
    if "navier_stokes_amd" in sys.modules:
        return sys.modules["navier_stokes_amd"]
    spec = importlib.util.spec_from_file_location("navier_stokes_amd", os.path.join(HERE, "__init__.py"),
                                                  submodule_search_locations=[HERE])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["navier_stokes_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def forced(nsg, N, nu, precision, eps=0.1, kf=2.5, cfl=0.5):
    return nsg.Solver(N, nu=nu, dt=1e-3, precision=precision, cfl=cfl, force_eps=eps, force_kf=kf)


def generate(nsg, Nd, Nl, kc, nu, spin_n, spin_time, settle, nsamples, every, seed=2, precision="fp32bulk", starts=1):
    """forced DNS -> (samples [(x, tau)], mean filtered E(k), filtered states to start LES runs from
    (`starts` of them, evenly spread over the window, the last one at its end), stats).  The
    statistically steady state is reached cheaply: the run starts on a spin_n^3 grid until
    t >= spin_time, is resampled (zero-padded) onto the Nd^3 grid and settles for `settle` steps
    before sampling starts."""
    t0 = time.perf_counter()
    spin = forced(nsg, spin_n, nu, precision)
    spin.init("k4", k0=2.0, energy=0.5, seed=seed)
    nspin = 0
    while spin.time()[0] < spin_time:
        spin.step(100)
        nspin += 100
        if nspin % 2000 == 0:
            log(f"[les] spin-up {nspin} steps, t = {spin.time()[0]:.3f}")
    dns = forced(nsg, Nd, nu, precision)
    spin.resample_to(dns)
    spin.close()
    dns.step(settle)
    t_spin = time.perf_counter() - t0
    log(f"[les] spin-up {nspin} steps at {spin_n}^3 + {settle} at {Nd}^3: {t_spin:.1f} s, t = {dns.time()[0]:.3f}")
    les = nsg.Solver(Nl, nu=nu)
    samples, spectra, enst, energy, ubs = [], [], [], [], []
    t_dns, n_dns, t_phys0 = 0.0, 0, dns.time()[0]
    for s in range(nsamples):
        t0 = time.perf_counter()
        dns.step(every)
        t_dns += time.perf_counter() - t0
        n_dns += every
        x, tau = dns.sgs_sample(les, kc)
        tot, ek, _ = les.measures()
        samples.append((x, tau))
        if (nsamples - 1 - s) % max(nsamples // starts, 1) == 0 and len(ubs) < starts:
            ubs.append(les.get_state())
        spectra.append(ek)
        enst.append(tot.enstrophy)
        energy.append(dns.measures()[0].energy)
        if (s + 1) % 16 == 0:
            log(f"[les] sample {s + 1}/{nsamples}, t = {dns.time()[0]:.3f}")
    t_phys = dns.time()[0] - t_phys0
    h = len(spectra) // 2
    e1, e2 = np.mean(spectra[:h], axis=0), np.mean(spectra[h:], axis=0)
    kk = np.arange(1, int(kc) + 1)
    stats = {"filtered_E_k_half_window_rel_diff": (np.abs(e1[kk] - e2[kk]) / (0.5 * (e1[kk] + e2[kk]))).tolist(),
             "spinup_steps_coarse": nspin, "spinup_grid": spin_n, "settle_steps": settle, "spinup_s": t_spin,
             "dns_ms_per_step": 1e3 * t_dns / n_dns, "dns_mean_dt": t_phys / n_dns,
             "dns_gpu_s_per_time_unit": t_dns / t_phys, "sampled_time_units": t_phys,
             "filtered_enstrophy": float(np.mean(enst)), "dns_energy_first_last": [energy[0], energy[-1]]}
    dns.close()
    les.close()
    return samples, np.mean(spectra, axis=0), ubs, stats


def sgs_dissipation(x, tau):
    """Pi = -<tau_ij S_ij>, S = sym(grad ub); x [n, 9] (x[3 i + j] = d_j u_i), tau [n, 6]"""
    g = x.astype(np.float64).reshape(-1, 3, 3)
    S = 0.5 * (g + g.transpose(0, 2, 1))
    t = tau.astype(np.float64)
    return -float(np.mean(t[:, 0] * S[:, 0, 0] + t[:, 1] * S[:, 1, 1] + t[:, 2] * S[:, 2, 2]
                          + 2 * (t[:, 3] * S[:, 0, 1] + t[:, 4] * S[:, 0, 2] + t[:, 5] * S[:, 1, 2])))


def rotations():
    """the 24 proper rotations of the cube (signed permutation matrices with det +1)"""
    import itertools
    out = []
    for perm in itertools.permutations(range(3)):
        for signs in itertools.product((1, -1), repeat=3):
            R = np.zeros((3, 3), dtype=np.int64)
            for a in range(3):
                R[a, perm[a]] = signs[a]
            if round(np.linalg.det(R)) == 1:
                out.append(R)
    return out


TAU_ORDER = ((0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2))


def rotate_sample(x, tau, N, R):
    """the sample seen in a frame rotated by R: fields move with the grid (x' = R x, periodic) and
    transform as tensors (g' = R g R^T, tau' = R tau R^T)"""
    q = np.indices((N, N, N)).reshape(3, -1)
    p = (R.T @ q) % N  # source point of each destination point
    src = (p[0] * N + p[1]) * N + p[2]
    g = x.astype(np.float64)[src].reshape(-1, 3, 3)
    g = np.einsum("ia,nab,jb->nij", R, g, R)
    t6 = tau.astype(np.float64)[src]
    T = np.empty((t6.shape[0], 3, 3))
    for c, (i, j) in enumerate(TAU_ORDER):
        T[:, i, j] = T[:, j, i] = t6[:, c]
    T = np.einsum("ia,nab,jb->nij", R, T, R)
    t_out = np.stack([T[:, i, j] for i, j in TAU_ORDER], axis=-1)
    return g.reshape(-1, 9).astype(np.float32), t_out.astype(np.float32)


def sgs_transfer(x, tau, N):
    """per-shell subgrid energy transfer T(b) = sum over the modes of the shell b = round(|k|) of
    Re(conj(S_ij) tau_ij) (half spectrum, the kz = 0 and N/2 planes counted once): the energy the
    term -P div(tau) moves into the resolved modes of that shell is -T(b) up to normalisation, since
    conj(u_i) (-i k_j tau_ij) = conj(i k_j u_i) tau_ij and tau is symmetric.  x [N^3, 9], tau [N^3, 6]
    (xx, yy, zz, xy, xz, yz); returns b = 0..N // 3."""
    g = x.astype(np.float64).reshape(N, N, N, 3, 3)
    S = 0.5 * (g + np.swapaxes(g, -1, -2))
    t = tau.astype(np.float64).reshape(N, N, N, 6)
    pairs = [(0, 0, 0, 1.0), (1, 1, 1, 1.0), (2, 2, 2, 1.0), (0, 1, 3, 2.0), (0, 2, 4, 2.0), (1, 2, 5, 2.0)]
    acc = 0.0
    for i, j, c, w in pairs:
        acc = acc + w * np.real(np.conj(np.fft.rfftn(S[..., i, j])) * np.fft.rfftn(t[..., c]))
    k = np.fft.fftfreq(N, 1.0 / N).astype(np.int64)
    kz = np.arange(N // 2 + 1, dtype=np.int64)
    k2 = k[:, None, None] ** 2 + k[None, :, None] ** 2 + kz[None, None, :] ** 2
    b = np.floor(np.sqrt(k2)).astype(np.int64)
    b += k2 > b * b + b  # round(|k|): k^2 in [b^2 - b + 1, b^2 + b]
    wz = np.where((kz == 0) | (kz == N // 2), 1.0, 2.0)[None, None, :]
    kmax = N // 3
    sel = b <= kmax
    return np.bincount(b[sel], weights=(acc * wz)[sel], minlength=kmax + 1)[: kmax + 1]


def train(nsg, samples, Nl, steps, batch, lr, holdout, seed=1, augment=0):
    """train on samples[:-holdout]; `augment` extra copies of each training sample in random cube
    rotations (the closure has no preferred axis)"""
    xs = np.stack([s[0] for s in samples])
    ts = np.stack([s[1] for s in samples])
    sig_in = float(np.sqrt(np.mean(xs.astype(np.float64) ** 2)))
    sig_out = [float(v) for v in np.sqrt(np.mean(ts.astype(np.float64) ** 2, axis=(0, 1)))]
    net = nsg.Network(Nl, max_batch=batch, seed=seed)
    net.set_norm(sig_in, sig_out)
    ntrain = len(samples) - holdout
    for x, t in samples:
        net.add_sample(x, t)
    l_train0 = net.eval(0, ntrain)
    l_hold0 = net.eval(ntrain, holdout) if holdout else None
    # rotated copies go after the held-out samples; training draws from [0, ntrain) and the copies
    rng = np.random.default_rng(seed)
    rots = rotations()[1:]
    for i, (x, t) in enumerate(samples[:ntrain]):
        for r in rng.choice(len(rots), size=augment, replace=False):
            net.add_sample(*rotate_sample(x, t, Nl, rots[r]))
        if augment and (i + 1) % 32 == 0:
            log(f"[les] rotated copies of {i + 1}/{ntrain} training samples added")
    nall = net.samples
    t0 = time.perf_counter()
    hist = []
    chunk = 50
    done = 0
    while done < steps:
        k = min(chunk, steps - done)
        rate = lr * (0.1 if done >= 0.8 * steps else 1.0)
        if augment:
            # [0, ntrain) and the rotated copies [ntrain + holdout, nall): alternate the two ranges
            # in proportion to their sizes
            n_aug = nall - ntrain - holdout
            use_aug = ((done // chunk) % (augment + 1)) != 0
            first, count = (ntrain + holdout, n_aug) if use_aug else (0, ntrain)
        else:
            first, count = 0, ntrain
        hist.extend(net.train(k, batch=batch, lr=rate, seed=seed + done, first=first, count=count).tolist())
        done += k
        if done % 1000 == 0:
            log(f"[les] training step {done}/{steps}, loss {np.mean(hist[-50:]):.4f}")
    t_train = time.perf_counter() - t0
    res = {"sigma_in": sig_in, "sigma_out": sig_out, "train_samples": ntrain, "holdout_samples": holdout,
           "loss_train_initial": l_train0, "loss_train_final": net.eval(0, ntrain),
           "loss_holdout_initial": l_hold0, "loss_holdout_final": net.eval(ntrain, holdout) if holdout else None,
           "steps": steps, "batch": batch, "lr": lr, "train_s": t_train, "rotated_copies_per_sample": augment,
           "loss_history_every_50": hist[::50]}
    # SGS dissipation of the model against the truth on the training samples; the gain that
    # matches them calibrates the output (a-priori, training data only)
    # and per shell, the subgrid transfer spectra of the two: their ratio (clipped to [1/2, 2], 1 where
    # the model's transfer is negligible) is the per-shell gain of a second calibration
    pi_true, pi_model = 0.0, 0.0
    tr_true, tr_model = 0.0, 0.0
    for x, t in samples[:ntrain]:
        y = net.forward((x / sig_in)[None])[0] * np.array(sig_out)
        pi_true += sgs_dissipation(x, t)
        pi_model += sgs_dissipation(x, y)
        tr_true = tr_true + sgs_transfer(x, t, Nl)
        tr_model = tr_model + sgs_transfer(x, y, Nl)
    ok = np.abs(tr_model) > 1e-6 * np.max(np.abs(tr_model))
    shell_gain = np.where(ok, tr_true / np.where(ok, tr_model, 1.0), 1.0)
    shell_gain = np.clip(shell_gain, 0.5, 2.0)
    shell_gain[0] = 1.0
    res.update({"sgs_dissipation_true": pi_true / ntrain, "sgs_dissipation_model": pi_model / ntrain,
                "dissipation_gain": pi_true / pi_model if pi_model > 0 else 1.0,
                "sgs_transfer_true_by_shell": (tr_true / ntrain).tolist(),
                "sgs_transfer_model_by_shell": (tr_model / ntrain).tolist(),
                "shell_gain": shell_gain.tolist()})
    return net, res


def scaled_network(nsg, net, gain, Nl, batch):
    """a copy of `net` whose output is multiplied by `gain` (last layer's weights and bias)"""
    p = net.get_params().astype(np.float64)
    n_last = 6 * 32 * 27 + 6
    p[-n_last:] *= gain
    out = nsg.Network(Nl, max_batch=batch)
    si, so = net.norm
    out.set_norm(si, so)
    out.set_params(p.astype(np.float32))
    return out


def run_les(nsg, Nl, ub, nu, steps, net=None, avg_from=0.2, eps=0.1, gain=None):
    s = forced(nsg, Nl, nu, "fp64")
    s.set_state(ub)
    if net is not None:
        s.attach_closure(net)
        if gain is not None:
            s.set_closure_gain(gain)
    spectra, t0w = [], time.perf_counter()
    t_start = s.time()[0]
    blow = None
    nchunk = 50
    for i in range(0, steps, nchunk):
        try:
            s.step(nchunk)
        except nsg.NSGError as e:
            blow = f"step {i}: {e}"
            break
        if i >= avg_from * steps:
            spectra.append(s.measures()[1])
    wall = time.perf_counter() - t0w
    t_phys = s.time()[0] - t_start
    tot = s.measures()[0] if blow is None else None
    s.close()
    ek = np.mean(spectra, axis=0) if spectra else None
    return {"E_k": ek, "blowup": blow, "gpu_s": wall, "time_units": t_phys,
            "ms_per_step": 1e3 * wall / steps, "final_energy": tot.energy if tot else None}


def spectral_error(ek, ref, kc):
    k = np.arange(1, int(kc) + 1)
    if ek is None:
        return None, None
    r = np.abs(ek[k] - ref[k]) / ref[k]
    return float(np.max(r)), float(np.mean(r))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--dns", type=int, default=512)
    ap.add_argument("--les", type=int, default=64)
    ap.add_argument("--kc", type=float, default=None, help="filter cutoff (default les // 3)")
    ap.add_argument("--nu", type=float, default=1e-3)
    ap.add_argument("--spin-n", type=int, default=256, help="grid of the spin-up run")
    ap.add_argument("--spin-time", type=float, default=20.0, help="spin-up length in time units")
    ap.add_argument("--settle", type=int, default=1500, help="steps on the DNS grid before sampling")
    ap.add_argument("--samples", type=int, default=96)
    ap.add_argument("--every", type=int, default=80)
    ap.add_argument("--holdout", type=int, default=8)
    ap.add_argument("--train-steps", type=int, default=3000)
    ap.add_argument("--augment", type=int, default=3, help="rotated copies of each training sample")
    ap.add_argument("--batch", type=int, default=4)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--les-steps", type=int, default=5000)
    ap.add_argument("--les-starts", type=int, default=1,
                    help="LES runs per variant, from filtered states spread over the window (E(k) averaged)")
    ap.add_argument("--out", default=None)
    a = ap.parse_args()
    kc = a.kc if a.kc else a.les // 3
    nsg = load_nsg()
    samples, ek_ref, ubs, dstats = generate(nsg, a.dns, a.les, kc, a.nu, a.spin_n, a.spin_time, a.settle, a.samples,
                                            a.every, starts=a.les_starts)
    log(f"[les] {len(samples)} samples; DNS {dstats['dns_ms_per_step']:.2f} ms/step")
    net, tstats = train(nsg, samples, a.les, a.train_steps, a.batch, a.lr, a.holdout, augment=a.augment)
    log(f"[les] training: loss {tstats['loss_train_initial']:.4f} -> {tstats['loss_train_final']:.4f}, holdout "
        f"{tstats['loss_holdout_initial']} -> {tstats['loss_holdout_final']}, {tstats['train_s']:.1f} s")
    # measures are volume integrals over (2 pi)^3 while eps_in is a volume average
    nu_eff = 0.1 * (2.0 * np.pi) ** 3 / (2.0 * dstats["filtered_enstrophy"])
    cal = scaled_network(nsg, net, tstats["dissipation_gain"], a.les, a.batch)
    runs = {}
    gain = np.array(tstats["shell_gain"])
    for name, net_, nu_, g_ in [("cnn_closure", net, a.nu, None), ("cnn_closure_dissipation_calibrated", cal, a.nu, None),
                                ("cnn_closure_shell_calibrated", net, a.nu, gain),
                                ("no_model", None, a.nu, None), ("eddy_viscosity", None, nu_eff, None)]:
        per = [run_les(nsg, a.les, ub, nu_, a.les_steps, net_, gain=g_) for ub in ubs]
        r = dict(per[-1])
        eks = [p["E_k"] for p in per if p["E_k"] is not None]
        r["E_k"] = np.mean(eks, axis=0) if eks else None
        if len(per) > 1:
            r["E_k_per_start"] = [p["E_k"].tolist() if p["E_k"] is not None else None for p in per]
            r["E_k_max_rel_err_per_start"] = [spectral_error(p["E_k"], ek_ref, kc)[0] for p in per]
            r["blowup"] = [p["blowup"] for p in per]
        emax, emean = spectral_error(r["E_k"], ek_ref, kc)
        r.update({"nu": nu_, "E_k_max_rel_err_k_le_kc": emax, "E_k_mean_rel_err_k_le_kc": emean, "starts": len(per)})
        log(f"[les] {name}: max rel E(k) err {emax}, mean {emean}, {r['ms_per_step']:.3f} ms/step, blowup {r['blowup']}")
        runs[name] = r
    les_s_per_tu = runs["cnn_closure"]["gpu_s"] / max(runs["cnn_closure"]["time_units"], 1e-30)
    report = {
        "config": {"dns": a.dns, "les": a.les, "kc": kc, "nu": a.nu, "forcing": "eps_in=0.1 on |k|<=2.5",
                   "spin_n": a.spin_n, "spin_time": a.spin_time, "settle": a.settle, "samples": a.samples,
                   "sample_every": a.every,
                   "les_steps": a.les_steps, "dns_precision": "fp32bulk", "les_precision": "fp64"},
        "dns": dstats, "training": tstats,
        "filtered_dns_E_k": ek_ref[: int(kc) + 2].tolist(),
        "les": {k: {kk: (vv.tolist() if isinstance(vv, np.ndarray) else vv) for kk, vv in v.items()} for k, v in runs.items()},
        "gpu_time_ratio_dns_over_les_per_time_unit": dstats["dns_gpu_s_per_time_unit"] / les_s_per_tu,
    }
    line = json.dumps(report)
    if a.out:
        with open(a.out, "w") as f:
            f.write(line + "\n")
    print(line)


if __name__ == "__main__":
    main()
