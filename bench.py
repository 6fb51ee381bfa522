#!/usr/bin/env python3
"""Benchmark: RK4 steps/s of the forced 512^3 box, FP64 vs FP32-bulk, on one MI355X.

Workload (the 512^3 forced stationary turbulence configuration): periodic [0,2pi)^3, N=512,
nu=1e-3, constant-power forcing eps_in=0.1 on 0<|k|<=2.5, RK4 with the reference's Crank-Nicolson
viscous factor, CFL 0.5 adaptive step computed on the device, random-phase k^4 exp(-2(k/2)^2)
start (Philox seed 2, 1/2<|u|^2> = 0.5).  The state is resident in HBM before timing starts.

Prints one JSON line.  `value` is FP32-bulk steps/s; the FP64 plain path is measured the same way
and reported under "fp64".  `roofline` describes the kernel with the largest share of step time,
timed with HIP events on the library's own stream.  `cpu_baseline` times the FP64 CPU oracle
(a C restatement of the reference algorithm, OpenMP) on a bounded sample: one full RK4-CN step of
the same forced workload at 512^3.  `--cpu-plan` runs the whole CPU plan instead (256^3 and 512^3
median of 3 steps after 1 warm-up, and the Taylor-Green configuration end to end).

"ensemble_96_boxes" reports the batched-ensemble workload: 96 independent FP32-bulk boxes (32 at
32^3, 32 at 48^3, 24 at 64^3, 8 at 96^3), 100 RK4 steps, autotuned launch schedules, one HIP graph
replay per ensemble step, against the same steps launched box by box.

"integrators" times the Dormand-Prince schemes on the same forced 512^3 workload: RK5-CN (six
stages, fixed dt = the RK4 run's mean CFL step) in FP64 and FP32-bulk, and the adaptive DP5 loop
(seven stages per try, FP64) from that dt.

"increment_stats" times the post-processing statistics (longitudinal velocity increments at nine
separations: moments pass and 1000-bin histogram pass) of the bench's 512^3 FP32-bulk final state
in real space.

"closure_les" times the learned-closure LES on the 64^3 grid (the forced DNS state filtered with
cutoff 21, the BF16 MFMA network evaluated every step) and reports the ratio of GPU seconds per
unit of simulated time, DNS over LES, plus the network's forward throughput against the MI355X
dense BF16 peak.
"""
from __future__ import annotations

import argparse
import importlib.util
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
HBM_PEAK_GBS = 8000.0
METRIC = "RK4 steps/s at 512^3 FP64 vs FP32-bulk, % HBM peak; closure-LES speedup"
PASS_NAMES = ["pass1_invx_curl", "pass2_invy", "pass3_z_c2r_uxw_r2c", "pass4_fwdy", "pass5_fwdx_epilogue"]


def load_nsg():
    pkg = os.path.join(ROOT, "3d_navier_stokes_amd")
    spec = importlib.util.spec_from_file_location("navier_stokes_amd", os.path.join(pkg, "__init__.py"),
                                                  submodule_search_locations=[pkg])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["navier_stokes_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def log(msg):
    print(msg, file=sys.stderr, flush=True)


KERNEL_OF_PASS = [r"pass1_kernel<{T}, {N},", r"passy_kernel<{T}, {N}, .*, 1>$", r"passz_(rhs|full)_kernel<{T}, {N},",
                  r"passy_kernel<{T}, {N}, .*, -1>$", r"pass5p?_kernel<{T}, {N},"]


def measured_traffic(N, precision, pass_index):
    """HBM bytes per launch of a pass kernel from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE
    summary (calibrated against a plain device copy), or None when no summary covers it."""
    import re

    path = os.path.join(ROOT, "profiles", f"step{N}_hbm_traffic.json")
    if not os.path.exists(path):
        return None
    pat = KERNEL_OF_PASS[pass_index].format(T="double" if precision == "fp64" else "float", N=N)
    for name, v in json.load(open(path))["kernels"].items():
        if re.match(pat, name):
            return v["traffic_bytes"]
    return None


def run_mode(nsg, N, precision, steps, warmup, prof_iters):
    s = nsg.Solver(N, nu=1e-3, dt=1e-3, scheme="rk4cn", precision=precision, cfl=0.5, force_eps=0.1,
                   force_kf=2.5)
    s.init("k4", k0=2.0, energy=0.5, seed=2)
    log(f"[bench] {precision}: warmup {warmup} steps")
    s.step(warmup)
    tb = s.time()[0]
    t0 = time.perf_counter()
    s.step(steps)  # returns after the stream has drained
    el = time.perf_counter() - t0
    mean_dt = (s.time()[0] - tb) / steps
    ms, by = s.profile(prof_iters)
    log(f"[bench] {precision}: {steps / el:.2f} steps/s; pass ms {['%.3f' % m for m in ms[:5]]}")
    passes = {}
    for i, name in enumerate(PASS_NAMES):
        gbs = by[i] / (ms[i] * 1e-3) / 1e9 if ms[i] > 0 else 0.0
        passes[name] = {"ms": ms[i], "bytes": by[i], "GBps": gbs, "frac": gbs / HBM_PEAK_GBS}
    dom = max(range(5), key=lambda i: ms[i])
    achieved = by[dom] / (ms[dom] * 1e-3) / 1e9
    roof = {"bound": "hbm", "kernel": PASS_NAMES[dom], "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": measured_traffic(N, precision, dom),
            "bytes_per_launch": by[dom], "ms_per_launch": ms[dom]}
    step_gb = s.step_bytes() / 1e9
    out = {"steps_per_s": steps / el, "ms_per_step": 1e3 * el / steps, "mean_dt": mean_dt, "roofline": roof, "passes": passes,
           "step_bytes_GB": step_gb, "step_GBps": step_gb / (el / steps)}
    return s, out


def run_integrators(nsg, N, steps, warmup, dt_cfl):
    """the Dormand-Prince integrators on the same forced workload: RK5-CN (fixed dt = the RK4 run's
    mean CFL dt) in FP64 and FP32-bulk, and adaptive DP5 in FP64 from that dt (accepted steps and
    simulated time per second, rejections included)"""
    out = {}
    for prec in ("fp64", "fp32bulk"):
        s = nsg.Solver(N, nu=1e-3, dt=dt_cfl, scheme="rk5cn", precision=prec, force_eps=0.1, force_kf=2.5)
        s.init("k4", k0=2.0, energy=0.5, seed=2)
        s.step(warmup)
        t0 = time.perf_counter()
        s.step(steps)
        el = time.perf_counter() - t0
        ms, by = s.profile(1)
        gb = s.step_bytes() / 1e9
        out[f"rk5cn_{prec}"] = {"steps_per_s": steps / el, "ms_per_step": 1e3 * el / steps, "step_bytes_GB": gb,
                                "step_GBps": gb / (el / steps),
                                "pass_GBps": {n: by[i] / (ms[i] * 1e-3) / 1e9 for i, n in enumerate(PASS_NAMES)}}
        log(f"[bench] rk5cn {prec}: {steps / el:.2f} steps/s, {gb / (el / steps):.0f} GB/s over the step")
        s.close()
    s = nsg.Solver(N, nu=1e-3, dt=dt_cfl, scheme="dp5", force_eps=0.1, force_kf=2.5)
    s.init("k4", k0=2.0, energy=0.5, seed=2)
    s.step(warmup)
    a0, r0, _ = s.dp_stats()
    tb = s.time()[0]
    n = max(2, steps // 2)
    t0 = time.perf_counter()
    s.step(n)
    el = time.perf_counter() - t0
    a1, r1, err = s.dp_stats()
    t1, dt1, _ = s.time()
    out["dp5_fp64"] = {"accepted_steps_per_s": n / el, "tries_per_step": (n + r1 - r0) / n,
                       "sim_time_per_s": (t1 - tb) / el, "mean_dt": (t1 - tb) / n, "dt_next": dt1, "last_err": err,
                       "rk4_cfl_dt": dt_cfl}
    log(f"[bench] dp5 fp64: {n / el:.2f} accepted steps/s, {(n + r1 - r0) / n:.2f} tries/step, mean dt "
        f"{(t1 - tb) / n:.3e} (RK4 CFL dt {dt_cfl:.3e})")
    s.close()
    return out


def run_incr_stats(nsg, u, iters=3, solver=None):
    """velocity-increment statistics (the reference's post-processing) of one 512^3 real-space
    snapshot resident in HBM: moment pass and histogram pass, each three kernels (one per
    direction) streaming the AoS field, plus a bounded CPU-oracle sample for scale"""
    N = u.shape[0]
    with nsg.IncrementStats(N) as st:
        st.precompute(u)
        st.set_ranges()
        st.histogram(u)
        stats, _, counts = st.results(ranges=False)
        ms_m = st.time_pass(False, iters)
        ms_h = st.time_pass(True, iters)
    state = None
    if solver is not None:
        # in situ: the solver's c2r writes the field as three planes, each direction's kernel reads
        # its own component; wall time of the calls (c2r of the state included, one sync each)
        with nsg.IncrementStats(N) as st:
            st.precompute(solver)
            t0 = time.perf_counter()
            for _ in range(iters):
                st.precompute(solver)
            ms_sm = 1e3 * (time.perf_counter() - t0) / iters
            st.set_ranges()
            st.histogram(solver)
            t0 = time.perf_counter()
            for _ in range(iters):
                st.histogram(solver)
            ms_sh = 1e3 * (time.perf_counter() - t0) / iters
        state = {"moments_ms_incl_c2r": ms_sm, "histogram_ms_incl_c2r": ms_sh}
    pts = float(N) ** 3
    out = {"N": N, "moments_ms": ms_m, "histogram_ms": ms_h,
           "field_GB": 24 * pts / 1e9, "moments_field_GBps": 24 * pts / (ms_m * 1e-3) / 1e9,
           "histogram_field_GBps": 24 * pts / (ms_h * 1e-3) / 1e9,
           "moments_stream_GBps": 72 * pts / (ms_m * 1e-3) / 1e9,
           "histogram_stream_GBps": 72 * pts / (ms_h * 1e-3) / 1e9,
           "from_solver_state": state,
           "increments_per_pass": float(counts.sum()),
           "r1_stats_min_max_mean_sd_skew_kurt": [float(x) for x in stats[0]]}
    log(f"[bench] increment stats {N}^3: moments {ms_m:.2f} ms, histogram {ms_h:.2f} ms per snapshot "
        f"({out['moments_stream_GBps']:.0f} / {out['histogram_stream_GBps']:.0f} GB/s streamed)"
        + (f"; from the solver state (c2r included) {state['moments_ms_incl_c2r']:.2f} / "
           f"{state['histogram_ms_incl_c2r']:.2f} ms" if state else ""))
    return out


ENSEMBLE_SIZES = [32] * 32 + [48] * 32 + [64] * 24 + [96] * 8


def ensemble_boxes():
    """96 uneven boxes: 32 at 32^3, 32 at 48^3, 24 at 64^3, 8 at 96^3; box b: nu = 1e-3 + 4e-3 b/95,
    FP32-bulk, RK4, CFL 0.5 (k^4 start: seed 1000+b, k0 = 2 + b mod 3)"""
    return [dict(N=N, nu=1e-3 + 4e-3 * b / 95, dt=1e-3, precision="fp32bulk", cfl=0.5)
            for b, N in enumerate(ENSEMBLE_SIZES)]


def run_ensemble(nsg, steps, warmup):
    """aggregate box-steps/s of the 96-box ensemble: autotuned schedules, one HIP graph per step, vs
    the same steps launched box by box with direct launches (same schedules)"""
    boxes = ensemble_boxes()

    def fresh(e):
        for b in range(len(boxes)):
            e.init(b, "k4", k0=2.0 + b % 3, energy=0.5, seed=1000 + b)

    with nsg.Ensemble(boxes) as e:
        fresh(e)
        log("[bench] ensemble: autotune")
        chosen = e.autotune(iters=3, use_cache=False)
        fresh(e)
        e.step(warmup)
        ms_graph = e.time_steps(steps, serial=False)
        gbytes = e.step_bytes()
    with nsg.Ensemble(boxes) as r:
        for g, c in enumerate(chosen):
            r.set_schedule(g, c)
        fresh(r)
        r.step_serial(warmup)
        ms_serial = r.time_steps(steps, serial=True)
    nb = len(boxes)
    out = {"boxes": nb, "sizes": {"32": 32, "48": 32, "64": 24, "96": 8}, "steps": steps,
           "schedules": dict(zip(["32", "48", "64", "96"], chosen)),
           "graph_ms_per_ensemble_step": ms_graph / steps,
           "graph_box_steps_per_s": nb * steps / (ms_graph * 1e-3),
           "serial_ms_per_ensemble_step": ms_serial / steps,
           "serial_box_steps_per_s": nb * steps / (ms_serial * 1e-3),
           "speedup_vs_serial": ms_serial / ms_graph,
           "algorithmic_GB_per_ensemble_step": gbytes / 1e9,
           "graph_GBps": gbytes / (ms_graph * 1e-3 / steps) / 1e9}
    log(f"[bench] ensemble: graph {out['graph_box_steps_per_s']:.0f} box-steps/s, serial "
        f"{out['serial_box_steps_per_s']:.0f}, x{out['speedup_vs_serial']:.2f}")
    return out


CNN_CH = (9, 32, 32, 32, 6)
BF16_PEAK_TFLOPS = 2500.0


def run_closure(nsg, dns, r_dns, Nl=64, steps=200, warmup=20):
    """closure-LES cost against the DNS: the DNS state filtered onto the Nl^3 grid, the network
    attached (its cost does not depend on the weights: small random ones keep the run tame), and
    GPU seconds per unit of simulated time of both runs; plus the network's MFMA throughput"""
    les = nsg.Solver(Nl, nu=1e-3, dt=1e-3, precision="fp64", cfl=0.5, force_eps=0.1, force_kf=2.5)
    dns.sgs_sample(les, Nl // 3, want_x=False, want_tau=False)
    t_dns, dt_dns, _ = dns.time()
    les.set_time(t_dns, dt_dns * dns.N / Nl)
    net = nsg.Network(Nl, seed=1)
    net.set_norm(1.8, [0.035] * 6)
    net.set_params(net.get_params() * 0.1)
    les.attach_closure(net)
    les.step(warmup)
    tb = les.time()[0]
    t0 = time.perf_counter()
    les.step(steps)
    el = time.perf_counter() - t0
    dt_les = (les.time()[0] - tb) / steps
    ms_fwd = net.time_forward(1, 20)
    les.close()
    net.close()
    # the same forward over a training batch of 8 samples (more items per CU than one sample's)
    net8 = nsg.Network(Nl, max_batch=8, seed=1)
    ms_fwd8 = net8.time_forward(8, 10)
    net8.close()
    flop = 2.0 * Nl ** 3 * 27 * sum(a * b for a, b in zip(CNN_CH[:-1], CNN_CH[1:]))
    tflops = flop / (ms_fwd * 1e-3) / 1e12
    tflops8 = 8 * flop / (ms_fwd8 * 1e-3) / 1e12
    dns_s_per_tu = r_dns["ms_per_step"] * 1e-3 / r_dns["mean_dt"]
    les_s_per_tu = el / steps / dt_les
    out = {"les_grid": Nl, "les_ms_per_step": 1e3 * el / steps, "les_mean_dt": dt_les,
           "dns_ms_per_step": r_dns["ms_per_step"], "dns_mean_dt": r_dns["mean_dt"],
           "gpu_time_ratio_dns_over_les": dns_s_per_tu / les_s_per_tu,
           "cnn_forward_ms": ms_fwd, "cnn_forward_GFLOP": flop / 1e9,
           "roofline": {"bound": "mfma", "kernel": "conv3p_kernel (4 layers, one forward, 1 sample)",
                        "achieved": round(tflops, 2), "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                        "frac": round(tflops / BF16_PEAK_TFLOPS, 5), "traffic": None},
           "cnn_forward_batch8": {"ms": ms_fwd8, "achieved_TFLOPs": round(tflops8, 2),
                                  "frac": round(tflops8 / BF16_PEAK_TFLOPS, 5)}}
    log(f"[bench] closure LES {Nl}^3: {out['les_ms_per_step']:.3f} ms/step, GPU-time ratio DNS/LES "
        f"{out['gpu_time_ratio_dns_over_les']:.0f}, CNN forward {ms_fwd:.3f} ms ({tflops:.1f} TFLOP/s), "
        f"batch 8 {ms_fwd8:.3f} ms ({tflops8:.1f} TFLOP/s)")
    return out


def _cpu_host():
    """host description for the CPU baseline: threads used, logical CPUs, CPU model"""
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return os.cpu_count() or 1, model


def _oracle_step_time(O, uh, dt, nsteps=1):
    t0 = time.perf_counter()
    O.step(uh, 1e-3, dt, 1, "rk4cn", 0.1, 2.5, nsteps=nsteps)
    return time.perf_counter() - t0


def cpu_baseline(N):
    """Bounded sample of the BASELINE CPU plan: one full FP64 RK4-CN step (4 nonlinear terms, stage
    combinations, forcing, Crank-Nicolson update) of the forced workload at N^3 on the oracle (the
    C restatement of the reference algorithm; OpenMP; its own batched radix-4/3/2 FFT, no FFTW in the
    image), after a 64^3 warm-up step."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O  # the checker; timed here only as the reported CPU baseline

    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    O.set_threads(threads)
    _oracle_step_time(O, O.ic_k4(64, k0=2.0, energy=0.5, seed=2), 1e-3)
    uh = O.ic_k4(N, k0=2.0, energy=0.5, seed=2)
    el = _oracle_step_time(O, uh, O.cfl_dt(uh, 0.5))
    del uh
    # the post-processing statistics restatement (sequential, one core) on a 128^3 snapshot
    n_st = 128
    u = O.realspace(O.ic_k4(n_st, 3.0, 0.5, 5), 0)
    t1 = time.perf_counter()
    O.incr_moments(u)
    el_st = time.perf_counter() - t1
    ncpu, model = _cpu_host()
    return {"value": 1.0 / el, "unit": "steps/s", "cores": threads, "kind": "port",
            "sample": f"one full FP64 RK4-CN step of the forced workload at {N}^3 (4 nonlinear terms = 36 3-D real "
                      f"FFTs, stage combinations, forcing, update) took {el:.2f} s on {threads} threads "
                      f"({ncpu} logical CPUs, {model}); own batched radix-4/3/2 FFT (no FFTW in the image)",
            "increment_stats_moments": {"points_per_s": n_st ** 3 / el_st, "cores": 1,
                                        "sample": f"one moments pass over a {n_st}^3 snapshot took {el_st:.3f} s"}}


def cpu_plan():
    """The whole CPU-baseline plan: FP64 RK4-CN wall time per step at 256^3 and 512^3 (median of 3
    after 1 warm-up step, the forced workload) and the Taylor-Green configuration end to end (64^3,
    nu = 1/1600, dt = 1e-3, 200 steps with the measures of the 21 saves)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O

    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    O.set_threads(threads)
    out = {}
    for N in (256, 512):
        uh = O.ic_k4(N, k0=2.0, energy=0.5, seed=2)
        dt = O.cfl_dt(uh, 0.5)
        _oracle_step_time(O, uh, dt)
        ts = sorted(_oracle_step_time(O, uh, dt) for _ in range(3))
        out[f"rk4cn_fp64_{N}"] = {"s_per_step_median": ts[1], "s_per_step": ts, "steps_per_s": 1.0 / ts[1]}
        log(f"[bench] cpu plan {N}^3: {ts[1]:.2f} s/step (median of 3)")
        del uh
    t0 = time.perf_counter()
    u = O.ic_taylor_green(64)
    O.measures(u, 1.0 / 1600.0)
    for _ in range(20):
        u = O.step(u, 1.0 / 1600.0, 1e-3, nsteps=10)
        O.measures(u, 1.0 / 1600.0)
    el = time.perf_counter() - t0
    out["taylor_green_64_200_steps"] = {"s": el, "steps_per_s": 200.0 / el}
    log(f"[bench] cpu plan Taylor-Green 64^3 200 steps: {el:.2f} s")
    ncpu, model = _cpu_host()
    out["host"] = {"threads": threads, "logical_cpus": ncpu, "cpu": model,
                   "fft": "oracle's own batched radix-4/3/2 transform (FFTW absent)"}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--prof-iters", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-n", type=int, default=512)
    ap.add_argument("--cpu-plan", action="store_true",
                    help="only run the whole CPU-baseline plan (256^3/512^3 steps, Taylor-Green end to end)")
    ap.add_argument("--ensemble-steps", type=int, default=100, help="0 skips the 96-box ensemble")
    ap.add_argument("--no-closure", dest="closure", action="store_false", help="skip the closure-LES timing")
    ap.add_argument("--no-incr-stats", dest="incr_stats", action="store_false",
                    help="skip the velocity-increment statistics timing")
    ap.add_argument("--no-integrators", dest="integrators", action="store_false",
                    help="skip the RK5-CN / Dormand-Prince timing")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="only run one device-to-device copy of this many MiB (HBM counter calibration)")
    ap.add_argument("--incr-only", type=int, default=0,
                    help="only time the increment statistics of a k^4 start field of this size")
    args = ap.parse_args()
    if args.cpu_plan:
        print(json.dumps({"cpu_plan": cpu_plan()}), flush=True)
        return
    if args.incr_only:
        nsg = load_nsg()
        with nsg.Solver(args.incr_only, precision="fp32bulk") as s:
            s.init("k4", k0=2.0, energy=0.5, seed=2)
            u = s.realspace(0)
            print(json.dumps({"increment_stats": run_incr_stats(nsg, u, solver=s)}))
        return
    if args.copy_probe_mib:
        import torch

        n = args.copy_probe_mib << 20
        a = torch.ones(n, dtype=torch.uint8, device="cuda:0")
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": n}))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs one box on one GPU; --gpus N>1 is not supported "
                                   "(multi-GPU slab decomposition is a design note only)"}))
        sys.exit(2)

    nsg = load_nsg()
    N = args.n
    s64, r64 = run_mode(nsg, N, "fp64", args.steps, args.warmup, args.prof_iters)
    s32, r32 = run_mode(nsg, N, "fp32bulk", args.steps, args.warmup, args.prof_iters)

    # FP32-bulk error vs FP64 from one identical state: one step each, energy spectrum and totals
    s32.set_state(s64.get_state())
    t64, dt64, _ = s64.time()
    s32.set_time(t64, dt64)
    s64.step(1)
    s32.step(1)
    m64, ek64, _ = s64.measures()
    m32, ek32, _ = s32.measures()
    # shells holding at least 1e-10 of the peak shell's energy: below that, FP32 round-off of the
    # transforms is an absolute noise floor and a relative error measures the floor
    sel = ek64 >= 1e-10 * ek64.max()
    ek_err = float(np.max(np.abs(ek32[sel] - ek64[sel]) / ek64[sel]))
    kc = N // 3
    k = np.arange(1, kc + 1)
    selk = k[sel[1:kc + 1]]
    ek_err_kc = float(np.max(np.abs(ek32[selk] - ek64[selk]) / ek64[selk]))
    n_resolved = int(np.count_nonzero(sel))
    u64 = s64.get_state()
    u32 = s32.get_state()
    state_err = float(np.linalg.norm(u32 - u64) / np.linalg.norm(u64))
    s64.close()
    clos = run_closure(nsg, s32, r32) if args.closure else None
    incr = run_incr_stats(nsg, s32.realspace(0), solver=s32) if args.incr_stats else None
    s32.close()
    integ = run_integrators(nsg, N, max(4, args.steps // 2), 2, r64["mean_dt"]) if args.integrators else None

    ens = run_ensemble(nsg, args.ensemble_steps, 3) if args.ensemble_steps > 0 else None

    cpu = None
    if not args.no_cpu_baseline:
        log(f"[bench] cpu baseline at {args.cpu_n}^3 ...")
        cpu = cpu_baseline(args.cpu_n)
        log(f"[bench] cpu baseline {cpu['value']:.4f} steps/s")

    line = {
        "metric": METRIC,
        "value": round(r32["steps_per_s"], 3),
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r32["ms_per_step"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32-bulk (f32 FFTs/products, f64 state/update/reductions)",
        "data": "synthetic",
        "config": {"workload": f"forced isotropic turbulence {N}^3, nu=1e-3, constant-power forcing eps_in=0.1 "
                               f"on |k|<=2.5, RK4 (reference RK4-CN), CFL 0.5, k^4 random-phase start seed 2",
                   "N": N, "scheme": "rk4cn", "parallelism": "single box, one GPU"},
        "fp64": {"value": round(r64["steps_per_s"], 3), "ms_per_step": round(r64["ms_per_step"], 3),
                 "roofline": r64["roofline"], "passes": r64["passes"], "step_bytes_GB": r64["step_bytes_GB"],
                 "step_GBps": r64["step_GBps"]},
        "fp32bulk_over_fp64": round(r32["steps_per_s"] / r64["steps_per_s"], 3),
        "fp32bulk_error_vs_fp64_one_step": {"E_k_max_rel_resolved_shells": ek_err, "E_k_max_rel_k_le_kc": ek_err_kc,
                                            "resolved_shells": n_resolved,
                                            "energy_rel": abs(m32.energy - m64.energy) / m64.energy,
                                            "dissipation_rel": abs(m32.dissipation - m64.dissipation) / m64.dissipation,
                                            "state_rel_l2": state_err},
        "roofline": r32["roofline"],
        "passes": r32["passes"],
        "step_bytes_GB": r32["step_bytes_GB"],
        "step_GBps": r32["step_GBps"],
        "integrators": integ,
        "increment_stats": incr,
        "ensemble_96_boxes": ens,
        "closure_les": clos,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
