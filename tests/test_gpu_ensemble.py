"""Ensembles (batched boxes, one HIP graph per step) and the FFT schedule autotuner, through the C ABI.

Bars:
  * every FFT schedule of a size meets the single-box parity bars against the FP64 oracle
    (FP64 rel L2 <= 1e-12; FP32-bulk rhs rel L2 <= 2e-6);
  * the persistent z-pass grid of a launch schedule changes no bit of a step;
  * a box stepped inside a batch / graph is BITWISE equal to the same box stepped alone under the
    same schedule (same kernels, per-box data, no cross-box reductions);
  * graph replays interleaved with eager work (measures, state reads, time reads) match the
    per-box direct launches bitwise on every replay;
  * the autotuner leaves state and time untouched.
"""
import numpy as np
import pytest

from conftest import rel

pytestmark = pytest.mark.gpu

ALT_NS = [8, 12, 16, 24, 32, 48, 64, 96, 128, 256]


def k4(oracle, N, seed, k0=2.0):
    return oracle.ic_k4(N, k0, 0.5, seed)


@pytest.mark.parametrize("N", ALT_NS)
@pytest.mark.parametrize("sched", [1, 2])
def test_schedule_rhs_parity(nsg, oracle, N, sched):
    assert nsg.schedule_count(N) == 9
    uh = k4(oracle, N, 77 + N)
    want = oracle.rhs(uh)
    with nsg.Solver(N) as s:
        s.set_schedule(sched)
        assert rel(s.rhs(uh), want) < 1e-12
    with nsg.Solver(N, precision="fp32bulk") as s:
        s.set_schedule(sched)
        assert rel(s.rhs(uh), want) < 2e-6


@pytest.mark.parametrize("N", [12, 16, 32, 48, 128])
@pytest.mark.parametrize("sched", [1, 2])
def test_schedule_steps_parity(nsg, oracle, N, sched):
    uh = k4(oracle, N, 5 + N)
    with nsg.Solver(N, nu=0.01, dt=2e-3, force_eps=0.1) as s:
        s.set_schedule(sched)
        s.set_state(uh)
        s.step(2)
        got = s.get_state()
    want = oracle.step(uh, 0.01, 2e-3, 1, "rk4cn", 0.1, 2.5, nsteps=2)
    assert rel(got - uh, want - uh) < 1e-12


@pytest.mark.parametrize("sched", [1, 2])
def test_schedule_steps_256(nsg, oracle, sched):
    """the 256^3 FFT schedules 1 (4 elements per thread) and 2 (wider column tiles) in the stage
    passes: two forced RK4 steps in FP64 agree with schedule 0's (itself at 1e-12 of the oracle's
    right-hand side at 256^3) to 1e-12 on the increment, and FP32-bulk under the same schedule
    agrees with that FP64 state to 1e-8, the one-step FP32-bulk bar against the oracle"""
    N = 256
    uh = k4(oracle, N, 21)
    kw = dict(nu=1e-3, dt=1e-3, force_eps=0.1)
    out = {}
    for s_, prec in ((0, "fp64"), (sched, "fp64"), (sched, "fp32bulk")):
        with nsg.Solver(N, precision=prec, **kw) as s:
            s.set_schedule(s_)
            s.set_state(uh)
            s.step(2 if prec == "fp64" else 1)
            out[s_, prec] = s.get_state()
    want = out[0, "fp64"]
    assert rel(out[sched, "fp64"] - uh, want - uh) < 1e-12
    with nsg.Solver(N, **kw) as s:
        s.set_state(uh)
        s.step(1)
        one = s.get_state()
    assert rel(out[sched, "fp32bulk"], one) < 1e-8


def test_schedule_out_of_range(nsg):
    with nsg.Solver(16) as s:
        with pytest.raises(nsg.NSGError):
            s.set_schedule(9)
        with pytest.raises(nsg.NSGError):
            s.set_schedule(-1)
    assert nsg.schedule_count(16) == 9 and nsg.schedule_count(512) == 3 and nsg.schedule_count(17) == 0


@pytest.mark.parametrize("N,prec", [(256, "fp32bulk"), (128, "fp64")])
def test_z_grid_schedules_bitwise(nsg, oracle, N, prec):
    """launch schedules F and 2F (F = FFT schedules of N) differ from schedule 0 only in the
    persistent z-pass grid (N = 256 FP32: 4 and 12 blocks per CU instead of 8, 8192 blocks of work
    on grids of 2048, 1024 and 3072 blocks; N = 128 FP64: 8 and 12 instead of 4, 2048 blocks of
    work on grids of 1024, 2048 and 2048), which must change no bit"""
    nf = nsg.schedule_count(N) // 3
    assert nf == 3
    uh = k4(oracle, N, 11)
    got = {}
    for sched in (0, nf, 2 * nf):
        with nsg.Solver(N, nu=1e-3, dt=1e-3, cfl=0.5, precision=prec) as s:
            s.set_schedule(sched)
            s.set_state(uh)
            s.step(2)
            got[sched] = (s.get_state(), s.time())
    for sched in (nf, 2 * nf):
        assert np.array_equal(got[sched][0], got[0][0]), f"schedule {sched}"
        assert got[sched][1] == got[0][1]


def mixed_boxes():
    """small uneven ensemble: two batch groups of FP32-bulk (32^3, 48^3), one FP64 32^3 group, one
    forced IF-RK4 group; per-box nu, dt and CFL differ"""
    boxes = []
    for b in range(3):
        boxes.append(dict(N=32, nu=1e-3 + 4e-3 * b / 10, dt=1e-3, precision="fp32bulk", cfl=0.5))
    for b in range(2):
        boxes.append(dict(N=48, nu=2e-3 + 1e-3 * b, dt=5e-4, precision="fp32bulk"))
    boxes.append(dict(N=32, nu=5e-3, dt=1e-3, precision="fp64", cfl=0.3))
    boxes.append(dict(N=32, nu=4e-3, dt=1e-3, precision="fp32bulk", cfl=0.5))  # joins group 0
    boxes.append(dict(N=24, nu=1e-2, dt=2e-3, scheme="ifrk4", force_eps=0.1, force_kf=2.5))
    return boxes


def init_mixed(ens, boxes):
    for b, p in enumerate(boxes):
        ens.init(b, "k4", k0=2.0 + (b % 3), energy=0.5, seed=1000 + b)


def test_ensemble_grouping(nsg):
    boxes = mixed_boxes()
    with nsg.Ensemble(boxes) as e:
        assert len(e) == len(boxes)
        assert e.groups == 4
        assert e.group_of(0) == (0, 0) and e.group_of(2) == (0, 2) and e.group_of(6) == (0, 3)
        assert e.group_of(3) == (1, 0) and e.group_of(5) == (2, 0) and e.group_of(7) == (3, 0)


def test_ensemble_graph_equals_single_box(nsg):
    """each box after 3 graph steps == the same box stepped alone (bitwise), including adaptive dt"""
    boxes = mixed_boxes()
    with nsg.Ensemble(boxes) as e:
        init_mixed(e, boxes)
        e.step(3)
        got = [e.get_state(b) for b in range(len(boxes))]
        times = [e.time(b) for b in range(len(boxes))]
    for b, p in enumerate(boxes):
        with nsg.Solver(**p) as s:
            s.init("k4", k0=2.0 + (b % 3), energy=0.5, seed=1000 + b)
            s.step(3)
            assert np.array_equal(got[b], s.get_state()), f"box {b}"
            assert times[b] == s.time(), f"box {b}"


def test_ensemble_graph_replays_vs_serial(nsg):
    """graph replays with eager work in between vs per-box direct launches: bitwise every replay"""
    boxes = mixed_boxes()
    with nsg.Ensemble(boxes) as g, nsg.Ensemble(boxes) as r:
        init_mixed(g, boxes)
        init_mixed(r, boxes)
        for rep in range(4):
            g.step(1 + rep % 2)
            r.step_serial(1 + rep % 2)
            for b in range(len(boxes)):
                tg, ekg, _ = g.measures(b)
                tr, ekr, _ = r.measures(b)
                assert tuple(tg) == tuple(tr) and np.array_equal(ekg, ekr), f"replay {rep} box {b}"
                assert g.time(b) == r.time(b)
            for b in (0, 3, 5, 7):
                assert np.array_equal(g.get_state(b), r.get_state(b)), f"replay {rep} box {b}"


@pytest.mark.parametrize("N", [12, 24, 48])
def test_batch_cfl_per_box(nsg, N):
    """adaptive dt is per box: three boxes of one batch with different amplitudes get different dt,
    each equal (bitwise) to the box run alone.  At N=12 a wavefront of the z pass holds lines of two
    boxes, so its max|u| reduction must split them."""
    boxes = [dict(N=N, nu=0.01, dt=1.0, cfl=0.5) for _ in range(3)]
    energies = [0.5, 2.0, 0.125]
    with nsg.Ensemble(boxes) as e:
        assert e.groups == 1
        for b in range(3):
            e.init(b, "k4", k0=2.0, energy=energies[b], seed=40 + b)
        e.step(2)
        got = [(e.get_state(b), e.time(b)) for b in range(3)]
    dts = [t[1][1] for t in got]
    assert len(set(dts)) == 3
    for b in range(3):
        with nsg.Solver(**boxes[b]) as s:
            s.init("k4", k0=2.0, energy=energies[b], seed=40 + b)
            s.step(2)
            assert np.array_equal(s.get_state(), got[b][0]) and s.time() == got[b][1], f"box {b}"


def test_ensemble_against_oracle(nsg, oracle):
    """two boxes of one FP64 batch, fixed dt, against the FP64 oracle"""
    boxes = [dict(N=24, nu=0.01, dt=2e-3), dict(N=24, nu=0.03, dt=1e-3, visc_pow=1)]
    with nsg.Ensemble(boxes) as e:
        u0 = [k4(oracle, 24, 11), k4(oracle, 24, 12, k0=3.0)]
        for b in range(2):
            e.set_state(b, u0[b])
        e.step(2)
        for b, p in enumerate(boxes):
            want = oracle.step(u0[b], p["nu"], p["dt"], nsteps=2)
            assert rel(e.get_state(b) - u0[b], want - u0[b]) < 1e-12
            t, dt, _ = e.time(b)
            assert abs(t - 2 * p["dt"]) < 1e-15 and dt == p["dt"]
            to, eko, _ = oracle.measures(want, p["nu"])
            tot, ek, _ = e.measures(b)
            assert np.allclose(np.array(tot)[[0, 1, 2, 4]], to[[0, 1, 2, 4]], rtol=1e-12)
            assert rel(e.realspace(b, 0), oracle.realspace(want, 0)) < 1e-12


def test_autotune_restores_state(nsg, oracle):
    with nsg.Solver(48, nu=1e-3, dt=1e-3, precision="fp32bulk", cfl=0.5) as s:
        s.init("k4", k0=2.0, energy=0.5, seed=3)
        s.step(1)
        before, tb = s.get_state(), s.time()
        chosen, times = s.autotune(iters=2, use_cache=False)
        assert 0 <= chosen < 9 and times is not None and len(times) == 9 and min(times) > 0
        assert np.array_equal(s.get_state(), before) and s.time() == tb
        s.step(2)
        got = s.get_state()
    with nsg.Solver(48, nu=1e-3, dt=1e-3, precision="fp32bulk", cfl=0.5) as s:
        s.init("k4", k0=2.0, energy=0.5, seed=3)
        s.step(1)  # schedule 0, as before the tuning above
        s.set_schedule(chosen)
        s.step(2)
        assert np.array_equal(s.get_state(), got)


def test_ensemble_autotune_then_graph(nsg):
    boxes = mixed_boxes()
    with nsg.Ensemble(boxes) as e, nsg.Ensemble(boxes) as r:
        init_mixed(e, boxes)
        init_mixed(r, boxes)
        chosen = e.autotune(iters=1, use_cache=False)
        assert len(chosen) == e.groups
        for g, c in enumerate(chosen):
            r.set_schedule(g, c)
        e.step(2)
        r.step_serial(2)
        for b in range(len(boxes)):
            assert np.array_equal(e.get_state(b), r.get_state(b)), f"box {b}"


def test_ensemble_errors(nsg):
    with pytest.raises(nsg.NSGError, match="box 1"):
        nsg.Ensemble([dict(N=32), dict(N=33)])
    boxes = [dict(N=16, nu=0.01, dt=1e-3), dict(N=16, nu=0.01, dt=1e-3)]
    with nsg.Ensemble(boxes) as e:
        e.init(0, "k4", k0=2.0, energy=0.5, seed=1)
        bad = np.zeros((16, 16, 9, 3), dtype=np.complex128)
        bad[1, 0, 0, 1] = np.nan
        e.set_state(1, bad)
        with pytest.raises(nsg.NSGError, match="box 1") as ei:
            e.step(1)
        assert ei.value.code == -4
        with pytest.raises(nsg.NSGError):
            e.time(2)


def config3_boxes():
    """96 uneven boxes: 32 at 32^3, 32 at 48^3, 24 at 64^3, 8 at 96^3; box b: seed 1000+b,
    k0 = 2 + (b mod 3), nu_b = 1e-3 + 4e-3 b/95, FP32-bulk"""
    sizes = [32] * 32 + [48] * 32 + [64] * 24 + [96] * 8
    return [dict(N=N, nu=1e-3 + 4e-3 * b / 95, dt=1e-3, precision="fp32bulk", cfl=0.5) for b, N in enumerate(sizes)]


def test_config3_ensemble_sample(nsg):
    """the 96-box ensemble: 2 graph steps; boxes 0, 40, 70, 95 bitwise equal to a lone run; every box
    finite with energy decaying (unforced)"""
    boxes = config3_boxes()
    with nsg.Ensemble(boxes) as e:
        assert e.groups == 4
        e0 = []
        for b in range(96):
            e.init(b, "k4", k0=2.0 + b % 3, energy=0.5, seed=1000 + b)
            e0.append(e.measures(b)[0].energy)
        e.step(2)
        for b in range(96):
            tot = e.measures(b)[0]
            assert np.isfinite(tot.energy) and tot.energy < e0[b]
        got = {b: e.get_state(b) for b in (0, 40, 70, 95)}
    for b, st in got.items():
        with nsg.Solver(**boxes[b]) as s:
            s.init("k4", k0=2.0 + b % 3, energy=0.5, seed=1000 + b)
            s.step(2)
            assert np.array_equal(s.get_state(), st), f"box {b}"
