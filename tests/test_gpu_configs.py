"""Parity at the benchmark configurations themselves (gpu).

* production_l6_L4_n441 (coupled l <= 6, L <= 4, M = 0; n = 441, 72 channels):
  full-size H.psi in both gauges, CF4-Pade steps (GMRES restart 30, tol 1e-12)
  from the device ground state, and tSURFF frames at r = 80 (flux, channel
  projection, Volkov accumulation), each against the oracle;
* batched_l4_L3_n441 (coupled l <= 4, L <= 3; n = 441, 33 channels): a B = 16
  pump-probe scan with uneven per-member step counts, every member against the
  plain engine, two members against the oracle;
* the adjoint benchmark: 512 piecewise-linear field samples over 1000 a.u.
  (seeded N(0, 0.01^2), sin^2 window) on the 85-point grid, J = He+(n = 2)
  tSURFF yield + ionisation probability, checkpoints every 64 steps, over a
  6400-step window, against differences along 8 seeded directions.
"""
import os
import time

import numpy as np
import pytest
import torch
from threadpoolctl import threadpool_limits

import oracle
from oracle import arnoldi as oracle_arnoldi
from oracle import implicit as oi
from oracle import surface as osurf
from laqphysics_amd import hip_backend, implicit, workloads
from laqphysics_amd.hip_backend import HipHamiltonian
from tests.helpers import random_state, relative_error

pytestmark = pytest.mark.gpu

HOST_THREADS = 16   # the host cores one GPU's share of the box provides


def _say(*what):
    print('CASE', *what, flush=True)


@pytest.fixture(scope='module')
def production(device):
    w = workloads.production_l6_L4_n441()
    op = HipHamiltonian(w.h, device)
    E, psi0 = workloads.ground_state(w.h, device)
    return w, op, E, psi0


@pytest.mark.parametrize('velocity', [True, False])
def test_production_apply_against_oracle(production, velocity):
    w, op, E, psi0 = production
    _say('production_l6_L4_n441 H.psi, velocity', velocity)
    h = w.h
    field = (0., 0., 0.05)
    for seed in (49, 386):
        x = random_state(h.size, seed)
        y = op.host(op.apply(op.state(x), field, velocity))
        assert relative_error(y, oracle.tensor_apply(h, x, field, velocity)) <= 1e-12


def test_production_cf4_steps_against_oracle(production):
    """Three CF4-Pade steps (GMRES restart 30, tol 1e-12, the ground energy as
    reference energy) from the ground state at mid-pump: identical GMRES
    counts, |dpsi| <= 1e-10."""
    w, op, E, psi0 = production
    h = w.h
    _say('production_l6_L4_n441 CF4 x3 vs oracle')
    restart = w.extra['gmres_restart']
    op.separable = implicit.SeparablePreconditioner(op)
    eng = oi.OracleEngine(h)
    with threadpool_limits(HOST_THREADS, user_api='blas'):
        eng.separable = oi.SeparablePreconditioner(eng)
        x = psi0.copy()
        state = op.state(psi0)
        t0 = 8 * np.pi / 1.5 / 2        # the middle of the pump
        for i in range(3):
            t = t0 + i * w.dt
            start = time.perf_counter()
            x, a = oi.cf4_step(eng, x, w.avec, t, w.dt, E, tol=w.tolerance, restart=restart)
            host_s = time.perf_counter() - start
            state, b = implicit.cf4_step(op, state, w.avec, t, w.dt, E, tol=w.tolerance, restart=restart)
            print(f'step {i}: GMRES {a["linear_iterations"]} (oracle) / {b["linear_iterations"]} (device), '
                  f'oracle step {host_s:.1f} s', flush=True)
            assert a['linear_iterations'] == b['linear_iterations']
    err = np.linalg.norm(op.host(state) - x)
    print('|dpsi| after 3 steps', err, flush=True)
    assert err <= 1e-10


def test_production_surface_frames_against_oracle(production):
    """tSURFF at r = 80: two frames -- velocity-gauge flux on the device,
    projection on the He+ n <= 2 duals (sparse-LU CF4 between the frame
    times), Volkov accumulation -- against the oracle's flux, contraction and
    accumulation of the same two states."""
    from laqphysics_amd.ionic import Ionic, backward_duals
    from laqphysics_amd.surface import SurfaceContractor, VolkovAccumulator, directions
    w, op, E, psi0 = production
    h = w.h
    base = h.uncoupled
    _say('production_l6_L4_n441 tSURFF frames at r = 80')
    h.prepare_surface(w.surface)
    base.prepare_surface(w.surface)
    op2 = HipHamiltonian(h, op.device)     # an engine built after prepare_surface carries the surface rows
    labels = [(1, 0, 0), (2, 0, 0), (2, 1, 0)]
    unique = list(dict.fromkeys((n, l, abs(m)) for n, l, m in labels))
    columns = [unique.index((n, l, abs(m))) for n, l, m in labels]
    ion = Ionic(base, {m for n, l, m in unique})
    t0 = 60.
    times = t0 + np.arange(2) * w.dt
    chis = backward_duals(ion, unique, times, w.avec)
    energy = np.linspace(.2, 2., 7)
    angles = directions(True, 5, 1)
    con = SurfaceContractor(base, ion, labels, unique, columns, True, h.transform, op.device)
    acc = VolkovAccumulator(con, energy, angles, w.avec)
    rcon = osurf.Contractor(base, ion, labels, unique, columns, True, h.transform)
    racc = osurf.Volkov(base, rcon.outerstates, len(labels), energy, angles[0], angles[1], w.avec)
    # states with flux through r = 80: the ground state plus a small
    # everywhere-nonzero part, different at the two frames
    for i, t in enumerate(times):
        x = psi0 + 1e-3 * random_state(h.size, 11 + i)
        state = op2.state(x)
        frame = op2.flux_device(state, w.avec(t))
        rframe = osurf.flux(h, x, w.avec(t))
        assert relative_error(op2.host(frame.permute(2, 1, 0)), rframe) <= 1e-12
        acc.add(i, t, con(chis[i], frame))
        racc.add(i, t, rcon(chis[i], rframe))
        print(f'frame {i} at t = {t}: flux and projection done', flush=True)
    scale = np.max(abs(racc.amplitude))
    assert scale > 0
    assert np.max(abs(acc.host_amplitude() - racc.amplitude)) <= 1e-10 * scale


def _scan(B=16, seed=2024):
    """The pump-probe scan: delays U[0, 400] a.u. after the pump and probe CEPs
    U[0, 2 pi) from one seed; full step counts 2500 + 2500 delay/400."""
    from laqphysics_amd.pulses import Pulse
    rng = np.random.default_rng(seed)
    delays = rng.uniform(0., 400., B)
    ceps = rng.uniform(0., 2 * np.pi, B)
    pump = Pulse(omega=1.5, cycles=8, field=.03)
    members = [[(pump, 'z'), (Pulse(omega=.114, cycles=4, field=.01, start=pump.duration + d, cep=c), 'z')]
               for d, c in zip(delays, ceps)]
    counts = [int(2500 + round(2500 * d / 400.)) for d in delays]
    return members, counts, delays, pump


def test_batched_scan_B16_uneven_against_plain_and_oracle(device):
    """B = 16 members of the scan, each started inside its own probe pulse
    (different field and phase per member) with a prefix of 2-5 steps that
    keeps the scan's uneven counts: every member equals the plain engine,
    the two shortest the oracle, at 1e-10."""
    from laqphysics_amd.batched import BatchedPropagator
    from laqphysics_amd.pulses import vector_function
    w = workloads.batched_l4_L3_n441()
    h = w.h
    _say('batched_l4_L3_n441 B=16 uneven scan')
    op = HipHamiltonian(h, device)
    members, full, delays, pump = _scan()
    counts = [2 + (c - 2500) * 4 // 2501 for c in full]          # 2 .. 5, in the order of the full counts
    assert min(counts) == 2 and max(counts) >= 4
    t0 = [pump.duration + d + 10. for d in delays]
    E, psi0 = workloads.ground_state(h, device)
    batch = BatchedPropagator(op, [psi0] * 16, members, t0, counts, w.dt, tol=w.tolerance,
                              maxdim=w.krylov_dimension, tune=True)
    start = time.perf_counter()
    batch.run()
    print(f'batched run {time.perf_counter() - start:.1f} s, attempts per member '
          f'{batch.status()["attempts"].tolist()}', flush=True)
    status = batch.status()
    assert list(status['steps']) == counts and not status['alive'].any()
    shortest = sorted(range(16), key=lambda b: (counts[b], b))[:2]
    for b in range(16):
        avec = vector_function(members[b])
        state = op.state(psi0)
        x = psi0.copy()
        for k in range(counts[b]):
            t = t0[b] + k * w.dt
            state, _ = hip_backend.step(lambda s, y: op.apply(y, avec(s), True), state, t, w.dt,
                                        tol=w.tolerance, maxdim=w.krylov_dimension)
            if b in shortest:
                start = time.perf_counter()
                x, _ = oracle_arnoldi.step(lambda s, y: oracle.tensor_apply(h, y, avec(s), True), x, t, w.dt,
                                           tol=w.tolerance, maxdim=w.krylov_dimension)
                print(f'member {b} step {k}: oracle {time.perf_counter() - start:.1f} s', flush=True)
        got = batch.psi[b].cpu().numpy()
        assert np.linalg.norm(got - op.host(state)) <= 1e-10, b
        if b in shortest:
            err = np.linalg.norm(got - x)
            print(f'member {b}: {counts[b]} steps, |batched - oracle| = {err:.2e}', flush=True)
            assert err <= 1e-10, b


def _five_point(J, theta, v, h):
    return (8 * (J(theta + h * v) - J(theta - h * v)) - (J(theta + 2 * h * v) - J(theta - 2 * h * v))) / (12 * h)


def test_adjoint_benchmark_setup_against_differences(device):
    """The adjoint benchmark's own setup over a 6400-step window (t = 468 ..
    532, the middle of the window where the sin^2 weight is largest): the
    gradient against central differences h = 1e-6 and five-point differences
    h = 1e-4 along seeded directions (2 and 1 of them; all 8 and 2 with
    HEX_LONG_TESTS set: every J costs a 6400-step forward run and the duals).

    J = P_ion + Y carries ~1e-14 of rounding (measured: the central stencil's
    deviations reach 1.2e-8 at h = 1e-6 while the five-point ones stay below
    5e-10), so the two-point stencil at h = 1e-6 has a floor of
    ~1e-14/1e-6 = 1e-8 absolute and the five-point one at h = 1e-4 of
    ~(18/12) 1e-14/1e-4 = 1.5e-10.  Both hold the gradient to rtol 1e-6 plus
    three times that floor; the measured errors are printed."""
    from laqphysics_amd.adjoint import PiecewiseLinearControl, Problem
    w = workloads.plain_l2_n85()
    h = w.h
    _say('adjoint benchmark setup: 512 samples, 6400 steps, 8 directions')
    op = HipHamiltonian(h, device)
    E, psi0 = workloads.ground_state(h, device)
    control = PiecewiseLinearControl(np.random.default_rng(7).normal(0., .01, 512), 1000.)
    p = Problem(op, psi0, control, w.dt, 6400, t0=468., tol=w.tolerance, maxdim=w.krylov_dimension,
                checkpoint_every=64, radius=w.surface)
    p.add_yield(w.surface, labels=[(2, 0, 0), (2, 1, 0)])
    start = time.perf_counter()
    g, J0 = p.gradient()
    print(f'J = {J0:.6e}, |g| = {np.linalg.norm(g):.3e}, gradient {time.perf_counter() - start:.1f} s', flush=True)
    assert np.isfinite(g).all() and J0 > 0
    # A(t) = -int_0^t E: every sample before the window's end shifts A inside
    # it, the samples after it do not enter
    assert np.all(g[control.tk > 532. + 2.] == 0) and np.any(g[control.tk < 468.] != 0)
    inside = control.tk < 532. + 2.
    theta = control.theta.copy()
    p.forced_dims = list(p.dims)

    def J(th):
        p.control.theta = np.asarray(th, float)
        return p.objective(p.forward(keep=False)[0])

    rng = np.random.default_rng(8)
    worst2, worst5 = 0., 0.
    central, fivepoint = (8, 2) if os.environ.get('HEX_LONG_TESTS') else (2, 1)
    for d in range(central):
        v = rng.normal(size=theta.size) * inside
        v /= np.linalg.norm(v)
        fd2 = (J(theta + 1e-6 * v) - J(theta - 1e-6 * v)) / 2e-6
        err2 = abs(g @ v - fd2)
        worst2 = max(worst2, err2 / abs(fd2))
        print(f'direction {d}: g.v = {g @ v:.9e}, central h=1e-6 {fd2:.9e} (|diff| {err2:.2e})', flush=True)
        assert err2 <= 1e-6 * abs(fd2) + 3e-8
        if d < fivepoint:
            fd5 = _five_point(J, theta, v, 1e-4)
            err5 = abs(g @ v - fd5)
            worst5 = max(worst5, err5 / abs(fd5))
            print(f'direction {d}: five-point h=1e-4 {fd5:.9e} (|diff| {err5:.2e})', flush=True)
            assert err5 <= 1e-6 * abs(fd5) + 5e-10
    print(f'worst relative deviation: central {worst2:.2e}, five-point {worst5:.2e}', flush=True)
    p.forced_dims = None
    p.control.theta = theta
