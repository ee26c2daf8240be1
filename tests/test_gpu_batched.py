"""Batched graph-captured propagation against the plain device engine and the
oracle (gpu)."""
import numpy as np
import pytest

import oracle
from oracle import arnoldi as oracle_arnoldi
from laqphysics_amd import hip_backend, workloads
from laqphysics_amd.batched import BatchedPropagator
from laqphysics_amd.hip_backend import HipHamiltonian
from laqphysics_amd.pulses import Pulse, vector_function

pytestmark = pytest.mark.gpu


def _say(*what):
    print('CASE', *what, flush=True)


def _members():
    return [[(Pulse(omega=1.6, cycles=6, field=.02, start=s, cep=c), 'z')]
            for s, c in ((0., 0.), (1., .7), (2.5, 2.1))]


def test_batched_members_match_plain_engine_and_oracle(device):
    """Three members with different pulses (delay, CEP) and uneven step counts:
    every member equals its own plain-engine and oracle runs."""
    _say('batched vs plain')
    w = workloads.plain_l2_n85()
    h = w.h
    op = HipHamiltonian(h, device)
    members = _members()
    avecs = [vector_function(m) for m in members]
    t0 = [3.0, 3.2, 4.0]
    counts = [12, 7, 10]
    x0 = workloads.initial_state(h)
    batch = BatchedPropagator(op, [x0, x0, x0], members, t0, counts, w.dt, tol=w.tolerance,
                              maxdim=w.krylov_dimension, tune=False)
    batch.run()
    status = batch.status()
    assert list(status['steps']) == counts and not status['alive'].any()
    for b in range(3):
        x = x0.copy()
        state = op.state(x0)
        for k in range(counts[b]):
            t = t0[b] + k * w.dt
            state, info = hip_backend.step(lambda s, y: op.apply(y, avecs[b](s), True), state, t, w.dt,
                                           tol=w.tolerance, maxdim=w.krylov_dimension)
            x, _ = oracle_arnoldi.step(lambda s, y: oracle.tensor_apply(h, y, avecs[b](s), True), x, t, w.dt,
                                       tol=w.tolerance, maxdim=w.krylov_dimension)
        got = batch.psi[b].cpu().numpy()
        assert np.linalg.norm(got - op.host(state)) <= 1e-12
        assert np.linalg.norm(got - x) <= 1e-10


def test_batched_subdivision_follows_the_reference_recursion(device):
    """A step too long for the Krylov order is halved recursively (tol/2 per
    level) on the device, as python/torch_backend.py:160-167 does on the host."""
    _say('batched subdivision')
    w = workloads.plain_l2_n85()
    h = w.h
    op = HipHamiltonian(h, device)
    members = _members()[:2]
    avecs = [vector_function(m) for m in members]
    dt, maxdim = .02, 10
    x0 = workloads.initial_state(h)
    batch = BatchedPropagator(op, [x0, x0], members, [3., 3.5], [2, 1], dt, tol=1e-10, maxdim=maxdim, tune=False)
    batch.run()
    status = batch.status()
    assert list(status['steps']) == [2, 1]
    assert np.all(status['substeps'] > status['steps'])
    for b in range(2):
        x = x0.copy()
        for k in range(int(status['steps'][b])):
            t = [3., 3.5][b] + k * dt
            x, info = oracle_arnoldi.step(lambda s, y: oracle.tensor_apply(h, y, avecs[b](s), True), x, t, dt,
                                          tol=1e-10, maxdim=maxdim)
        assert np.linalg.norm(batch.psi[b].cpu().numpy() - x) <= 1e-10


def test_batched_replays_are_reproducible(device):
    _say('batched replay determinism')
    w = workloads.plain_l2_n85()
    h = w.h
    op = HipHamiltonian(h, device)
    m = _members()[0]
    x0 = workloads.initial_state(h)
    runs = []
    for _ in range(2):
        batch = BatchedPropagator(op, [x0, x0], [m, m], [3., 3.], [5, 5], w.dt, tol=w.tolerance,
                                  maxdim=w.krylov_dimension, tune=False)
        batch.run(chunk=1)
        runs.append(batch.psi.cpu().numpy())
    np.testing.assert_array_equal(runs[0], runs[1])
    np.testing.assert_array_equal(runs[0][0], runs[0][1])
