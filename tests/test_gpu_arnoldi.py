"""Device Arnoldi propagation against the oracle and the golden traces (gpu)."""
import numpy as np
import pytest

import oracle
from oracle import arnoldi as oracle_arnoldi
from laqphysics_amd import hip_backend, workloads
from laqphysics_amd.hip_backend import HipHamiltonian
from tests.helpers import golden_field, golden_problem

pytestmark = pytest.mark.gpu


def _say(*what):
    print('CASE', *what, flush=True)


@pytest.mark.parametrize('i', range(6))
def test_device_arnoldi_reproduces_golden_step(golden_dir, device, i):
    _say('golden arnoldi', i)
    d = np.load(golden_dir / 'debug_modules' / f'3d_step{i:04d}.npz')
    h = golden_problem()
    op = HipHamiltonian(h, device)
    A = golden_field(i, float(d['dt']))
    y, info = hip_backend.exponential_action(lambda v: op.apply(v, A, True), op.state(d['input']),
                                             float(d['dt']), tol=1e-12, maxdim=60)
    assert info['dimension'] == 12
    np.testing.assert_allclose(op.host(y), d['output'], atol=1e-13)


def test_device_propagation_matches_oracle_plain_l2_n85(device):
    """40 midpoint Arnoldi steps of the 85-point, 11-channel problem from the
    synthetic initial state: ||d psi|| <= 1e-10 and identical Krylov choices."""
    _say('plain_l2_n85 propagation')
    w = workloads.plain_l2_n85()
    h = w.h
    avec = w.avec
    t0 = 300 * w.dt  # inside the pulse, so the dipole terms are active
    op = HipHamiltonian(h, device)
    x = workloads.initial_state(h)
    state = op.state(x)
    for i in range(40):
        t = t0 + i * w.dt
        x, a = oracle_arnoldi.step(lambda s, y: oracle.tensor_apply(h, y, avec(s), True), x, t, w.dt,
                                   tol=w.tolerance, maxdim=w.krylov_dimension)
        state, b = hip_backend.step(lambda s, y: op.apply(y, avec(s), True), state, t, w.dt,
                                    tol=w.tolerance, maxdim=w.krylov_dimension)
        assert a['dimension'] == b['dimension']
    assert np.linalg.norm(op.host(state) - x) <= 1e-10
