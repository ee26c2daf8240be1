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


@pytest.mark.parametrize('N', [5000, 100_000, 300_000])
def test_cgs2_commit_column_against_numpy(device, N):
    """One CGS2 column (hex_cgs2_commit: both passes, the Hessenberg column and
    the next basis vector) against NumPy's CGS2 in the reference's order
    (python/propagate.py:17-25), for every basis size up to 16, on states of
    5 to 98 reduction blocks and one at the 293-block size."""
    import ctypes as ct
    import torch
    from laqphysics_amd import native
    from laqphysics_amd.hip_backend import _stream, _vp
    print(f'CASE cgs2 column N={N}', flush=True)
    h = golden_problem()
    op = HipHamiltonian(h, device)
    rng = np.random.default_rng(N)
    lib = native.lib()
    for k in range(1, 17):
        Qh = rng.normal(size=(k + 1, N)) + 1j * rng.normal(size=(k + 1, N))
        Qh[:k] = np.linalg.qr(Qh[:k].T)[0].T
        vh = rng.normal(size=N) + 1j * rng.normal(size=N) + Qh[:k].T @ (rng.normal(size=k) * 10)
        Q = torch.tensor(Qh, device=device)
        v = torch.tensor(vh, device=device)
        hd = torch.zeros(64, dtype=torch.complex128, device=device)
        native.check(lib.hex_cgs2_commit(op._engine, _vp(Q), N, k, _vp(v), N, _vp(hd), _vp(Q[k]), _stream(device)),
                     'cgs2')
        w = vh.copy()
        z1 = Qh[:k].conj() @ w
        w = w - Qh[:k].T @ z1
        z2 = Qh[:k].conj() @ w
        w = w - Qh[:k].T @ z2
        nrm = np.linalg.norm(w)
        got = hd.cpu().numpy()
        np.testing.assert_allclose(got[:k], z1 + z2, rtol=0, atol=1e-12 * np.linalg.norm(vh))
        assert abs(got[k].real - nrm) <= 1e-12 * nrm and got[k].imag == 0
        np.testing.assert_allclose(Q[k].cpu().numpy(), w / nrm, rtol=0, atol=1e-12 / np.sqrt(N))
        np.testing.assert_allclose(v.cpu().numpy(), w, rtol=0, atol=1e-12 * nrm / np.sqrt(N) * 10)
