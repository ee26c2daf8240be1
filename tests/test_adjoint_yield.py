"""The oracle's tSURFF yield term of the adjoint objective (oracle/adjoint_yield.py,
PyTorch on the CPU, gradient by autograd), pinned: its ionic duals against the
sparse-LU backward propagation, its value against the oracle's tSURFF pipeline,
its explicit field derivative against finite differences (CPU).  The device
yield is checked against it in tests/test_gpu_yield.py."""
import types

import numpy as np
import pytest
import torch

from oracle import surface as osurf
from oracle.adjoint_yield import TorchYield
from laqphysics_amd import workloads
from laqphysics_amd.adjoint import PiecewiseLinearControl, SurfaceYield
from laqphysics_amd.ionic import backward_duals

LABELS = [(2, 0, 0), (2, 1, 0)]


@pytest.fixture(scope='module')
def setup():
    w = workloads.plain_l2_n85()
    h = w.h
    rng = np.random.default_rng(4)
    nsteps, dt = 12, .05
    control = PiecewiseLinearControl(rng.normal(0., .5, 6), nsteps * dt)
    problem = types.SimpleNamespace(h=h, engine=types.SimpleNamespace(device=torch.device('cpu')),
                                    control=control, nsteps=nsteps, dt=dt, t0=0.)
    energy = np.linspace(.3, 1.2, 5)
    from laqphysics_amd.surface import directions
    angles = directions(True, 4, 1)
    sy = TorchYield(SurfaceYield(problem, w.surface, LABELS, energy, angles))
    shape = (nsteps + 1, h.nc, len(h.surface_indices), h.n)
    F0 = torch.as_tensor(rng.normal(size=shape) + 1j * rng.normal(size=shape)) * 1e-3
    Fd = torch.as_tensor(rng.normal(size=shape) + 1j * rng.normal(size=shape)) * 1e-3
    return w, h, control, sy, energy, angles, F0, Fd


def _oracle_yield(h, sy, control, energy, angles, F0, Fd):
    base = h.uncoupled
    unique = list(dict.fromkeys((n, l, abs(m)) for n, l, m in LABELS))
    columns = [unique.index((n, l, abs(m))) for n, l, m in LABELS]
    chis = backward_duals(sy.sy.ion, unique, sy.sy.times, control.avec)
    con = osurf.Contractor(base, sy.sy.ion, LABELS, unique, columns, True, h.transform)
    vol = osurf.Volkov(base, con.outerstates, len(LABELS), energy, angles[0], angles[1], control.avec)
    for i, t in enumerate(sy.sy.times):
        frame = (F0[i] + control.vector(t) * Fd[i]).numpy().transpose(2, 1, 0)
        vol.add(i, t, con(chis[i], frame))
    wE = np.gradient(energy)
    wE[0] = wE[-1] = (energy[1] - energy[0]) / 2
    weight = wE[:, None] * np.sqrt(2 * energy)[:, None] * angles[2][None, :]
    return float(np.sum(weight[None] * abs(vol.amplitude) ** 2)), chis


def test_duals_match_the_sparse_lu_backward_propagation(setup):
    w, h, control, sy, energy, angles, F0, Fd = setup
    theta = torch.as_tensor(control.theta)
    unique = list(dict.fromkeys((n, l, abs(m)) for n, l, m in LABELS))
    ref = backward_duals(sy.sy.ion, unique, sy.sy.times, control.avec)
    got = sy.duals(theta)
    for a, b in zip(got, ref):
        np.testing.assert_allclose(a.numpy(), b, rtol=1e-11, atol=1e-13)


def test_yield_value_matches_oracle_pipeline(setup):
    w, h, control, sy, energy, angles, F0, Fd = setup
    ref, _ = _oracle_yield(h, sy, control, energy, angles, F0, Fd)
    got = float(sy.value(torch.as_tensor(control.theta), F0, Fd))
    assert ref > 0
    assert abs(got - ref) <= 1e-11 * ref


def test_explicit_field_derivative_matches_differences(setup):
    w, h, control, sy, energy, angles, F0, Fd = setup
    theta = torch.tensor(control.theta, requires_grad=True)
    sy.value(theta, F0, Fd).backward()
    v = np.random.default_rng(9).normal(size=control.K)
    v /= np.linalg.norm(v)
    step = 1e-4
    J = lambda th: float(sy.value(torch.as_tensor(th), F0, Fd))
    th = control.theta
    fd = (8 * (J(th + step * v) - J(th - step * v)) - (J(th + 2 * step * v) - J(th - 2 * step * v))) / (12 * step)
    assert abs(theta.grad.numpy() @ v - fd) <= 1e-7 * abs(fd) + 1e-15

