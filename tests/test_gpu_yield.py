"""The device yield term of the adjoint objective (hex_yield_value /
hex_yield_gradient) against the oracle's PyTorch yield with its autograd
gradient, and against finite differences of the device value (gpu)."""
import types

import numpy as np
import pytest
import torch

from oracle.adjoint_yield import TorchYield
from laqphysics_amd import workloads
from laqphysics_amd.adjoint import PiecewiseLinearControl, SurfaceYield
from laqphysics_amd.surface import directions

pytestmark = pytest.mark.gpu

LABELS = [(2, 0, 0), (2, 1, 0)]


def _setup(device, nsteps, K, scale, seed):
    print('CASE yield', nsteps, K, scale, seed, flush=True)
    w = workloads.plain_l2_n85()
    h = w.h
    rng = np.random.default_rng(seed)
    dt = .05
    control = PiecewiseLinearControl(rng.normal(0., scale, K), nsteps * dt * .8)
    problem = types.SimpleNamespace(h=h, engine=types.SimpleNamespace(device=torch.device(device)),
                                    control=control, nsteps=nsteps, dt=dt, t0=.1)
    sy = SurfaceYield(problem, w.surface, LABELS, np.linspace(.3, 1.2, 5), directions(True, 4, 1))
    shape = (nsteps + 1, h.nc, len(h.surface_indices), h.n)
    F0 = torch.as_tensor(rng.normal(size=shape) + 1j * rng.normal(size=shape)) * 1e-3
    Fd = torch.as_tensor(rng.normal(size=shape) + 1j * rng.normal(size=shape)) * 1e-3
    return control, sy, F0, Fd


@pytest.mark.parametrize('nsteps,K,scale,seed', [(12, 6, .5, 4), (30, 9, 2., 11)])
def test_device_yield_matches_oracle_value_and_gradient(device, nsteps, K, scale, seed):
    control, sy, F0, Fd = _setup(device, nsteps, K, scale, seed)
    Y_ref, g_ref, G0_ref, Gd_ref = TorchYield(sy).gradient(control.theta, F0, Fd)
    F0d, Fdd = F0.to(device), Fd.to(device)
    Y, g, G0, Gd = sy.gradient(control.theta, F0d, Fdd)
    assert Y_ref > 0
    assert abs(Y - Y_ref) <= 1e-12 * Y_ref
    assert np.abs(g - g_ref).max() <= 1e-10 * np.abs(g_ref).max()
    for got, ref in ((G0, G0_ref), (Gd, Gd_ref)):
        got = got.cpu()
        assert float((got - ref).abs().max()) <= 1e-11 * float(ref.abs().max())


def test_device_yield_gradient_matches_its_own_differences(device):
    control, sy, F0, Fd = _setup(device, 16, 7, 1., 23)
    F0d, Fdd = F0.to(device), Fd.to(device)
    _, g, _, _ = sy.gradient(control.theta, F0d, Fdd)
    v = np.random.default_rng(5).normal(size=control.K)
    v /= np.linalg.norm(v)
    h = 1e-4
    J = lambda th: sy.value(th, F0d, Fdd)
    th = control.theta
    fd = (8 * (J(th + h * v) - J(th - h * v)) - (J(th + 2 * h * v) - J(th - 2 * h * v))) / (12 * h)
    assert abs(g @ v - fd) <= 1e-7 * abs(fd) + 1e-15


def test_frame_gradients_are_the_derivatives_along_the_frames(device):
    """dY along a frame direction (dF0, dFd) = 2 Re <G, dF> for both parts."""
    control, sy, F0, Fd = _setup(device, 10, 5, .5, 8)
    F0d, Fdd = F0.to(device), Fd.to(device)
    _, _, G0, Gd = sy.gradient(control.theta, F0d, Fdd)
    rng = np.random.default_rng(3)
    for which in (0, 1):
        d = torch.as_tensor(rng.normal(size=F0.shape) + 1j * rng.normal(size=F0.shape), device=device) * 1e-3
        G = (G0, Gd)[which]
        h = 1e-3

        def J(s):
            a, b = (F0d + s * d, Fdd) if which == 0 else (F0d, Fdd + s * d)
            return sy.value(control.theta, a.contiguous(), b.contiguous())
        fd = (8 * (J(h) - J(-h)) - (J(2 * h) - J(-2 * h))) / (12 * h)
        an = 2 * float(torch.vdot(G.reshape(-1), d.reshape(-1)).real)
        assert abs(an - fd) <= 1e-7 * abs(fd) + 1e-15
