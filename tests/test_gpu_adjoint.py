"""Adjoint gradient against central finite differences of the device forward
solve and of the oracle's CPU forward solve (gpu)."""
import numpy as np
import pytest
import torch

import oracle
from oracle import arnoldi as oracle_arnoldi
from laqphysics_amd import workloads
from laqphysics_amd.adjoint import Objective, PiecewiseLinearControl, Problem
from laqphysics_amd.hip_backend import HipHamiltonian

pytestmark = pytest.mark.gpu


def _say(*what):
    print('CASE', *what, flush=True)


def _problem(device, steps=150, K=12, seed=3):
    w = workloads.plain_l2_n85()
    h = w.h
    op = HipHamiltonian(h, device)
    rng = np.random.default_rng(seed)
    T = steps * w.dt
    control = PiecewiseLinearControl(rng.normal(0., 1., K), T)
    # the inner region r1, r2 <= 1.2 a.u. makes J sensitive to the field within a few a.u. of time
    return Problem(op, workloads.initial_state(h), control, w.dt, steps, tol=w.tolerance, maxdim=w.krylov_dimension,
                   checkpoint_every=64, radius=1.2)


def _J(problem, theta):
    problem.control.theta = np.asarray(theta, float)
    return problem.objective(problem.forward(keep=False)[0])


def test_control_vector_potential_is_minus_integral_of_field():
    c = PiecewiseLinearControl(np.random.default_rng(1).normal(size=9), 4.)
    t = np.linspace(0, 4, 4001)
    E = np.array([c.electric(s) for s in t])
    A = np.array([c.vector(s) for s in t])
    from scipy.integrate import cumulative_trapezoid
    np.testing.assert_allclose(A, -cumulative_trapezoid(E, t, initial=0), atol=1e-6)


def _central(J, theta, v, h):
    """Fourth-order central difference of J along v.  At the gradient scale of
    these problems (|dJ| ~ 1e-3) a second-order stencil at h = 1e-6 is limited
    by rounding, (|J| eps)/h ~ 1e-9 absolute, i.e. ~1e-6 relative; h = 1e-4 with
    the five-point stencil has truncation O(h^4) and rounding ~1e-11."""
    return (8 * (J(theta + h * v) - J(theta - h * v)) - (J(theta + 2 * h * v) - J(theta - 2 * h * v))) / (12 * h)


def test_adjoint_gradient_matches_central_differences(device):
    _say('adjoint vs FD (device forward)')
    p = _problem(device)
    theta = p.control.theta.copy()
    g, J = p.gradient()
    assert any(len(leaves) > 1 for leaves in p.dims), 'the field is meant to force subdivided steps'
    p.forced_dims = list(p.dims)
    rng = np.random.default_rng(11)
    for d in range(3):
        v = rng.normal(size=theta.size)
        v /= np.linalg.norm(v)
        fd = _central(lambda th: _J(p, th), theta, v, 1e-4)
        assert abs(g @ v - fd) <= 1e-6 * abs(fd) + 1e-12, (g @ v, fd)
    p.forced_dims = None
    p.control.theta = theta


def test_adjoint_gradient_matches_oracle_cpu_differences(device):
    _say('adjoint vs FD (oracle forward)')
    p = _problem(device, steps=100)
    h = p.h
    g, J = p.gradient()
    dims = list(p.dims)
    theta = p.control.theta.copy()
    v = np.random.default_rng(5).normal(size=theta.size)
    v /= np.linalg.norm(v)
    weight = p.weight.cpu().numpy()

    def J_cpu(th):
        c = PiecewiseLinearControl(th, p.control.T)
        x = workloads.initial_state(h)
        for n in range(p.nsteps):
            for t, dt, m in dims[n]:
                A = c.avec(t + dt / 2)
                x = oracle_arnoldi.exponential_action_fixed(lambda y: oracle.tensor_apply(h, y, A, True), x, dt, m)
        return 1. - np.vdot(x, weight * x).real
    fd = _central(J_cpu, theta, v, 1e-4)
    assert abs(g @ v - fd) <= 1e-6 * abs(fd) + 1e-12, (g @ v, fd)
    assert abs(J - J_cpu(theta)) <= 1e-10


def test_autograd_function_returns_the_costate_gradient(device):
    _say('adjoint autograd')
    p = _problem(device, steps=70)
    theta = torch.tensor(p.control.theta, requires_grad=True)
    J = Objective.apply(theta, p)
    J.backward()
    g, J2 = p.gradient()
    np.testing.assert_allclose(theta.grad.numpy(), g, rtol=1e-12, atol=1e-15)
    assert abs(float(J) - J2) <= 1e-14
