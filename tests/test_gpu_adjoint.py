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
    assert abs(J.item() - J2) <= 1e-14


def _yield_problem(device, ionization_weight, steps=80):
    from laqphysics_amd.coupled_angular import CoupledHelium
    from laqphysics_amd.fedvr import make_grid
    from laqphysics_amd.helium3d import Helium
    from laqphysics_amd.surface import directions
    grid = make_grid([0, 1, 2, 4, 6, 8], 3, tail=6, ecs_angle=.3)
    h = CoupledHelium(Helium(grid, 2, M=0, cutoff_radii=[1, 2]), 2)
    op = HipHamiltonian(h, device)
    dt = .02
    control = PiecewiseLinearControl(np.random.default_rng(21).normal(0., 1., 10), steps * dt)
    # tol 1e-13: the discrete Krylov map is then exact to far below the differences
    p = Problem(op, workloads.initial_state(h), control, dt, steps, tol=1e-13, maxdim=12, checkpoint_every=32,
                ionization_weight=ionization_weight)
    p.add_yield(5., labels=[(2, 0, 0), (2, 1, 0)], energy=np.linspace(.2, 1.5, 8), angles=directions(True, 6, 1))
    return p


@pytest.mark.parametrize('ionization_weight', [0., 1.])
def test_adjoint_yield_gradient_matches_central_differences(device, ionization_weight):
    """The He+(n=2) tSURFF yield term (alone, and with the ionisation
    probability): costate sources at every step plus the explicit field
    dependence of the duals, phases and flux, against differences of the
    device forward solve."""
    _say('adjoint yield vs FD', ionization_weight)
    p = _yield_problem(device, ionization_weight)
    theta = p.control.theta.copy()
    g, J = p.gradient()
    assert J > 0 and np.isfinite(g).all()
    p.forced_dims = list(p.dims)
    rng = np.random.default_rng(13)
    for d in range(2):
        v = rng.normal(size=theta.size)
        v /= np.linalg.norm(v)
        # J = 1 - <psi, W psi> carries ~2e-15 of rounding (a sum of ~1e5 terms
        # near 1), which the five-point stencil at h = 1e-3 amplifies by
        # (1 + 8 + 8 + 1)/12/h: a floor of ~3e-12 absolute (its truncation is O(h^4))
        fd = _central(lambda th: _J(p, th), theta, v, 1e-3)
        assert abs(g @ v - fd) <= 1e-6 * abs(fd) + 4e-12, (g @ v, fd, J)
    p.forced_dims = None
    p.control.theta = theta
