"""Pin the CF4/Pade/GMRES oracle with the reference's closed-form checks (CPU)."""
import numpy as np
import pytest
from scipy.integrate import solve_ivp

from oracle import implicit as oi
from laqphysics_amd.fedvr import make_grid
from laqphysics_amd.helium3d import Helium
from tests.helpers import random_state


def test_separable_resolvent_inverts_exact_ionic_sum():
    # tests/test_implicit.py:25-35 of the reference
    h = Helium(make_grid([0, .5, 1, 2, 4], 3, ecs_angle=.4, tail=6), 1, M=None)
    pre = oi.SeparablePreconditioner(oi.OracleEngine(h))
    x = random_state(h.size, 816)
    alpha = .02 / (3 + 1j * np.sqrt(3))
    shift = 2.9
    y = pre.apply(x, alpha, shift).reshape(h.shape, order='F')
    out = np.empty_like(y)
    for c, ((l, m), (ll, mm)) in enumerate(h.ch):
        t = h.grid.kinetic.toarray()
        r = h.r
        h1 = t + np.diag(-2 * h.cut / r + l * (l + 1) / (2 * r * r))
        h2 = t + np.diag(-2 * h.cut / r + ll * (ll + 1) / (2 * r * r))
        out[:, :, c] = y[:, :, c] + 1j * alpha * (h1 @ y[:, :, c] + y[:, :, c] @ h2.T + shift * y[:, :, c])
    np.testing.assert_allclose(out.ravel(order='F'), x, atol=3e-10, rtol=3e-10)


class DenseEngine:
    """Mock engine (tests/test_implicit.py:41-46 of the reference)."""

    def __init__(self, H, D, a2):
        self.H, self.D, self.a2_coefficient = H, D, a2
        self.diagonal = np.diag(H).copy()

    def apply(self, x, field, velocity):
        m = self.H + np.einsum('a,aij->ij', field, self.D) + self.a2_coefficient * np.dot(field, field) * np.eye(5)
        return m @ x


@pytest.mark.parametrize('a2', [.5, 1.])
def test_cf4_fourth_order_scalar_phase_and_time_ordering(a2):
    # tests/test_implicit.py:37-55 of the reference
    rng = np.random.default_rng(677)
    z = rng.normal(size=(3, 5, 5))
    D = (z + z.transpose(0, 2, 1)) / 2
    H = np.diag(np.linspace(-2.9, 1.1, 5)).astype(complex) - .02j * np.diag(np.arange(5))
    engine = DenseEngine(H, D, a2)
    field = lambda t: np.array([.4 * np.cos(1.7 * t), .3 * np.sin(2.1 * t), .2 * t])
    initial = rng.normal(size=5) + 1j * rng.normal(size=5)
    initial /= np.linalg.norm(initial)
    exact = solve_ivp(lambda t, x: -1j * engine.apply(x, field(t), True), [0, .4], initial, method='DOP853',
                      rtol=1e-12, atol=1e-14).y[:, -1]
    errors = []
    for dt in [.1, .05]:
        x = initial.copy()
        for i in range(round(.4 / dt)):
            x, info = oi.cf4_step(engine, x, field, i * dt, dt, -2.9, tol=1e-13)
        errors.append(np.linalg.norm(x - exact))
    assert errors[1] < errors[0] / 12


def test_gmres_against_direct_solve():
    rng = np.random.default_rng(31)
    A = np.diag(np.geomspace(1, 1000, 12)).astype(complex)
    A += .03 * (rng.normal(size=A.shape) + 1j * rng.normal(size=A.shape))
    b = rng.normal(size=12) + 1j * rng.normal(size=12)
    x, info = oi.gmres(lambda y: A @ y, b, 1 / np.diag(A), tol=1e-12, restart=16)
    assert info['relative_residual'] < 1e-12
    np.testing.assert_allclose(x, np.linalg.solve(A, b), rtol=1e-10, atol=1e-11)


def test_gmres_failure_raises_reference_message():
    rng = np.random.default_rng(3)
    A = rng.normal(size=(40, 40)) + 1j * rng.normal(size=(40, 40))
    b = rng.normal(size=40) + 0j
    with pytest.raises(RuntimeError, match='GMRES did not converge'):
        oi.gmres(lambda y: A @ y, b, np.ones(40), tol=1e-14, restart=2, max_restarts=2)
