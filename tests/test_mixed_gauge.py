"""Mixed length/velocity gauge: host tables and the oracle, on the CPU.

Pins (python/mixed_gauge.py; the reference's tests/test_mixed_gauge.py):
  * uniform limits: a switch of zero width at the origin is the velocity
    gauge, a switch beyond the grid the length gauge -- both against the C
    restatement of the reference kernel;
  * the projected rhat_a rhat_b operators sum to one and give
    <p0|cos^2 theta|p0> = 3/5 only with the intermediate shell included;
  * the one-electron mixed Hamiltonian approximates exp(i A r f cos) H_V
    exp(-i A r f cos) + E r f cos, the defect falling at least threefold per
    element-order step;
  * the coupled oracle is the projection of the product oracle.
"""
from types import SimpleNamespace

import numpy as np
import pytest
from scipy.special import eval_legendre

import oracle
from oracle import mixed as oracle_mixed
from laqphysics_amd.coupled_angular import CoupledHelium
from laqphysics_amd.fedvr import make_grid
from laqphysics_amd.helium3d import Helium
from laqphysics_amd.ionic import Ionic
from laqphysics_amd.mixed_gauge import MixedIonic, angular_squares, parameters, profile
from helpers import random_state

A = np.array([.03, -.04, .05])
E = np.array([.02, .03, -.01])


def test_parameters_are_A_E_and_doubled_products():
    c = parameters(lambda t: A, lambda t: E, 0.)
    assert c.shape == (12,)
    assert np.array_equal(c[:3], A) and np.array_equal(c[3:6], E)
    assert np.allclose(c[6:], [A[0] ** 2, A[1] ** 2, A[2] ** 2, 2 * A[0] * A[1], 2 * A[0] * A[2], 2 * A[1] * A[2]])


@pytest.mark.parametrize('inner,outer,velocity', [(0., 0., True), (100., 101., False)])
def test_uniform_limits_are_the_native_gauges(inner, outer, velocity):
    h = Helium(make_grid([0, .5, 1, 2, 4, 8], 4), 1, M=None)
    x = random_state(h.size, 511)
    coeff = parameters(lambda t: A, lambda t: E, 0.)
    got = oracle_mixed.apply(h, x, coeff, inner, outer)
    want = oracle.tensor_apply(h, x, A if velocity else E, velocity)
    np.testing.assert_allclose(got, want, atol=3e-11, rtol=3e-12)


def test_profile_limits():
    r = make_grid([0, 1, 3, 6], 5, ecs_angle=.3, tail=6).r
    q, F, c0, c2 = profile(r, 0., 0.)
    assert np.array_equal(q, r) and not np.any(F) and np.all(c0 == 1) and not np.any(c2)
    q, F, c0, c2 = profile(r, 1., 4.)
    inside, outside = r.real < 1, r.real > 4
    assert np.allclose(q[inside], 0) and np.allclose(F[inside], r[inside])
    assert np.array_equal(q[outside], r[outside]) and not np.any(F[outside]) and not np.any(c2[outside])
    with pytest.raises(ValueError, match='inner < outer'):
        profile(r, 3., 2.)


def test_angular_squares_include_the_intermediate_shell():
    states = [(l, m) for l in range(3) for m in range(-l, l + 1)]
    squares = angular_squares(states)
    np.testing.assert_allclose(sum(squares[:3]), np.eye(len(states)), atol=1e-14)
    assert abs(squares[2, states.index((1, 0)), states.index((1, 0))] - .6) < 1e-14


def test_one_electron_gauge_identity_converges_radially():
    """exp(i A r f cos) H_V(A) exp(-i A r f cos) - E r f cos against H_mixed,
    on a real grid with lmax = 6 partial waves (Gauss-Legendre in cos)."""
    z, w = np.polynomial.legendre.leggauss(48)
    a, e, lmax = .03, .02, 6
    harmonics = np.array([np.sqrt((2 * l + 1) / 2) * eval_legendre(l, z) for l in range(lmax + 1)])
    errors = []
    for order in (4, 6, 8):
        g = make_grid([0, .5, 1, 2, 3, 4, 6, 8, 12], order)
        one = SimpleNamespace(grid=g, n=len(g.r), r=g.r, cut=np.ones(len(g.r)),
                              ch=[((l, 0), (0, 0)) for l in range(lmax + 1)])
        ion = Ionic(one)
        mixed = MixedIonic(one, inner=2, outer=4)
        psi = ion.final_states([(1, 0, 0)])[:, 0]
        q, F, c0, c2 = profile(g.r, 2, 4)
        U = np.einsum('lz,jz,rz,z->rlj', harmonics, harmonics, np.exp(1j * a * F[:, None] * z[None, :]), w)

        def gauge(x):
            return np.einsum('rlj,jr->lr', U, x.reshape(lmax + 1, -1)).reshape(-1)
        v = gauge(psi)
        coeff = parameters(lambda t: np.array([0, 0, a]), lambda t: np.array([0, 0, e]), 0.)
        defect = mixed.matrix(coeff) @ v - gauge(ion.matrix([0, 0, a]) @ psi) - e * (mixed.Df[2] @ v)
        errors.append(np.linalg.norm(defect))
        H = mixed.matrix(coeff)
        assert np.max(abs((H - H.conj().T).data), initial=0) < 1e-11     # Hermitian on a real grid
    assert errors[1] < errors[0] / 3 and errors[2] < errors[1] / 3, errors
    assert errors[-1] < 2e-7


def test_coupled_oracle_is_the_projected_product_oracle():
    h = Helium(make_grid([0, 1, 2, 4, 6], 3, ecs_angle=.3, tail=5), 2, M=0, cutoff_radii=[2., 4.])
    c = CoupledHelium(h, 2)
    x = random_state(c.size, 882)
    coeff = parameters(lambda t: np.array([0, 0, .03]), lambda t: np.array([0, 0, .02]), 0.)
    small = oracle_mixed.apply(c, x, coeff, 1., 4.)
    full = c.project(oracle_mixed.apply(h, c.expand(x), coeff, 1., 4.))
    np.testing.assert_allclose(small, full, atol=1e-13, rtol=1e-13)


def test_bipolar_squares_are_the_projected_product_squares():
    """The bipolar basis forms rhat_a rhat_b from its rank-2 channel tensors
    (python/mixed_gauge.py:45-51); on the full all-M product space they equal
    the product basis's Kronecker squares seen through the Clebsch-Gordan
    matrix, for both electrons and all six pairs."""
    from laqphysics_amd.bipolar import BipolarHelium
    from laqphysics_amd.helium3d import Helium
    from laqphysics_amd.mixed_gauge import two_electron_squares
    grid = make_grid([0, 1, 2, 3], 3)
    b = BipolarHelium(grid, 2, 4, M=None)
    p = Helium(grid, 2, M=None)
    U = b.transform.toarray()
    assert U.shape[0] == p.nc and np.allclose(U.T @ U, np.eye(b.nc), atol=1e-14)
    for sb, sp in zip(two_electron_squares(b), two_electron_squares(p)):
        for qb, qp in zip(sb, sp):
            np.testing.assert_allclose(qb.toarray(), U.T @ qp.toarray() @ U, atol=2e-15)
