"""Known answers and invariants of the host-side operator tables (CPU only)."""
import numpy as np
import pytest

import oracle
from laqphysics_amd.angular import C, channels, coulomb
from laqphysics_amd.bipolar import BipolarHelium
from laqphysics_amd.coupled_angular import CoupledHelium
from laqphysics_amd.fedvr import derivative, make_grid
from laqphysics_amd.helium3d import Helium
from laqphysics_amd.pulses import Pulse
from laqphysics_amd import workloads
from tests.helpers import random_state


def test_dvr_hydrogen_analytic_energies():
    from scipy.linalg import eigvalsh
    g = make_grid([0, .5, 1, 2, 4, 8, 16, 32], 8)
    for l in (0, 1, 2):
        h = g.kinetic.toarray() + np.diag(-2 / g.r + l * (l + 1) / (2 * g.r * g.r))
        np.testing.assert_allclose(eigvalsh(h)[0], -2 / (l + 1) ** 2, atol=2e-7)


def test_polynomial_derivative():
    x = np.array([-1., -.4, .2, .8, 1.])
    np.testing.assert_allclose(derivative(x) @ (x ** 4), 4 * x ** 3, atol=1e-13)


def test_high_order_irecs_has_no_growing_free_modes():
    from scipy.linalg import eigvals
    g = make_grid([0, .5, 1, 2, 4, 8, 16, 20], 6, ecs_angle=.5, tail=40, alpha=.8)
    assert eigvals(g.kinetic.toarray()).imag.max() < 1e-8


@pytest.mark.parametrize('builder', ['plain_l2_n85', 'batched_l4_L3_n441'])
def test_workload_tails_have_no_growing_free_modes(builder):
    from scipy.linalg import eigvals
    g = workloads.ALL[builder]().h.grid
    assert eigvals(g.kinetic.toarray()).imag.max() < 1e-8


def test_irecs_bridge_uses_real_element_mass():
    g = make_grid([0, 1, 2, 4, 8], 6, ecs_angle=.5, tail=20)
    inside = g.r.real <= 8
    physical = np.zeros(len(g.r), complex)
    physical[inside] = np.sin(np.pi * g.r[inside].real / 16)
    coeff = physical * np.sqrt(g.weights)
    prob = np.sum(abs(coeff) ** 2 * g.interior_weights / abs(g.weights))
    np.testing.assert_allclose(prob, 4, atol=1e-8)


def test_gaunt_sum_identity():
    basis = channels(3, M=0)
    rng = np.random.default_rng(2026)
    for _ in range(100):
        bra = basis[rng.integers(len(basis))]
        ket = basis[rng.integers(len(basis))]
        for lam in range(7):
            (a, b), (c, d) = bra, ket
            explicit = sum((-1.) ** q * C(a, lam, q, c) * C(b, lam, -q, d) for q in range(-lam, lam + 1))
            assert coulomb(bra, ket, lam) == explicit


def test_gaunt_known_value():
    # <1 0|C^1_0|0 0> = 1/sqrt(3)
    np.testing.assert_allclose(C((1, 0), 1, 0, (0, 0)), 1 / np.sqrt(3), rtol=1e-15)


def test_workload_sizes():
    w = workloads.plain_l2_n85()
    assert (w.h.n, w.h.nc) == (85, 11)
    w = workloads.batched_l4_L3_n441()
    assert (w.h.n, w.h.nc) == (441, 33)


@pytest.mark.parametrize('angle', [0., .4])
def test_coupled_basis_equals_projected_product(angle):
    # tests/test_coupled_angular.py:6-17 of the reference
    u = Helium(make_grid([0, .5, 1, 2, 4, 6, 8], 3, ecs_angle=angle, tail=6), 2, M=0, cutoff_radii=[1, 2])
    c = CoupledHelium(u, 2)
    x = random_state(c.size, 619)
    for velocity in (False, True):
        lhs = oracle.tensor_apply(c, x, (0, 0, .15), velocity)
        rhs = c.project(oracle.tensor_apply(u, c.expand(x), (0, 0, .15), velocity))
        np.testing.assert_allclose(lhs, rhs, rtol=3e-13, atol=3e-11)


@pytest.mark.parametrize('M,L,natural', [(0, 2, True), (None, 2, False)])
def test_bipolar_basis_equals_projected_product(M, L, natural):
    # tests/test_bipolar.py:7-24 of the reference
    grid = make_grid([0, 1, 2, 4, 6, 8], 3, tail=6, ecs_angle=.3)
    full = Helium(grid, 2, M=M, cutoff_radii=[1, 2])
    b = BipolarHelium(grid, 2, L, M, cutoff_radii=[1, 2], natural_parity=natural)
    U = b.transform
    np.testing.assert_allclose((U.T @ U).toarray(), np.eye(b.nc), atol=1e-14)
    x = random_state(b.size, 484)
    fields = [[0, 0, 0], [0, 0, .07]] if M == 0 else [[0, 0, 0], [.03, -.04, .07]]
    for f in fields:
        for velocity in (False, True):
            lhs = oracle.tensor_apply(b, x, f, velocity)
            rhs = b.project(oracle.tensor_apply(full, b.expand(x), f, velocity))
            np.testing.assert_allclose(lhs, rhs, atol=3e-11, rtol=3e-12)


def test_field_vector_potential_consistency():
    # tests/test_physics.py:8-15 of the reference
    p = Pulse(cycles=8)
    t = np.linspace(.001, p.duration - .001, 1000)
    h = 1e-5
    np.testing.assert_allclose(-(p.vector(t + h) - p.vector(t - h)) / (2 * h), p.electric(t), atol=2e-10)
    assert abs(p.vector(p.duration + 100)) < 1e-14
    near_one = Pulse(cycles=1 + 1e-13, field=.01, cep=.3)
    np.testing.assert_allclose(near_one.vector(near_one.duration),
                               near_one.field * near_one.duration * np.cos(near_one.cep) / 4, atol=1e-13)


def test_surface_commutator_identity():
    # tests/test_surface.py:26-35 of the reference, on the oracle apply
    h = Helium(make_grid([0, 1, 2, 4, 6, 8], 3), 1, M=None, cutoff_radii=[1, 2])
    idx = h.prepare_surface(5.)
    p = random_state(h.size, 557)
    A = np.array([.12, .23, .34])
    theta = (h.r.real > 5.).astype(float)[None, :, None]
    u = p.reshape(h.shape, order='F')
    from oracle.surface import flux as oracle_flux
    ref = (oracle.tensor_apply(h, (theta * u).ravel(order='F'), A, True).reshape(h.shape, order='F')
           - theta * oracle.tensor_apply(h, p, A, True).reshape(h.shape, order='F'))[:, idx, :]
    np.testing.assert_allclose(oracle_flux(h, p, A), ref, atol=2e-12, rtol=2e-12)
