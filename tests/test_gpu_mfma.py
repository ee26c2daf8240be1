"""Radial walks on the matrix cores (launch shape 16) against the oracle and
the VALU staged walks (gpu).

The matrix-core pass runs both radial walks as FP64 16 x 16 x 4 GEMMs over
16-row blocks of the band; it must give the oracle's H.psi on every grid
kind: irECS tails (complex blocks), grids whose last block is partial,
elements of several orders, both gauges, the adjoint, the term masks of the
adjoint gradient and the B-member apply.
"""
import numpy as np
import pytest
import torch

import oracle
from laqphysics_amd import workloads
from laqphysics_amd.fedvr import make_grid
from laqphysics_amd.helium3d import Helium
from laqphysics_amd.hip_backend import HipHamiltonian
from tests.helpers import random_state, relative_error

pytestmark = pytest.mark.gpu

MFMA = 16
REL = 1e-12


def _say(*what):
    print('CASE', *what, flush=True)


def _ops(h, device):
    staged = HipHamiltonian(h, device)
    staged.set_launch(0)
    mfma = HipHamiltonian(h, device)
    mfma.set_launch(MFMA)
    return staged, mfma


@pytest.mark.parametrize('angle', [0., .4])
@pytest.mark.parametrize('velocity', [False, True])
def test_mfma_against_oracle_ecs_all_axes(device, angle, velocity):
    """Order-3 elements (n = 25: one full and one partial block), all M."""
    _say('mfma ecs all axes', angle, velocity)
    h = Helium(make_grid([0, .5, 1, 2, 4, 6, 8], 3, ecs_angle=angle, tail=8), 2, M=None, cutoff_radii=[1, 2])
    staged, op = _ops(h, device)
    x = random_state(h.size, 386)
    for f in [(0., 0., 0.), (.13, .23, .34)]:
        y = op.host(op.apply(op.state(x), f, velocity))
        assert relative_error(y, oracle.tensor_apply(h, x, f, velocity)) <= REL
        assert relative_error(y, staged.host(staged.apply(staged.state(x), f, velocity))) <= 1e-13


@pytest.mark.parametrize('velocity', [False, True])
def test_mfma_against_oracle_plain_l2_n85(device, velocity):
    _say('mfma plain_l2_n85', velocity)
    h = workloads.plain_l2_n85().h
    staged, op = _ops(h, device)
    for seed in (49, 619):
        x = random_state(h.size, seed)
        f = (0., 0., .05)
        y = op.host(op.apply(op.state(x), f, velocity))
        assert relative_error(y, oracle.tensor_apply(h, x, f, velocity)) <= REL


@pytest.mark.parametrize('angle', [0., .4])
def test_mfma_adjoint_and_terms(device, angle):
    """H^dagger against the Kronecker conjugate transpose, and the term masks
    (base only, field coupling only, affine) against the staged walks."""
    _say('mfma adjoint and terms', angle)
    from oracle.kron import reference_sparse
    h = Helium(make_grid([0, .5, 1, 2, 4, 6, 8], 3, ecs_angle=angle, tail=8), 1, M=None, cutoff_radii=[1, 2])
    staged, op = _ops(h, device)
    x = random_state(h.size, 21)
    f = (.13, .23, .34)
    for velocity in (False, True):
        ref = reference_sparse(h, f, velocity).conj().T @ x
        y = op.host(op.apply_adjoint(op.state(x), f, velocity))
        assert relative_error(y, ref) <= REL
    xs = op.state(x)
    for terms in (1, 2, 3, 7):
        for adjoint in (False, True):
            a = op.apply_terms(xs, f, True, terms=terms, a=.3 - .2j, b=1.5, adjoint=adjoint)
            b = staged.apply_terms(staged.state(x), f, True, terms=terms, a=.3 - .2j, b=1.5, adjoint=adjoint)
            assert relative_error(op.host(a), staged.host(b)) <= 1e-13, (terms, adjoint)


def test_mfma_production_l6_n441(device):
    """Full-size production operator (n = 441, 72 channels), both gauges."""
    _say('mfma production_l6_L4_n441')
    h = workloads.production_l6_L4_n441().h
    op = HipHamiltonian(h, device)
    op.set_launch(MFMA)
    x = random_state(h.size, 49)
    for velocity in (True, False):
        f = (0., 0., .05)
        y = op.host(op.apply(op.state(x), f, velocity))
        assert relative_error(y, oracle.tensor_apply(h, x, f, velocity)) <= REL


def test_mfma_batched_members(device):
    """The B-member apply on the matrix cores equals the single-state applies
    bit for bit (each member its own field), n = 441."""
    _say('mfma batched B=4 batched_l4_L3_n441')
    h = workloads.batched_l4_L3_n441().h
    op = HipHamiltonian(h, device)
    op.set_launch(MFMA)
    xs = torch.stack([op.state(random_state(h.size, s)) for s in (1, 2, 3, 4)])
    fields = [(0., 0., .05), (0., 0., -.02), (0., 0., 0.), (0., 0., .31)]
    got = op.apply_batched(xs, fields, True)
    for b in range(4):
        assert torch.equal(got[b], op.apply(xs[b], fields[b], True)), b
    y = op.host(got[0])
    assert relative_error(y, oracle.tensor_apply(h, op.host(xs[0]), fields[0], True)) <= REL
