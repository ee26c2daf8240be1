"""H.psi on the HIP engine against the oracle (gpu)."""
import json

import numpy as np
import pytest
import torch

import oracle
from laqphysics_amd import native, workloads
from laqphysics_amd.bipolar import BipolarHelium
from laqphysics_amd.coupled_angular import CoupledHelium
from laqphysics_amd.fedvr import make_grid
from laqphysics_amd.helium3d import Helium
from laqphysics_amd.hip_backend import HipHamiltonian
from tests.helpers import golden_field, golden_problem, random_state, relative_error

pytestmark = pytest.mark.gpu

# per-apply bar: ||d(H psi)|| / ||H psi|| <= 1e-12
REL = 1e-12


def _say(*what):
    print('CASE', *what, flush=True)


@pytest.mark.parametrize('i', range(6))
def test_engine_reproduces_golden_columns(golden_dir, device, i):
    _say('golden columns', i)
    d = np.load(golden_dir / 'debug_modules' / f'3d_step{i:04d}.npz')
    h = golden_problem()
    op = HipHamiltonian(h, device)
    A = golden_field(i, float(d['dt']))
    Q = d['arnoldi_Q']
    for k in range(Q.shape[1]):
        y = op.host(op.apply(op.state(Q[:, k]), A, True))
        np.testing.assert_allclose(y, d[f'Hq{k:03d}'], rtol=0, atol=1e-14)


@pytest.mark.parametrize('angle', [0., .4])
@pytest.mark.parametrize('velocity', [False, True])
def test_engine_against_oracle_ecs_all_axes(device, angle, velocity):
    _say('ecs all axes', angle, velocity)
    h = Helium(make_grid([0, .5, 1, 2, 4, 6, 8], 3, ecs_angle=angle, tail=8), 2, M=None, cutoff_radii=[1, 2])
    op = HipHamiltonian(h, device)
    x = random_state(h.size, 386)
    for f in [(0., 0., 0.), (.13, .23, .34)]:
        y = op.host(op.apply(op.state(x), f, velocity))
        assert relative_error(y, oracle.tensor_apply(h, x, f, velocity)) <= REL


@pytest.mark.parametrize('seed', [49, 386, 619])
@pytest.mark.parametrize('velocity', [False, True])
def test_engine_against_oracle_plain_l2_n85(device, seed, velocity):
    _say('plain_l2_n85', seed, velocity)
    h = workloads.plain_l2_n85().h
    op = HipHamiltonian(h, device)
    x = random_state(h.size, seed)
    f = (0., 0., .05)
    y = op.host(op.apply(op.state(x), f, velocity))
    assert relative_error(y, oracle.tensor_apply(h, x, f, velocity)) <= REL


def test_engine_against_oracle_batched_l4_n441(device):
    _say('batched_l4_L3_n441 apply')
    h = workloads.batched_l4_L3_n441().h
    op = HipHamiltonian(h, device)
    x = random_state(h.size, 619)
    f = (0., 0., .05)
    y = op.host(op.apply(op.state(x), f, True))
    assert relative_error(y, oracle.tensor_apply(h, x, f, True)) <= REL


def test_engine_against_oracle_circular_l5_n301(device):
    """The full-size circular-polarisation operator (bipolar, all M, both
    parities: 771 channels, n = 301, N = 69.9M) with an x-y field, and the
    adjoint identity <y, H x> = <H^dagger y, x> at that size."""
    _say('circular_l5_n301 apply')
    h = workloads.circular_l5_n301().h
    op = HipHamiltonian(h, device)
    x = random_state(h.size, 49)
    f = (.02, .015, 0.)
    y = op.host(op.apply(op.state(x), f, True))
    assert relative_error(y, oracle.tensor_apply(h, x, f, True)) <= REL
    z = random_state(h.size, 386)
    xs, zs = op.state(x), op.state(z)
    lhs = torch.vdot(zs, op.apply(xs, f, True)).item()
    rhs = torch.vdot(op.apply_adjoint(zs, f, True), xs).item()
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


def test_bipolar_all_m_circular_field(device):
    _say('bipolar all M')
    grid = make_grid([0, 1, 2, 4, 6, 8], 3, tail=6, ecs_angle=.3)
    h = BipolarHelium(grid, 2, 2, M=None, cutoff_radii=[1, 2])
    op = HipHamiltonian(h, device)
    x = random_state(h.size, 908)
    for velocity in (False, True):
        f = (.03, -.04, .07)
        y = op.host(op.apply(op.state(x), f, velocity))
        assert relative_error(y, oracle.tensor_apply(h, x, f, velocity)) <= REL


def test_dropin_tensor_apply_host_buffers(device):
    _say('drop-in tensor_apply')
    h = CoupledHelium(Helium(make_grid([0, .5, 1, 2, 4, 6, 8], 3, ecs_angle=.4, tail=6), 2, M=0,
                             cutoff_radii=[1, 2]), 2)
    x = random_state(h.size, 12)
    for velocity in (False, True):
        y = native.tensor_apply(h, x, (0, 0, .15), velocity)
        assert relative_error(y, oracle.tensor_apply(h, x, (0, 0, .15), velocity)) <= REL


def test_apply_rejects_wrong_state(device):
    import torch
    h = golden_problem()
    op = HipHamiltonian(h, device)
    with pytest.raises(ValueError):
        op.apply(torch.zeros(h.size + 1, dtype=torch.complex128, device=device))


@pytest.mark.parametrize('angle', [0., .4])
def test_device_flux_against_oracle(device, angle):
    # tests/test_torch_backend.py:8-17 of the reference (flux part)
    _say('flux', angle)
    from oracle.surface import flux as oracle_flux
    h = Helium(make_grid([0, .5, 1, 2, 4, 6, 8], 3, ecs_angle=angle, tail=8), 1, M=None, cutoff_radii=[1, 2])
    h.prepare_surface(5)
    op = HipHamiltonian(h, device)
    x = random_state(h.size, 386)
    for f in [(0., 0., 0.), (.13, .23, .34)]:
        np.testing.assert_allclose(op.flux(op.state(x), f), oracle_flux(h, x, f), atol=2e-12, rtol=2e-12)


def test_device_flux_plain_l2_n85(device):
    _say('flux plain_l2_n85')
    from oracle.surface import flux as oracle_flux
    w = workloads.plain_l2_n85()
    h = w.h
    h.prepare_surface(w.surface)
    op = HipHamiltonian(h, device)
    x = random_state(h.size, 49)
    f = (0., 0., .05)
    np.testing.assert_allclose(op.flux(op.state(x), f), oracle_flux(h, x, f), atol=2e-12, rtol=2e-12)


@pytest.mark.parametrize('angle', [0., .4])
@pytest.mark.parametrize('velocity', [False, True])
def test_adjoint_apply_against_kronecker_conjugate_transpose(device, angle, velocity):
    _say('adjoint', angle, velocity)
    from oracle.kron import reference_sparse
    h = Helium(make_grid([0, .5, 1, 2, 4, 6, 8], 3, ecs_angle=angle, tail=8), 1, M=None, cutoff_radii=[1, 2])
    op = HipHamiltonian(h, device)
    x = random_state(h.size, 21)
    f = (.13, .23, .34)
    ref = reference_sparse(h, f, velocity).conj().T @ x
    y = op.host(op.apply_adjoint(op.state(x), f, velocity))
    assert relative_error(y, ref) <= REL


def test_adjoint_apply_plain_l2_n85_inner_product_identity(device):
    _say('adjoint plain_l2_n85')
    h = workloads.plain_l2_n85().h
    op = HipHamiltonian(h, device)
    x = random_state(h.size, 3)
    y = random_state(h.size, 4)
    f = (0., 0., .05)
    Hy = oracle.tensor_apply(h, y, f, True)
    Hdx = op.host(op.apply_adjoint(op.state(x), f, True))
    lhs, rhs = np.vdot(x, Hy), np.vdot(Hdx, y)
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


@pytest.mark.parametrize('rows', [0, 1, 2, 4, 8, 16])
def test_launch_shapes_agree_with_the_oracle(device, rows):
    """Every radial launch shape the autotuner may pick (direct walks with 1-8
    rows per block, the matrix-core walks, against the staged default) gives
    the oracle's H.psi."""
    _say('launch shape', rows)
    h = workloads.plain_l2_n85().h
    op = HipHamiltonian(h, device)
    x = random_state(h.size, 5)
    f = (0., 0., .05)
    ref = op.host(op.apply(op.state(x), f, True))
    op.set_launch(rows)
    got = op.host(op.apply(op.state(x), f, True))
    assert relative_error(got, oracle.tensor_apply(h, x, f, True)) <= REL
    assert relative_error(got, ref) <= 1e-14


def test_plain_autotune_picks_a_shape_and_caches_it(device, tmp_path, monkeypatch):
    """HipHamiltonian.autotune times every shape once, sets the choice, caches
    it per (order, n, nc, members) and reuses the cache; the apply after the
    choice is still the oracle's."""
    _say('plain autotune')
    from laqphysics_amd import batched
    monkeypatch.setattr(batched, 'TUNING', tmp_path / 'autotune.json')
    h = workloads.plain_l2_n85().h
    op = HipHamiltonian(h, device)
    shape = op.autotune()
    assert shape in (0, 1, 2, 4, 8, 16)
    table = json.loads((tmp_path / 'autotune.json').read_text())
    key = batched._key(op, 1)
    assert int(table[key]['rows_per_block']) == shape and len(table[key]['ms_per_apply']) == 6
    times = table[key]['ms_per_apply']
    assert times[str(shape)] <= min(times.values()) / .97 + 1e-12
    again = HipHamiltonian(h, device)
    assert again.autotune() == shape
    x = random_state(h.size, 6)
    f = (0., 0., .05)
    assert relative_error(again.host(again.apply(again.state(x), f, True)), oracle.tensor_apply(h, x, f, True)) <= REL


def test_batched_apply_equals_member_applies(device):
    """The B-member apply (each member its own field) returns exactly the
    single-state applies, on the n = 441 grid."""
    _say('batched apply B=4 batched_l4_L3_n441')
    h = workloads.batched_l4_L3_n441().h
    op = HipHamiltonian(h, device)
    xs = torch.stack([op.state(random_state(h.size, s)) for s in (1, 2, 3, 4)])
    fields = [(0., 0., .05), (0., 0., -.02), (0., 0., 0.), (0., 0., .31)]
    got = op.apply_batched(xs, fields, True)
    for b in range(4):
        ref = op.apply(xs[b], fields[b], True)
        assert torch.equal(got[b], ref), b
