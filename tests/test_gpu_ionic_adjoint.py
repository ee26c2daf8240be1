"""He+ ionic duals on the device (gpu).

Mirror of the reference's tests/test_ionic_adjoint.py: the multi-column
device adjoint (``IonicAdjoint``, python/ionic_adjoint.py:12-52) stepped by
CF4-Pade at linear tol 1e-12 against the oracle's dense-LU backward CF4
(python/implicit.py:111-123), velocity and mixed gauge with x, y and z fields,
on an ECS grid with an irECS tail.  Plus:
  * the operator against the scipy matrix H(f)^dagger for 1 .. 70 columns
    (every lane mapping of the multi-column product, and two column blocks);
  * the preconditioner against a dense solve of the field-free resolvent;
  * the endpoint sweep with block replay equal, bit for bit, to keeping every
    frame, and both equal to the oracle's duals;
  * the reference's preparation files: reopened and resumed bit for bit.
"""
from types import SimpleNamespace

import numpy as np
import pytest
import torch

from oracle import ionic as oracle_ionic
from laqphysics_amd.fedvr import make_grid
from laqphysics_amd.ionic import Ionic
from laqphysics_amd.ionic_adjoint import DeviceDuals, IonicAdjoint
from laqphysics_amd.mixed_gauge import MixedIonic, parameters

pytestmark = pytest.mark.gpu


def small_ion(mixed=False):
    g = make_grid([0, .5, 1, 2, 4, 8], 3, tail=8, ecs_angle=.4)
    h = SimpleNamespace(grid=g, r=g.r, n=len(g.r), cut=np.ones(len(g.r)),
                        ch=[((l, m), (0, 0)) for l in range(3) for m in range(-l, l + 1)])
    return MixedIonic(h, inner=1, outer=4) if mixed else Ionic(h)


def A(t):
    return np.array([.03 * np.cos(t), .02 * np.sin(1.3 * t), .01])


def E(t):
    return np.array([.03 * np.sin(t), -.026 * np.cos(1.3 * t), 0])


@pytest.mark.parametrize('mixed', [False, True])
def test_multi_rhs_adjoint_matches_dense_lu_oracle(device, mixed):
    print(f'CASE ionic adjoint CF4 vs oracle, mixed={mixed}', flush=True)
    ion = small_ion(mixed)
    field = (lambda t: parameters(A, E, t)) if mixed else A
    chi = ion.dual_final_states([(1, 0, 0), (2, 1, 1)])
    engine = IonicAdjoint(ion, 2, device)
    state = engine.state(chi)
    for i in range(8):
        t = .16 - i * .02
        chi = oracle_ionic.cf4_step(ion.matrix(field(t - oracle_ionic.C1 * .02)).conj().T.toarray(),
                                    ion.matrix(field(t - oracle_ionic.C2 * .02)).conj().T.toarray(), chi, -.02)
        state, info = engine.step(state, field, t, -.02)
        assert info['linear_residual'] < 1e-12 and state.dtype == torch.complex128
    np.testing.assert_allclose(engine.host(state), chi, atol=2e-11, rtol=2e-10)


@pytest.mark.parametrize('nrhs', [1, 2, 3, 5, 20, 33, 64, 70])
def test_apply_matches_the_sparse_operator(device, nrhs):
    print(f'CASE ionic apply, nrhs={nrhs}', flush=True)
    ion = small_ion(mixed=nrhs == 20)
    field = parameters(A, E, .3) if nrhs == 20 else A(.3)
    engine = IonicAdjoint(ion, nrhs, device)
    coef = engine.coefficients(field)
    rng = np.random.default_rng(nrhs)
    x = rng.normal(size=(engine.ndof, nrhs)) + 1j * rng.normal(size=(engine.ndof, nrhs))
    G = ion.matrix(field).conj().T
    a, b = .3 - .7j, 1.1 + .2j
    want = a * (G @ x) + b * x
    got = engine.host(engine.apply_affine(engine.state(x), coef, True, a, b))
    assert np.max(abs(got - want)) <= 1e-13 * np.max(abs(want))
    # a second field reuses the pattern with new combined values
    field2 = parameters(E, A, .1) if nrhs == 20 else -2 * A(.7)
    want2 = ion.matrix(field2).conj().T @ x
    got2 = engine.host(engine.apply(engine.state(x), engine.coefficients(field2)))
    assert np.max(abs(got2 - want2)) <= 1e-13 * np.max(abs(want2))


@pytest.mark.parametrize('nrhs', [2, 70])
def test_preconditioner_is_the_field_free_resolvent(device, nrhs):
    print(f'CASE ionic preconditioner, nrhs={nrhs}', flush=True)
    ion = small_ion()
    engine = IonicAdjoint(ion, nrhs, device)
    rng = np.random.default_rng(7)
    x = rng.normal(size=(engine.ndof, nrhs)) + 1j * rng.normal(size=(engine.ndof, nrhs))
    alpha, shift = .01 / (3 + 1j * np.sqrt(3)), .0123
    G0 = ion.h0.conj().T.toarray()
    want = np.linalg.solve(np.eye(engine.ndof) + 1j * alpha * (G0 + shift * np.eye(engine.ndof)), x)
    got = engine.host(engine.precondition(engine.state(x), alpha, shift))
    assert np.max(abs(got - want)) <= 1e-11 * np.max(abs(want))


def test_block_replay_equals_every_frame_and_the_oracle(device):
    print('CASE ionic device duals, block replay', flush=True)
    ion = small_ion(mixed=True)
    field = lambda t: parameters(A, E, t)
    unique = [(1, 0, 0), (2, 1, -1), (2, 1, 1)]
    times = np.linspace(0., .26, 14)
    every = DeviceDuals(ion, unique, times, field, substeps=2, block=1, device=device)
    blocks = DeviceDuals(ion, unique, times, field, substeps=2, block=4, device=device)
    want = oracle_ionic.backward(lambda t: ion.matrix(field(t)).conj().T.toarray(), ion.dual_final_states(unique),
                                 times, substeps=2)
    for k in range(len(times)):
        a, b = every.host(k), blocks.host(k)
        assert np.array_equal(a, b), k
        np.testing.assert_allclose(a, want[k], atol=2e-11, rtol=2e-10)
    assert every.maximum_residual < 1e-12 and blocks.maximum_residual < 1e-12


def test_preparation_files_resume(device, tmp_path):
    """The reference's preparation files (python/streaming_surface.py:171-200):
    a complete preparation reopens without a sweep, one stopped half way
    resumes from its lowest ready endpoint, both bit for bit; another recipe
    is refused."""
    import json
    import shutil
    print('CASE ionic device duals, preparation files', flush=True)
    ion = small_ion()
    unique = [(1, 0, 0), (2, 1, 1)]
    times = np.linspace(0., .26, 14)
    options = dict(substeps=1, block=3, device=device, configuration='signature')
    full = DeviceDuals(ion, unique, times, A, folder=tmp_path / 'a', **options)
    record = json.loads((tmp_path / 'a' / 'preparation.json').read_text())
    assert record['complete'] and record['lowest_ready_frame'] == 0
    manifest = json.loads((tmp_path / 'a' / 'ionic_endpoints' / 'array.json').read_text())
    assert manifest['shape'] == [len(full.ends), full.engine.ndof, 2]
    again = DeviceDuals(ion, unique, times, A, folder=tmp_path / 'a', **options)
    assert again.linear_iterations == 0
    shutil.copytree(tmp_path / 'a', tmp_path / 'b')
    record.update(lowest_ready_frame=6, complete=False)
    (tmp_path / 'b' / 'preparation.json').write_text(json.dumps(record))
    part = DeviceDuals(ion, unique, times, A, folder=tmp_path / 'b', **options)
    assert 0 < part.linear_iterations < full.linear_iterations
    for k in range(len(times)):
        assert np.array_equal(again.host(k), full.host(k)) and np.array_equal(part.host(k), full.host(k)), k
    with pytest.raises(ValueError, match='recipe changed'):
        DeviceDuals(ion, unique, times, A, folder=tmp_path / 'a', **{**options, 'block': 4})
