"""Pin the oracle: operator tables, the C restatement of the H.psi kernel and the
CPU Arnoldi against the reference's golden vectors, the reference's own Fortran
kernel (when built into oracle/_ref), the independent Kronecker assembly and
closed forms.  CPU only."""
import subprocess

import numpy as np
import pytest
from scipy.linalg import expm

import oracle
from oracle import arnoldi as oracle_arnoldi
from oracle.kron import reference_sparse
from laqphysics_amd.fedvr import make_grid
from laqphysics_amd.helium3d import Helium
from tests.helpers import golden_field, golden_problem, random_state

STEPS = range(6)


def _golden(golden_dir, i):
    return np.load(golden_dir / 'debug_modules' / f'3d_step{i:04d}.npz')


@pytest.mark.parametrize('i', STEPS)
def test_tables_reproduce_reference_hamiltonian(golden_dir, i):
    """The grid's GLL derivative is the closed form, not the reference's
    barycentric one, so the assembled matrix agrees to rounding: every entry
    within 2e-15 of the largest (measured: 8.9e-16)."""
    d = _golden(golden_dir, i)
    h = golden_problem()
    H = reference_sparse(h, golden_field(i, float(d['dt'])), True).toarray()
    ref = d['reference_hamiltonian']
    assert np.abs(H - ref).max() <= 2e-15 * np.abs(ref).max()
    np.testing.assert_array_equal(H != 0, ref != 0)


@pytest.mark.parametrize('i', STEPS)
def test_oracle_apply_reproduces_every_golden_column(golden_dir, i):
    d = _golden(golden_dir, i)
    h = golden_problem()
    A = golden_field(i, float(d['dt']))
    Q = d['arnoldi_Q']
    for k in range(Q.shape[1]):
        np.testing.assert_allclose(oracle.tensor_apply(h, Q[:, k], A, True), d[f'Hq{k:03d}'], rtol=0, atol=1e-14)
    np.testing.assert_allclose(oracle.tensor_apply(h, d['input'], A, True), d['reference_hamiltonian'] @ d['input'],
                               rtol=1e-13, atol=2e-15)


@pytest.mark.parametrize('i', STEPS)
def test_oracle_arnoldi_reproduces_golden_trace(golden_dir, i):
    d = _golden(golden_dir, i)
    h = golden_problem()
    A = golden_field(i, float(d['dt']))
    snap = {}
    y, info = oracle_arnoldi.exponential_action(lambda x: oracle.tensor_apply(h, x, A, True), d['input'],
                                                float(d['dt']), tol=1e-12, trace=snap.__setitem__)
    assert info['dimension'] == d['arnoldi_H'].shape[0]
    np.testing.assert_allclose(snap['arnoldi_H'], d['arnoldi_H'], atol=1e-13)
    np.testing.assert_allclose(snap['arnoldi_Q'], d['arnoldi_Q'], atol=1e-13)
    np.testing.assert_allclose(y, d['output'], atol=1e-14)
    np.testing.assert_allclose(y, d['reference_output'], atol=2e-11)


@pytest.mark.parametrize('angle', [0., .3])
@pytest.mark.parametrize('velocity', [False, True])
def test_oracle_against_independent_kronecker_assembly(angle, velocity):
    # tests/test_kernels.py:32-40 of the reference
    g = make_grid([0, .5, 1, 2, 4], 3, ecs_angle=angle, tail=5 if angle else 0)
    h = Helium(g, 1, M=None)
    f = np.array([.13, .23, .34])
    x = random_state(h.size, 49)
    ref = reference_sparse(h, f, velocity)
    np.testing.assert_allclose(oracle.tensor_apply(h, x, f, velocity), ref @ x, rtol=2e-13, atol=2e-12)
    if not angle:
        dense = ref.toarray()
        np.testing.assert_allclose(dense, dense.conj().T, atol=2e-12)


@pytest.mark.skipif(not oracle.have_reference_kernel(), reason='reference kernel not built into oracle/_ref')
@pytest.mark.parametrize('velocity', [False, True])
def test_oracle_matches_reference_fortran_kernel_bit_for_bit(velocity):
    g = make_grid([0, .5, 1, 2, 4, 6, 8], 3, ecs_angle=.4, tail=8)
    h = Helium(g, 2, M=None, cutoff_radii=[1, 2])
    x = random_state(h.size, 386)
    f = (.13, .23, .34)
    a = oracle.tensor_apply(h, x, f, velocity)
    b = oracle.ref_tensor_apply(h, x, f, velocity)
    assert np.max(abs(a - b)) <= 4e-15 * np.max(abs(b))


@pytest.mark.parametrize('plant', [1, 2, 3, 4, 5])
def test_planted_oracle_mistakes_fail_the_golden_checks(golden_dir, tmp_path, plant):
    """Each planted mistake of the C restatement is caught by the golden H or
    the Kronecker assembly, so those checks pin the oracle."""
    import ctypes as ct
    so = tmp_path / f'plant{plant}.so'
    subprocess.run(['gcc', '-O1', '-fopenmp', '-fPIC', '-shared', '-std=c11', f'-DORACLE_PLANT={plant}',
                    '-o', str(so), str(oracle.HERE / 'tensor_apply.c')], check=True)
    lib = ct.CDLL(str(so))
    worst = 0.
    for i in (0, 3):
        d = _golden(golden_dir, i)
        h = golden_problem()
        A = golden_field(i, float(d['dt']))
        for velocity in (False, True):
            y = oracle._call(lib, 'oracle_tensor_apply', h, d['input'], A, velocity)
            ref = reference_sparse(h, A, velocity) @ d['input']
            worst = max(worst, np.max(abs(y - ref)))
    g = make_grid([0, .5, 1, 2, 4], 3, ecs_angle=.3, tail=5)
    h = Helium(g, 2, M=None, cutoff_radii=[1, 2])
    x = random_state(h.size, 7)
    for velocity in (False, True):
        y = oracle._call(lib, 'oracle_tensor_apply', h, x, (.13, .23, .34), velocity)
        worst = max(worst, np.max(abs(y - reference_sparse(h, (.13, .23, .34), velocity) @ x)))
    assert worst > 1e-6


def test_arnoldi_against_dense_exponential():
    # tests/test_kernels.py:42-47 of the reference
    rng = np.random.default_rng(762)
    A = rng.normal(size=(20, 20)) + 1j * rng.normal(size=(20, 20))
    H = (A + A.conj().T) / 2 - .1j * np.diag(np.arange(20))
    x = rng.normal(size=20) + 1j * rng.normal(size=20)
    y, info = oracle_arnoldi.exponential_action(lambda z: H @ z, x, .3, tol=1e-12)
    np.testing.assert_allclose(y, expm(-.3j * H) @ x, atol=2e-11, rtol=2e-11)


def test_arnoldi_subdivides_on_small_krylov_space():
    rng = np.random.default_rng(5)
    A = rng.normal(size=(30, 30))
    H = (A + A.T) / 2
    x = rng.normal(size=30) + 0j
    y, info = oracle_arnoldi.step(lambda t, z: H @ z, x, 0., .5, tol=1e-8, maxdim=6)
    assert info.get('subdivided')
    np.testing.assert_allclose(y, expm(-.5j * H) @ x, atol=1e-6)


def test_fixed_dimension_arnoldi_equals_adaptive_at_the_adaptive_dimension(golden_dir):
    d = _golden(golden_dir, 2)
    h = golden_problem()
    A = golden_field(2, float(d['dt']))
    apply = lambda x: oracle.tensor_apply(h, x, A, True)
    y, info = oracle_arnoldi.exponential_action(apply, d['input'], float(d['dt']), tol=1e-12)
    z = oracle_arnoldi.exponential_action_fixed(apply, d['input'], float(d['dt']), info['dimension'])
    np.testing.assert_array_equal(y, z)
