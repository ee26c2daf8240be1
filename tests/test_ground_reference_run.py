"""Ground state and the reference's recorded 3D run, on the CPU.

Fixtures: tests/golden/3d_dt_refined/{config,run,history}.json are the
reference's own results of one complete run (results/3d_dt_refined/): product
basis lmax = 1, M = 0 (88 radial points, 6 channels), irECS tail at 0.5 rad,
velocity-gauge z pulse (1.5, 8 cycles, 0.005), Arnoldi dimension 48 at tol
1e-10, 2952 steps of dt 0.02998..., ground solve at tol 1e-9 with energy
-2.8972353187832827; history rows every 100 steps.

The oracle's ground solve (python/ground_s.py:55-89 restated) reproduces that
energy; without electron repulsion its energy must be twice the ion's lowest
level (a closed form of the separable problem); the compact L = 0 operator must
be the L = 0 block of the full operator; the first 100 steps of the oracle's
run must reproduce the reference's recorded history row.
"""
import json

import numpy as np
import pytest
from scipy.linalg import eigh

import oracle
from oracle import arnoldi as oracle_arnoldi
from oracle import ground as oracle_ground
from laqphysics_amd.coupled_angular import CoupledHelium
from laqphysics_amd.fedvr import make_grid
from laqphysics_amd.ground import GroundS
from laqphysics_amd.helium3d import Helium
from laqphysics_amd.pulses import vector_function
from laqphysics_amd.run import config_pulses, observables, schedule, setup, transfer
from helpers import random_state


@pytest.fixture(scope='module')
def recorded(golden_dir):
    d = golden_dir / '3d_dt_refined'
    return {name: json.loads((d / f'{name}.json').read_text()) for name in ('config', 'run', 'history')}


@pytest.fixture(scope='module')
def pair(recorded):
    return setup(recorded['config'])


@pytest.fixture(scope='module')
def oracle_ground_state(pair, recorded):
    hreal, _ = pair
    g = GroundS(hreal)
    return g, oracle_ground.solve(g, tol=recorded['config']['ground_tolerance'])


def test_reference_run_setup(pair, recorded):
    hreal, h = pair
    run, cfg = recorded['run'], recorded['config']
    assert (h.n, h.nc) == (run['nrad'], run['channels'])
    assert hreal.grid.ecs_angle == 0 and h.grid.ecs_angle == cfg['radial']['ecs_angle']
    T, nsteps, dt = schedule(config_pulses(cfg), cfg['dt'], cfg['surface_stride'], cfg.get('end_time'),
                             cfg['post_time'])
    assert nsteps == run['nsteps'] and dt == run['dt']
    assert list(h.prepare_surface(cfg['surface'])) == run['surface_indices']


def test_oracle_ground_energy_matches_reference_run(oracle_ground_state, recorded):
    g, (E, small, res, its) = oracle_ground_state
    assert abs(E - recorded['run']['ground_energy']) <= 1e-12
    assert res <= recorded['config']['ground_tolerance']
    assert abs(np.linalg.norm(small) - 1) <= 1e-12


def test_oracle_ground_without_repulsion_is_twice_the_ion():
    grid = make_grid([0., 1., 2., 4., 7., 10., 14.], 6)
    g = GroundS(Helium(grid, 1, M=0))
    g.vcoef = np.zeros_like(g.vcoef)
    E, small, res, its = oracle_ground.solve(g, tol=1e-10)
    T = grid.kinetic.toarray().real
    r = grid.r.real
    e0 = eigh(T + np.diag(-2 / r), eigvals_only=True)[0]
    assert abs(E - 2 * e0) <= 1e-11
    assert abs(E - (-4.)) < 1e-3          # two hydrogenic Z = 2 electrons


def test_compact_operator_is_the_L0_block():
    grid = make_grid([0., 1., 3., 6., 10.], 5)
    h = Helium(grid, 2, M=0, cutoff_radii=[4., 8.])
    g = GroundS(h)
    x = random_state(g.size, 7)
    full = oracle.tensor_apply(h, g.expand(x))
    assert np.linalg.norm(g.expand(oracle.tensor_apply(g, x)) - full) <= 1e-13 * np.linalg.norm(full)
    # the coupled embedding (l, l, L=0) is the coupled image of the product expansion
    hc = CoupledHelium(h, 2)
    emb = np.zeros(hc.shape, complex, order='F')
    u = x.reshape(g.shape, order='F')
    for l in range(g.nc):
        emb[:, :, hc.basis.index((l, l, 0))] = u[:, :, l]
    assert np.linalg.norm(hc.project(g.expand(x)) - emb.ravel(order='F')) <= 1e-14 * np.linalg.norm(x)
    f = GroundS.from_grid(grid, 2, [4., 8.])
    for name in ('d1', 'd2', 'vptr', 'vc', 'vl', 'vcoef', 'ptr', 'col', 'tv', 'cut'):
        assert np.array_equal(getattr(f, name), getattr(g, name)), name
    assert np.array_equal(f.rad, g.rad) and np.array_equal(f.diag, g.diag)


def test_oracle_reference_run_first_100_steps(pair, recorded, oracle_ground_state):
    hreal, h = pair
    cfg, row = recorded['config'], recorded['history'][0]
    g, (E, small, res, its) = oracle_ground_state
    gs = g.expand(small)
    x = transfer(hreal, h, gs)
    avec = vector_function(config_pulses(cfg))
    dt = recorded['run']['dt']
    for i in range(100):
        x, info = oracle_arnoldi.step(lambda s, y: oracle.tensor_apply(h, y, avec(s), True), x, i * dt, dt,
                                      tol=cfg['krylov_tolerance'], maxdim=cfg['krylov_dimension'])
    got = {**observables(h, x, gs), **info}
    assert row['step'] == 100 and got['dimension'] == row['dimension']
    for key in ('real_region_norm', 'ground_population'):
        assert abs(got[key] - row[key]) <= 1e-12, (key, got[key], row[key])
    assert abs(got['exchange_error'] - row['exchange_error']) <= 1e-10
    assert abs(got['residual_estimate'] - row['residual_estimate']) <= 1e-2 * row['residual_estimate']


def test_reference_run_workload_is_the_recorded_configuration(recorded):
    from laqphysics_amd.workloads import reference_dt_refined_l1_n88
    cfg = reference_dt_refined_l1_n88()
    storage_only = {'max_surface_gib', 'spectrum_energy'}    # raw-flux storage and spectrum post-processing
    assert set(recorded['config']) - set(cfg) == storage_only
    for key, value in cfg.items():
        assert recorded['config'][key] == value, key


def test_configuration_signature_is_the_recorded_one(recorded):
    from laqphysics_amd.run import signature
    assert signature(recorded['config']) == recorded['run']['signature']


def test_run_files_refuse_unsafe_reuse(recorded, tmp_path):
    """An existing output without ``resume`` and a resume under a changed
    configuration are refused before any device work
    (python/run3d.py:122-131)."""
    from laqphysics_amd import run as runmod
    cfg = recorded['config']
    np.savez(tmp_path / 'checkpoint.npz', psi=np.zeros(1), step=0, signature='0' * 64)
    with pytest.raises(FileExistsError, match='use --resume'):
        runmod.run(cfg, tmp_path)
    with pytest.raises(ValueError, match='refusing unsafe resume'):
        runmod.run(cfg, tmp_path, resume=True)
    with pytest.raises(ValueError, match='max_file_bytes'):
        runmod.file_limit(4096)
    with pytest.raises(ValueError, match='above file cap'):
        runmod.capped_npz(tmp_path / 'big.npz', 8192, x=np.zeros(2048))
    runmod.capped_npz(tmp_path / 'small.npz', 8192, x=np.arange(3))
    with np.load(tmp_path / 'small.npz') as d:
        assert list(d['x']) == [0, 1, 2]
