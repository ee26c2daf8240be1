"""The whole 85-point, 11-channel run (600 midpoint Arnoldi steps from t = 0,
tSURFF at r = 20 every step) through the device driver, from the 1S ground
state of the device ground solver as the configuration specifies, against
the same run through the oracle from the same state: history rows
(inner-region norm, ground population, exchange error, Krylov dimension)
every 100 steps and the channel amplitudes at the end, within 1e-10 (gpu)."""
import numpy as np
import pytest

import oracle
from oracle import arnoldi as oracle_arnoldi
from oracle import surface as osurf
from laqphysics_amd import workloads
from laqphysics_amd.hip_backend import HipHamiltonian
from laqphysics_amd.run import OnlineSurface, observables, propagate
from laqphysics_amd.surface import directions

pytestmark = pytest.mark.gpu


def test_full_plain_run_matches_oracle(device):
    print('CASE full plain_l2_n85 run, 600 steps, tSURFF r=20', flush=True)
    w = workloads.plain_l2_n85()
    h = w.h
    avec = w.avec
    labels = [(1, 0, 0), (2, 1, 0)]          # python/streaming_surface.py:163 default channels
    energy = np.linspace(.2, 1., 9)
    angles = directions(True, 6, 1)
    times = np.arange(w.steps + 1) * w.dt
    E, x0 = workloads.ground_state(h, device)
    assert abs(E + 2.9) < .05          # He 1S on this grid and basis
    op = HipHamiltonian(h, device)
    surf = OnlineSurface(op, w.surface, labels, times, avec, energy, angles)
    state, history = propagate(op, x0, avec, w.dt, w.steps, tol=w.tolerance, maxdim=w.krylov_dimension,
                               surface=surf, ground=x0)
    assert [r['step'] for r in history] == list(range(100, 601, 100))

    base = h.uncoupled
    rcon = osurf.Contractor(base, surf.ion, labels, surf.contractor.unique, surf.contractor.columns, True, h.transform)
    racc = osurf.Volkov(base, rcon.outerstates, len(labels), energy, angles[0], angles[1], avec)
    x = x0.copy()
    rows = []
    racc.add(0, 0., rcon(surf.chis[0], osurf.flux(h, x, avec(0.))))
    for i in range(w.steps):
        x, info = oracle_arnoldi.step(lambda s, y: oracle.tensor_apply(h, y, avec(s), True), x, times[i], w.dt,
                                      tol=w.tolerance, maxdim=w.krylov_dimension)
        racc.add(i + 1, times[i + 1], rcon(surf.chis[i + 1], osurf.flux(h, x, avec(times[i + 1]))))
        if (i + 1) % 100 == 0:
            rows.append({**observables(h, x, x0), **info})
    for got, ref in zip(history, rows):
        assert got['dimension'] == ref['dimension']
        for key in ('real_region_norm', 'ground_population', 'exchange_error'):
            assert abs(got[key] - ref[key]) <= 1e-10, (got['step'], key, got[key], ref[key])
    assert np.linalg.norm(op.host(state) - x) <= 1e-10
    amp, ref_amp = surf.amplitude(), racc.amplitude
    scale = np.max(abs(ref_amp))
    assert scale > 0
    assert np.max(abs(amp - ref_amp)) <= 1e-10 * max(scale, 1.)
    assert np.max(abs(amp - ref_amp)) <= 1e-8 * scale
