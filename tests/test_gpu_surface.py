"""tSURFF on the device (flux -> channel projection -> Volkov accumulation)
against the oracle (gpu)."""
import numpy as np
import pytest

import oracle
from oracle import arnoldi as oracle_arnoldi
from oracle import surface as osurf
from laqphysics_amd import hip_backend, workloads
from laqphysics_amd.bipolar import BipolarHelium
from laqphysics_amd.coupled_angular import CoupledHelium
from laqphysics_amd.fedvr import make_grid
from laqphysics_amd.helium3d import Helium
from laqphysics_amd.hip_backend import HipHamiltonian
from laqphysics_amd.ionic import Ionic, backward_duals
from laqphysics_amd.surface import SurfaceContractor, VolkovAccumulator, directions

pytestmark = pytest.mark.gpu


def _say(*what):
    print('CASE', *what, flush=True)


def _setup(kind):
    grid = make_grid([0, 1, 2, 4, 6, 8], 3, tail=6, ecs_angle=.3)
    if kind == 'coupled':
        base = Helium(grid, 2, M=0, cutoff_radii=[1, 2])
        h = CoupledHelium(base, 2)
        folded = True
        labels = [(1, 0, 0), (2, 1, 0), (2, 0, 0)]
    elif kind == 'product':
        base = h = Helium(grid, 1, M=None, cutoff_radii=[1, 2])
        folded = True
        labels = [(1, 0, 0), (2, 1, 0)]
    else:
        h = BipolarHelium(grid, 1, 2, M=None, cutoff_radii=[1, 2])
        base = h.uncoupled
        folded = False
        labels = [(1, 0, 0), (2, 1, -1), (2, 1, 0), (2, 1, 1)]
    h.prepare_surface(5.)
    if base is not h:
        base.prepare_surface(5.)
    unique = list(dict.fromkeys((n, l, abs(m) if folded else m) for n, l, m in labels))
    columns = [unique.index((n, l, abs(m) if folded else m)) for n, l, m in labels]
    magnetic = {m for n, l, m in unique} if folded else None
    ion = Ionic(base, magnetic)
    return h, base, ion, labels, unique, columns, folded


@pytest.mark.parametrize('kind', ['coupled', 'product', 'bipolar'])
def test_device_contraction_against_oracle(device, kind):
    _say('contract', kind)
    h, base, ion, labels, unique, columns, folded = _setup(kind)
    transform = getattr(h, 'transform', None)
    dev = SurfaceContractor(base, ion, labels, unique, columns, folded, transform, device)
    ref = osurf.Contractor(base, ion, labels, unique, columns, folded, transform)
    rng = np.random.default_rng(7)
    chi = rng.normal(size=(ion.na * base.n, len(unique))) + 1j * rng.normal(size=(ion.na * base.n, len(unique)))
    frame = rng.normal(size=(h.n, len(h.surface_indices), h.nc)) + 1j * rng.normal(size=(h.n, len(h.surface_indices), h.nc))
    np.testing.assert_allclose(dev(chi, frame).cpu().numpy(), ref(chi, frame), rtol=1e-12, atol=1e-12)


def test_device_volkov_against_oracle(device):
    _say('volkov')
    h, base, ion, labels, unique, columns, folded = _setup('coupled')
    con = SurfaceContractor(base, ion, labels, unique, columns, folded, h.transform, device)
    from laqphysics_amd.pulses import Pulse, vector_function
    avec = vector_function([(Pulse(omega=5., cycles=2, field=.15), 'z')])
    times = np.linspace(0, 2 * np.pi * 2 / 5, 17)
    energy = np.linspace(.2, 1.2, 7)
    angles = directions(True, 4, 1)
    acc = VolkovAccumulator(con, energy, angles, avec)
    ref = osurf.Volkov(base, con.outerstates, len(labels), energy, angles[0], angles[1], avec)
    rng = np.random.default_rng(5)
    for i, t in enumerate(times):
        q = rng.normal(size=con.shape) + 1j * rng.normal(size=con.shape)
        acc.add(i, t, q)
        ref.add(i, t, q)
    np.testing.assert_allclose(acc.host_amplitude(), ref.amplitude, rtol=1e-11, atol=1e-12)


def test_device_tsurff_pipeline_plain_l2_n85(device):
    """Propagate 30 steps with a surface sample after each, flux -> contraction ->
    accumulation entirely on the device; compare channel amplitudes with the
    oracle pipeline (C apply, NumPy Arnoldi, NumPy flux/contraction/accumulation)."""
    _say('tsurff pipeline plain_l2_n85')
    w = workloads.plain_l2_n85()
    h = w.h
    base = h.uncoupled
    h.prepare_surface(w.surface)
    base.prepare_surface(w.surface)
    labels = [(1, 0, 0), (2, 1, 0)]
    unique = list(dict.fromkeys((n, l, abs(m)) for n, l, m in labels))
    columns = [unique.index((n, l, abs(m))) for n, l, m in labels]
    ion = Ionic(base, {m for n, l, m in unique})
    avec = w.avec
    steps = 30
    t0 = 200 * w.dt
    times = t0 + np.arange(steps + 1) * w.dt
    chis = backward_duals(ion, unique, times, avec)
    energy = np.linspace(.2, 1., 9)
    angles = directions(True, 6, 1)
    op = HipHamiltonian(h, device)
    con = SurfaceContractor(base, ion, labels, unique, columns, True, h.transform, device)
    acc = VolkovAccumulator(con, energy, angles, avec)
    rcon = osurf.Contractor(base, ion, labels, unique, columns, True, h.transform)
    racc = osurf.Volkov(base, rcon.outerstates, len(labels), energy, angles[0], angles[1], avec)
    x = workloads.initial_state(h)
    state = op.state(x)
    for i in range(steps + 1):
        t = times[i]
        if i:
            x, _ = oracle_arnoldi.step(lambda s, y: oracle.tensor_apply(h, y, avec(s), True), x, times[i - 1], w.dt,
                                       tol=w.tolerance, maxdim=w.krylov_dimension)
            state, _ = hip_backend.step(lambda s, y: op.apply(y, avec(s), True), state, times[i - 1], w.dt,
                                        tol=w.tolerance, maxdim=w.krylov_dimension)
        acc.add(i, t, con(chis[i], op.flux_device(state, avec(t))))
        racc.add(i, t, rcon(chis[i], osurf.flux(h, x, avec(t))))
    scale = np.max(abs(racc.amplitude))
    assert np.max(abs(acc.host_amplitude() - racc.amplitude)) <= 1e-10 * max(scale, 1.)


@pytest.mark.parametrize('kind', ['coupled', 'product', 'bipolar'])
def test_frame_terms_and_flux_adjoint(device, kind):
    """The J part and the unit-field part of a frame add up to the full flux,
    and hex_flux_adjoint is their exact adjoint: <g0, F0 x> + <gd, Fd x> =
    <F^H (g0, gd), x> for random x, g0, gd (the costate sources of the
    adjoint yield rest on it)."""
    import torch
    _say('flux adjoint', kind)
    h = _setup(kind)[0]
    op = HipHamiltonian(h, device)
    rng = np.random.default_rng(31)
    x = op.state(rng.normal(size=h.size) + 1j * rng.normal(size=h.size))
    F0 = op.flux_device(x, (0., 0., 0.), terms=1)
    Fd = op.flux_device(x, (0., 0., 1.), terms=2)
    full = op.flux_device(x, (0., 0., .7))
    torch.testing.assert_close(F0 + .7 * Fd, full, rtol=0, atol=1e-13 * float(full.abs().max()))
    shape = tuple(F0.shape)
    g0 = torch.as_tensor(rng.normal(size=shape) + 1j * rng.normal(size=shape), device=x.device)
    gd = torch.as_tensor(rng.normal(size=shape) + 1j * rng.normal(size=shape), device=x.device)
    u = op.flux_adjoint(g0, gd)
    lhs = complex(torch.vdot(g0.reshape(-1), F0.reshape(-1)) + torch.vdot(gd.reshape(-1), Fd.reshape(-1)))
    rhs = complex(torch.vdot(u, x))
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)
    assert float(Fd.abs().max()) > 0   # the field rows are exercised
