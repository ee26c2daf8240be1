"""Pin the tSURFF oracle: channel contraction paths agree with each other, and
the online Volkov accumulation equals the closed-form reverse trapezoid (CPU)."""
import numpy as np
import pytest

from oracle import surface as osurf
from laqphysics_amd.coupled_angular import CoupledHelium
from laqphysics_amd.fedvr import make_grid
from laqphysics_amd.helium3d import Helium
from laqphysics_amd.ionic import Ionic
from laqphysics_amd.pulses import Pulse, vector_function
from laqphysics_amd.surface import directions


def _coupled():
    u = Helium(make_grid([0, 1, 2, 4, 6, 8], 3, tail=6, ecs_angle=.3), 2, M=0, cutoff_radii=[1, 2])
    c = CoupledHelium(u, 2)
    u.prepare_surface(5.)
    c.prepare_surface(5.)
    return u, c


def _labels():
    labels = [(1, 0, 0), (2, 1, 0), (2, 0, 0)]
    unique = list(dict.fromkeys((n, l, abs(m)) for n, l, m in labels))
    columns = [unique.index((n, l, abs(m))) for n, l, m in labels]
    return labels, unique, columns


def test_coupled_fast_path_equals_product_path_on_expanded_frame():
    u, c = _coupled()
    labels, unique, columns = _labels()
    ion = Ionic(u, {0})
    rng = np.random.default_rng(424)
    chi = rng.normal(size=(ion.na * u.n, len(unique))) + 1j * rng.normal(size=(ion.na * u.n, len(unique)))
    ns = len(u.surface_indices)
    frame = rng.normal(size=(u.n, ns, c.nc)) + 1j * rng.normal(size=(u.n, ns, c.nc))
    fast = osurf.Contractor(u, ion, labels, unique, columns, True, c.transform)
    plain = osurf.Contractor(u, ion, labels, unique, columns, True, None)
    expanded = np.tensordot(frame, c.transform.T, axes=([-1], [0]))
    np.testing.assert_allclose(fast(chi, frame), plain(chi, expanded), rtol=1e-12, atol=1e-12)


def test_online_accumulation_equals_closed_form_trapezoid():
    u, c = _coupled()
    labels, unique, columns = _labels()
    ion = Ionic(u, {0})
    con = osurf.Contractor(u, ion, labels, unique, columns, True, c.transform)
    avec = vector_function([(Pulse(omega=5., cycles=2, field=.15), 'z')])
    times = np.linspace(0, 2 * np.pi * 2 / 5, 17)
    energy = np.linspace(.2, 1.2, 7)
    theta, phi, _ = directions(True, 4, 1)
    rng = np.random.default_rng(11)
    q = rng.normal(size=(len(times), *con.shape)) + 1j * rng.normal(size=(len(times), *con.shape))
    acc = osurf.Volkov(u, con.outerstates, len(labels), energy, theta, phi, avec)
    for i, t in enumerate(times):
        acc.add(i, t, q[i])
    ref = osurf.integrate(q, u, con.outerstates, times, avec, energy, theta, phi)
    np.testing.assert_allclose(acc.amplitude, ref, rtol=2e-11, atol=2e-11)


def test_volkov_tables_match_closed_form_bessel_and_harmonics():
    """The projection functions sqrt(2/pi) (-i)^l sqrt(w) r j_l(k r) and the
    harmonics Y_lm(theta, phi) of the Volkov accumulation, against the
    textbook closed forms for l <= 2 (pins the SciPy special functions the
    oracle and the device tables share)."""
    u, c = _coupled()
    energy = np.array([.05, .3, 1.1])
    theta, phi, _ = directions(False, 5, 4)
    states = [(0, 0), (1, 0), (1, 1), (2, 0), (2, -1)]
    k, tests, harm, unit = osurf.volkov_tables(u, states, energy, theta, phi)
    r = u.r[u.surface_indices].real
    w = u.grid.weights[u.surface_indices].real
    x = k[:, None] * r[None, :]
    bessel = {0: np.sin(x) / x, 1: np.sin(x) / x ** 2 - np.cos(x) / x,
              2: (3 / x ** 3 - 1 / x) * np.sin(x) - 3 * np.cos(x) / x ** 2}
    ct, st = np.cos(theta), np.sin(theta)
    Y = {(0, 0): np.full_like(theta, 1 / np.sqrt(4 * np.pi)) + 0j,
         (1, 0): np.sqrt(3 / (4 * np.pi)) * ct + 0j,
         (1, 1): -np.sqrt(3 / (8 * np.pi)) * st * np.exp(1j * phi),
         (2, 0): np.sqrt(5 / (16 * np.pi)) * (3 * ct ** 2 - 1) + 0j,
         (2, -1): np.sqrt(15 / (8 * np.pi)) * st * ct * np.exp(-1j * phi)}
    for q, (l, m) in enumerate(states):
        expect = np.sqrt(2 / np.pi) * (-1j) ** l * np.sqrt(w)[None, :] * r[None, :] * bessel[l]
        np.testing.assert_allclose(tests[q], expect, rtol=1e-12, atol=1e-14)
        np.testing.assert_allclose(harm[q], Y[(l, m)], rtol=1e-13, atol=1e-14)
    np.testing.assert_allclose(k, np.sqrt(2 * energy), rtol=1e-15)
    np.testing.assert_allclose(np.linalg.norm(unit, axis=1), 1., rtol=1e-15)


def test_expm_duals_match_cf4_and_compose():
    """Exact-exponential dual propagation (python/surface3d.py:101-110) is the
    exponential midpoint rule: against finely sub-stepped CF4 duals its error
    falls fourfold per halving of the sub-step (second order), and with no
    field one frame of four sub-steps equals one sub-step of the whole frame
    (the exponentials commute)."""
    from laqphysics_amd.fedvr import make_grid
    from laqphysics_amd.helium3d import Helium
    from laqphysics_amd.ionic import Ionic, backward_duals
    from laqphysics_amd.pulses import Pulse, vector_function
    grid = make_grid(np.linspace(0., 20., 6), 6, ecs_angle=.4, tail=10)
    h = Helium(grid, 1, M=0, cutoff_radii=[6., 8.])
    ion = Ionic(h, {0})
    labels = [(1, 0, 0), (2, 1, 0)]
    avec = vector_function([(Pulse(omega=1.5, cycles=2, field=.05), 'z')])
    times = np.arange(11) * .2
    fine = backward_duals(ion, labels, times, avec, substeps=16, propagator='cf4')
    errors = [max(np.linalg.norm(a - b) for a, b in zip(backward_duals(ion, labels, times, avec, substeps=s,
                                                                       propagator='expm'), fine))
              for s in (2, 4, 8)]
    assert errors[-1] < 2e-6
    for coarse, finer in zip(errors, errors[1:]):
        assert 3.6 < coarse / finer < 4.4
    off = lambda t: np.zeros(3)
    one = backward_duals(ion, labels, times, off, substeps=1, propagator='expm')
    four = backward_duals(ion, labels, times, off, substeps=4, propagator='expm')
    assert max(np.linalg.norm(a - b) for a, b in zip(one, four)) < 1e-12
    with pytest.raises(ValueError, match='expm or cf4'):
        backward_duals(ion, labels, times, avec, propagator='rk4')
