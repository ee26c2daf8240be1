"""Named problem definitions (grids, channel bases, pulses, stepping).

Each builder returns a ``Workload`` holding the operator tables, the pulse
list and the stepping parameters.  The grid conventions follow the reference's
``make_grid`` (python/fedvr.py:36-77): an ECS region is a Laguerre-Radau tail
that starts at the last real edge, and ``n = sum(order) + tail`` after the
origin node is removed, so "k elements of order p up to R, irECS beyond R" maps
to real elements on [0, R] plus a tail sized to give the stated point count.
"""
from dataclasses import dataclass, field
from functools import lru_cache

import numpy as np

from .bipolar import BipolarHelium
from .coupled_angular import CoupledHelium
from .fedvr import make_grid
from .helium3d import Helium
from .pulses import Pulse, vector_function


@dataclass
class Workload:
    name: str
    h: object
    pulses: list
    dt: float
    steps: int
    integrator: str = 'arnoldi'
    krylov_dimension: int = 10
    tolerance: float = 1e-10
    surface: float = None
    extra: dict = field(default_factory=dict)

    @property
    def avec(self):
        return vector_function(self.pulses)


def _grid(radius, elements, order, tail, angle):
    return make_grid(np.linspace(0., radius, elements + 1), order, ecs_angle=angle, tail=tail)


@lru_cache(None)
def plain_l2_n85():
    """Coupled l1,l2 <= 2, L <= 2, M = 0 (nc = 11); 8 real order-8 elements on
    [0, 24] plus a 21-node irECS tail at 0.35 rad (n = 85); z-polarised
    sin^2 pulse omega 1.6, E0 0.02, 6 cycles; Arnoldi order 10, dt 0.01,
    600 steps; tSURFF surface at r = 20."""
    grid = _grid(24., 8, 8, 21, .35)
    h = CoupledHelium(Helium(grid, 2, M=0, cutoff_radii=[12., 16.]), 2)
    pulses = [(Pulse(omega=1.6, cycles=6, field=.02), 'z')]
    return Workload('plain_l2_L2_n85', h, pulses, dt=.01, steps=600, krylov_dimension=10, surface=20.)


@lru_cache(None)
def production_l6_L4_n441():
    """Coupled l1,l2 <= 6, L <= 4, M = 0 (nc = 72); 33 real order-12 elements on
    [0, 100] plus a 45-node irECS tail at 0.3 rad (n = 441); XUV pump
    (1.5, 8 cycles, 0.03) and, 200 a.u. after the pump ends, a probe
    (0.114, 4 cycles, 0.01); CF4-Pade with GMRES restart 30, tol 1e-12,
    dt 0.02, 4000 steps; surface at r = 80."""
    grid = _grid(100., 33, 12, 45, .3)
    h = CoupledHelium(Helium(grid, 6, M=0, cutoff_radii=[60., 75.]), 4)
    pump = Pulse(omega=1.5, cycles=8, field=.03)
    probe = Pulse(omega=.114, cycles=4, field=.01, start=pump.duration + 200.)
    return Workload('production_l6_L4_n441', h, [(pump, 'z'), (probe, 'z')], dt=.02, steps=4000,
                    integrator='cf4-pade', krylov_dimension=12, tolerance=1e-12, surface=80.,
                    extra={'gmres_restart': 30})


@lru_cache(None)
def batched_l4_L3_n441():
    """Coupled l1,l2 <= 4, L <= 3, M = 0 (nc = 33) on the n = 441 grid above;
    Arnoldi order 12, dt 0.02."""
    grid = _grid(100., 33, 12, 45, .3)
    h = CoupledHelium(Helium(grid, 4, M=0, cutoff_radii=[60., 75.]), 3)
    pump = Pulse(omega=1.5, cycles=8, field=.03)
    probe = Pulse(omega=.114, cycles=4, field=.01, start=pump.duration + 200.)
    return Workload('batched_l4_L3_n441', h, [(pump, 'z'), (probe, 'z')], dt=.02, steps=2500,
                    krylov_dimension=12, surface=80.)


@lru_cache(None)
def circular_l5_n301():
    """Bipolar l1,l2 <= 5, L <= 5, all M, both parities (nc = 771); 23 real
    order-10 elements on [0, 70] plus a 71-node irECS tail at 0.3 rad
    (n = 301); sigma+ pulse omega 1.6, E0 0.02, 6 cycles; Arnoldi order 12,
    dt 0.01, 3000 steps."""
    grid = _grid(70., 23, 10, 71, .3)
    h = BipolarHelium(grid, 5, 5, M=None, cutoff_radii=[40., 52.], natural_parity=False)
    pulses = [(Pulse(omega=1.6, cycles=6, field=.02), 'sigma+')]
    return Workload('circular_l5_L5_n301', h, pulses, dt=.01, steps=3000, krylov_dimension=12, surface=60.)


ALL = {w.__name__: w for w in (plain_l2_n85, production_l6_L4_n441, batched_l4_L3_n441, circular_l5_n301)}


def initial_state(h):
    """Normalised uncorrelated (1s)^2 product with screened charge 27/16 in the
    (l1, l2) = (0, 0) channel: the starting guess of the reference's ground
    solvers (python/helium3d.py:93-95, python/ground_s.py:64-65), used as a
    deterministic synthetic initial state."""
    r = np.asarray(h.r)
    w = np.asarray(h.grid.weights)
    u = r * np.exp(-1.6875 * r) * np.sqrt(w)
    x = np.zeros(h.shape, complex, order='F')
    basis = getattr(h, 'basis', None)
    if basis is not None:
        c = [k for k, b in enumerate(basis) if b[0] == 0 and b[1] == 0][0]
    else:
        c = h.ch.index(((0, 0), (0, 0)))
    x[:, :, c] = u[:, None] * u[None, :]
    x = x.ravel(order='F')
    return x / np.linalg.norm(x)


def ground_state(h, device='cuda:0', tol=1e-9):
    """The field-free 1S ground state of ``h``'s radial grid and cut-off,
    solved on the device on the grid's real version (python/ground_s.py via
    ``GroundS.from_grid``), placed in the L = 0 channels of ``h``'s basis and
    carried onto the ECS grid by sqrt(w_ecs / w_real) per electron
    (python/run3d.py:134-136).  Returns (ground energy, normalised host vector)."""
    from .ground import GroundS
    real = h.grid.real_version()
    g = GroundS.from_grid(real, h.lmax, h.cutoff_radii)
    E, small, res = g.solve(tol, expand=False, device=device)
    u = np.asarray(small).reshape(g.shape, order='F')
    x = np.zeros(h.shape, complex, order='F')
    basis = getattr(h, 'basis', None)
    for l in range(g.nc):
        if basis is not None and len(basis[0]) == 3:
            x[:, :, basis.index((l, l, 0))] = u[:, :, l]
        elif basis is not None:
            x[:, :, h.index[(l, l, 0, 0)]] = u[:, :, l]
        else:
            for c, ((a, m), (b, mm)) in enumerate(h.ch):
                if a == l and b == l and m == -mm:
                    x[:, :, c] = (-1) ** (l - m) / np.sqrt(2 * l + 1) * u[:, :, l]
    ratio = np.sqrt(h.grid.weights / real.weights)
    x = x * ratio[:, None, None] * ratio[None, :, None]
    x = x.ravel(order='F')
    return E, x / np.linalg.norm(x)


def reference_dt_refined_l1_n88():
    """The reference's recorded 3D run (results/3d_dt_refined/config.json):
    product basis lmax 1, M = 0 (6 channels); 16 real order-4 elements on
    [0, 28] plus a 24-node Laguerre tail at 0.5 rad, alpha 0.8 (n = 88);
    cut-off [12, 16]; z pulse (1.5, 8 cycles, 0.005), velocity gauge; 55 a.u.
    after the pulse; dt 0.03 -> 2952 steps with the surface stride 4; Arnoldi
    dimension 48 at tol 1e-10; ground solve at tol 1e-9; surface at r = 20."""
    return {
        'radial': {'edges': [0, .5, 1, 2, 4, 6, 8, 10, 12, 14, 16, 18, 20, 22, 24, 26, 28], 'order': 4, 'tail': 24,
                   'ecs_angle': .5, 'alpha': .8},
        'lmax': 1, 'M': 0, 'cutoff_radii': [12, 16], 'surface': 20,
        'pulses': [{'pulse': {'omega': 1.5, 'cycles': 8, 'field': .005}, 'polarization': 'z'}],
        'post_time': 55, 'dt': .03, 'surface_stride': 4, 'checkpoint_every': 100, 'krylov_dimension': 48,
        'krylov_tolerance': 1e-10, 'ground_tolerance': 1e-9,
    }


# The reference's production-step benchmark (scripts/benchmark_step.py on
# configs/probe_20260924/production/probe_w014_sigma_minus{,_l6}.json); its
# published seconds per CF4 step on an RTX 6000 Ada
# (results/probe_20260924/validation/production_step_benchmark_l{4,6}.json).
REFERENCE_PRODUCTION_SECONDS = {4: 2.6708806613460183, 6: 5.466443647630513}


def reference_production(lmax=4):
    """Bipolar lmax 4 (375 channels) or 6 (705 channels), total L <= 4, all M,
    both parities; 7 order-12 and 26 order-6 elements on [0, 64] plus a
    24-node tail at 0.5 rad, alpha 0.8 (n = 264); cut-off [32, 40], surface
    44; z pulse (1.2, 48 cycles, 0.08), sigma+ pulse (1.5, 201 cycles, 0.02,
    from 360), sigma- probe (0.14, 48 cycles, 0.02, from 1321.98); CF4-Pade
    at dt 0.06 to 3596.22 (59937 steps), linear tol 1e-10; ground tol 1e-10."""
    if lmax not in (4, 6):
        raise ValueError('the reference production benchmark has lmax 4 or 6')
    return {
        'radial': {'edges': [0, .25, .5, .75, 1, 1.25, 1.5, 2, 2.5, 3, 4, 5, 6, 8, 10, 12, 14, 16, 18, 20, 22, 24, 26,
                             28, 30, 32, 36., 40., 44., 48., 52., 56., 60., 64.],
                   'order': [12] * 7 + [6] * 26, 'tail': 24, 'ecs_angle': .5, 'alpha': .8},
        'lmax': lmax, 'M': None, 'cutoff_radii': [32., 40.], 'surface': 44.,
        'pulses': [{'pulse': {'omega': 1.2, 'cycles': 48, 'field': .08}, 'polarization': 'z'},
                   {'pulse': {'omega': 1.5, 'cycles': 201, 'field': 0.020035791491176656, 'start': 360.},
                    'polarization': 'sigma+'},
                   {'pulse': {'omega': .14, 'cycles': 48, 'field': .02, 'start': 1321.98}, 'polarization': 'sigma-'}],
        'dt': .06, 'surface_stride': 1, 'time_integrator': 'cf4-pade', 'linear_tolerance': 1e-10,
        'ground_tolerance': 1e-10, 'total_Lmax': 4, 'angular_representation': 'bipolar', 'natural_parity': False,
        'end_time': 3596.22,
    }


# the "ionic_channels" of configs/probe_20260924/production/probe_w014_sigma_minus.json
PRODUCTION_IONIC_CHANNELS = [(1, 0, 0), (2, 0, 0), (2, 1, -1), (2, 1, 0), (2, 1, 1), (3, 0, 0), (3, 1, -1), (3, 1, 0),
                             (3, 1, 1), (3, 2, -2), (3, 2, -1), (3, 2, 0), (3, 2, 1), (3, 2, 2), (5, 2, -2), (5, 2, 2),
                             (5, 4, -4), (5, 4, -2), (5, 4, 2), (5, 4, 4)]
# results/probe_20260924/validation/production_step_benchmark_l4.json: one
# backward ionic CF4 step of the 20 channels on 8 CPU threads (IonicAdjoint on
# the CPU, GMRES at tol 1e-12), at the middle of the probe pulse
REFERENCE_IONIC_STEP_SECONDS = 0.3629409035667777


def reference_production_ion(lmax=4):
    """He+ of the reference production configuration: its radial grid
    (n = 264), cut-off [32, 40] and every (l, m) with l <= lmax (25 states for
    lmax 4), as scripts/benchmark_step.py builds it (python/ionic_probe.py:24-29)."""
    from types import SimpleNamespace
    from .fedvr import make_grid
    from .helium3d import cutoff
    from .ionic import Ionic
    config = reference_production(lmax)
    grid = make_grid(**config['radial'])
    cut = cutoff(grid.r.real, *config['cutoff_radii'])
    h = SimpleNamespace(grid=grid, n=len(grid.r), r=grid.r, cut=cut,
                        ch=[((l, m), (0, 0)) for l in range(lmax + 1) for m in range(-l, l + 1)])
    return Ionic(h)
