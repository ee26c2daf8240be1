"""Shared builders for the parity tests (small grids from the reference's own tests)."""
import numpy as np

from laqphysics_amd.fedvr import make_grid
from laqphysics_amd.helium3d import Helium


def golden_problem():
    """Helium(make_grid([0,1,2],2), 1, M=None): the module-snapshot problem."""
    return Helium(make_grid([0, 1, 2], 2), 1, M=None)


def golden_field(i, dt):
    return np.array([.04 * np.cos(i * dt), .03 * np.sin(i * dt), .02])


def random_state(size, seed):
    rng = np.random.default_rng(seed)
    x = rng.normal(size=size) + 1j * rng.normal(size=size)
    return x / np.linalg.norm(x)


def relative_error(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))
