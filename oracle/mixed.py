"""ORACLE -- mixed-gauge H.psi on the CPU (test infrastructure only).

The two-electron mixed-gauge terms are one-electron operators, so on the
product channels of all (l, m) pairs they are the Kronecker sum of the
one-electron mixed-gauge change Delta = H_ion,mixed(c) - H_ion,0 of
``laqphysics_amd.mixed_gauge.MixedIonic`` (restating python/mixed_gauge.py:119-140):
    H_mixed(c) psi = H_0 psi + (Delta (x) 1 + 1 (x) Delta) psi,
H_0 by the C restatement of the reference kernel with no field -- the
construction of the reference's tests/test_mixed_gauge.py:50-58.  An M = 0
product operator is embedded into the all-M channels (z fields keep the M = 0
channels invariant); a coupled operator goes through its product expansion.
"""
import numpy as np

from . import tensor_apply


def apply(h, x, coeff, inner, outer):
    from laqphysics_amd.helium3d import Helium
    from laqphysics_amd.mixed_gauge import MixedIonic
    base = getattr(h, 'uncoupled', h)
    xp = h.expand(x) if base is not h else np.asarray(x)
    full = Helium(base.grid, base.lmax, M=None, cutoff_radii=base.cutoff_radii)
    index = [full.ch.index(pair) for pair in base.ch]
    n = full.n
    xf = np.zeros(full.shape, complex, order='F')
    xf[:, :, index] = np.asarray(xp).reshape(base.shape, order='F')
    xf = xf.ravel(order='F')
    ion = MixedIonic(full, None, inner, outer)
    na = len(ion.states)
    if [ion.states[i] for i in range(na)] != [a for a, b in full.ch[::na]]:
        raise RuntimeError('product channel order is not (state1, state2) row-major')
    delta = (ion.matrix(coeff) - ion.h0).toarray()
    X = xf.reshape(na, na, n, n).transpose(0, 3, 1, 2).reshape(na * n, na * n)
    mixed = (delta @ X + X @ delta.T).reshape(na, n, na, n).transpose(0, 2, 3, 1).reshape(-1)
    out = tensor_apply(full, xf, (0., 0., 0.), True) + mixed
    out = out.reshape(full.shape, order='F')[:, :, index].ravel(order='F')
    return h.project(out) if base is not h else out
