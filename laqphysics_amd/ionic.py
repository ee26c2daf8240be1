"""He+ ionic channel states and their backward (adjoint) propagation, on the host.

Mirror of python/surface3d.py:8-48 (``Ionic``), python/implicit.py:111-123
(``SparseCF4``), python/streaming_surface.py:125-146 (``AdjointStepper``,
LU variant) and the exact-exponential dual propagation of
python/surface3d.py:101-110.  These feed the tSURFF channel projection with the ionic duals
chi_c(t) = U_ion(T, t)^dagger phi_c; the reference runs them on the CPU with
sparse LU, and so does this module.  ``backward_duals(..., device=...)``
hands the CF4 propagation to the device (``ionic_adjoint.DeviceDuals``).
"""
import numpy as np
from scipy.linalg import expm
from scipy.sparse import bmat, block_diag, csr_matrix, diags, eye
from scipy.sparse.linalg import expm_multiply, splu

from .angular import cartesian, radial_hydrogen

SQRT3 = np.sqrt(3.)
C1 = .5 - SQRT3 / 6
C2 = .5 + SQRT3 / 6
A1 = (3 - 2 * SQRT3) / 12
A2 = (3 + 2 * SQRT3) / 12


class Ionic:
    """One-electron He+ Hamiltonian in the inner channels (python/surface3d.py:8-48)."""

    def __init__(self, h, magnetic_numbers=None):
        self.h = h
        self.states = sorted({a for a, b in h.ch if magnetic_numbers is None or a[1] in magnetic_numbers})
        self.n = h.n
        self.na = len(self.states)
        self.index = {a: i for i, a in enumerate(self.states)}
        T = h.grid.kinetic.tocsr()
        r = h.r
        self.h0 = block_diag([T + diags(-2 * h.cut / r + l * (l + 1) / (2 * r * r)) for l, m in self.states],
                             format='csr')
        blocks = [[[csr_matrix((self.n, self.n), dtype=complex) if a == b else None for b in range(self.na)]
                   for a in range(self.na)] for axis in range(3)]
        base = T @ diags(r) - diags(r) @ T
        for a, (l, m) in enumerate(self.states):
            for b, (ll, mm) in enumerate(self.states):
                coef = cartesian((l, m), (ll, mm))
                if np.max(abs(coef)) < 1e-14:
                    continue
                pm = 1j * (base + diags((l * (l + 1) - ll * (ll + 1)) / (2 * r)))
                for axis in range(3):
                    if abs(coef[axis]) > 1e-14:
                        blocks[axis][a][b] = coef[axis] * pm
        self.P = [bmat(axis, format='csr') for axis in blocks]
        self.T = T
        self.identity = eye(self.n * self.na, format='csr')

    def final_states(self, labels):
        out = []
        for n, l, m in labels:
            z = np.zeros((self.na, self.n), complex)
            z[self.index[(l, m)]] = self.h.r * radial_hydrogen(n, l, self.h.r) * np.sqrt(self.h.grid.weights)
            out.append(z.ravel())
        return np.array(out).T

    def dual_final_states(self, labels):
        # the analytic continuation pairs without conjugating the radial argument
        return self.final_states(labels).conj()

    def matrix(self, A):
        H = self.h0 + self.identity * np.dot(A, A) / 2
        for a, P in zip(A, self.P):
            if abs(a) > 0:
                H = H + a * P
        return H.tocsr()


class SparseCF4:
    """CF4 with sparse-LU Pade factors for small systems; negative dt allowed
    (python/implicit.py:111-123)."""

    def __init__(self, matrix_at, key_at=None):
        self.matrix_at = matrix_at
        self.key_at = key_at
        self.key = None
        self.factors = None

    def step(self, x, t, dt):
        key = (dt, self.key_at(t + C1 * dt), self.key_at(t + C2 * dt)) if self.key_at else None
        if key is None or key != self.key or self.factors is None:
            h1 = self.matrix_at(t + C1 * dt)
            h2 = self.matrix_at(t + C2 * dt)
            I = eye(h1.shape[0], format='csc')
            self.factors = []
            for H in (A2 * h1 + A1 * h2, A1 * h1 + A2 * h2):
                for root in (3 + 1j * SQRT3, 3 - 1j * SQRT3):
                    alpha = dt / root
                    # minimum-degree ordering on A^T + A: the tail's dense block and
                    # the l-coupling make COLAMD's fill five times larger here
                    self.factors.append(splu((I + 1j * alpha * H).tocsc(), permc_spec='MMD_AT_PLUS_A'))
            self.key = key
        for lu in self.factors:
            x = 2 * lu.solve(x) - x
        return x


def backward_duals(ion, unique, times, avec, substeps=1, propagator='cf4', progress=None, device=None, block=1,
                   folder=None, configuration=''):
    """chi(t_i) for every frame time, by backward adjoint propagation from the
    final time over ``substeps`` sub-steps per frame interval.

    ``device`` (a cuda device): the CF4 propagation runs there
    (python/surface_projection.py:29-33, ionic_device), with endpoints every
    ``block`` frames and block replay, kept in ``folder``'s preparation files
    when one is given; the result is indexable like the list.

    ``propagator='cf4'``: sparse-LU CF4 (python/streaming_surface.py:198-214
    with ionic_block_frames = 1); ``'expm'``: the exact exponential
    chi <- exp(i d H(A(t_i - (j + 1/2) d))^dagger) chi of each sub-step
    (python/surface3d.py:101-110): a dense matrix exponential for systems of
    at most 400 unknowns, else its action on the columns.
    ``progress(done, total)`` is called every 512 frames."""
    if propagator not in ('cf4', 'expm'):
        raise ValueError('ionic propagator must be expm or cf4')
    if device is not None:
        if propagator != 'cf4':
            raise ValueError('accelerated ionic propagation requires cf4')
        from .ionic_adjoint import DeviceDuals
        return DeviceDuals(ion, unique, times, avec, substeps, block, device, progress=progress, folder=folder,
                           configuration=configuration)
    chi = ion.dual_final_states(unique)
    out = [None] * len(times)
    out[-1] = chi.copy()
    if len(times) < 2:
        return out
    sample_dt = float(times[1] - times[0])
    d = sample_dt / substeps
    if propagator == 'cf4':
        stepper = SparseCF4(lambda t: ion.matrix(avec(t)).conj().T, key_at=lambda t: tuple(avec(t)))
    # small systems: dense exponentials, cached while the field repeats (after
    # the pulse); larger ones: the action of the exponential on the dual
    # columns (scipy expm_multiply), as the reference computes it
    dense = ion.na * ion.n <= 400
    cache = {}
    for i in range(len(times) - 1, 0, -1):
        for sub in range(substeps):
            if propagator == 'cf4':
                chi = stepper.step(chi, times[i] - sub * d, -d)
            elif dense:
                A = avec(times[i] - (sub + .5) * d)
                key = tuple(A)
                U = cache.get(key)
                if U is None:
                    U = expm(1j * d * ion.matrix(A).conj().T.toarray())
                    if len(cache) > 64:
                        cache.clear()
                    cache[key] = U
                chi = U @ chi
            else:
                A = avec(times[i] - (sub + .5) * d)
                if not np.any(A):
                    # field-free sub-steps share one operator: its exponential once
                    if 'free' not in cache:
                        cache['free'] = expm(1j * d * ion.matrix(A).conj().T.toarray())
                    chi = cache['free'] @ chi
                else:
                    op = csr_matrix(1j * d * ion.matrix(A).conj().T)
                    chi = expm_multiply(op, chi, traceA=op.diagonal().sum())
        out[i - 1] = chi.copy()
        done = len(times) - i
        if progress is not None and done % 512 == 0:
            progress(done, len(times) - 1)
    return out
