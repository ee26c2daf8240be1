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
from scipy.sparse import block_diag, csr_matrix, diags, eye, kron
from scipy.sparse.linalg import expm_multiply, splu

from .angular import cartesian, radial_hydrogen

SQRT3 = np.sqrt(3.)
C1 = .5 - SQRT3 / 6
C2 = .5 + SQRT3 / 6
A1 = (3 - 2 * SQRT3) / 12
A2 = (3 + 2 * SQRT3) / 12


def channel_dipole(states, threshold=1e-14):
    """Cartesian components of C^1 between one-electron states (l, m): three
    (na, na) matrices, entries at or below ``threshold`` dropped."""
    comps = np.array([[cartesian(a, b) for b in states] for a in states]).transpose(2, 0, 1)
    comps[np.abs(comps) <= threshold] = 0.0
    return [csr_matrix(c) for c in comps]


class Ionic:
    """He+ (one electron, Z = 2) in the inner channels (l, m) of a two-electron
    basis: the operator of the reference's ``Ionic`` (python/surface3d.py:8-48).

    With ``R = diag(r)``, ``B = [T, R]`` and ``Lam[a, b] = l_a(l_a+1) - l_b(l_b+1)``
    the field-free part and the velocity-gauge coupling of component ``x`` are

        h0  = sum_a |a><a| (x) (T + diag(-2 cut / r + l_a(l_a+1) / 2r^2)),
        P_x = i [C_x (x) B + (C_x o Lam) (x) diag(1 / 2r)],

    and ``matrix(A) = h0 + |A|^2/2 + sum_x A_x P_x``.
    """

    def __init__(self, h, magnetic_numbers=None):
        self.h = h
        inner = {a for a, _ in h.ch}
        self.states = sorted(a for a in inner if magnetic_numbers is None or a[1] in magnetic_numbers)
        self.index = {a: i for i, a in enumerate(self.states)}
        self.n, self.na = h.n, len(self.states)
        r = np.asarray(h.r)
        T = h.grid.kinetic.tocsr()
        self.T = T
        ls = np.array([l for l, _ in self.states])
        self.h0 = block_diag([T + diags(-2 * h.cut / r + l * (l + 1) / (2 * r * r)) for l in ls], format='csr')
        lam = (ls * (ls + 1))[:, None] - (ls * (ls + 1))[None, :]
        B = csr_matrix(T @ diags(r) - diags(r) @ T)
        half = diags(1.0 / (2 * r))
        self.P = [csr_matrix(1j * (kron(Cx, B) + kron(Cx.multiply(lam), half))) for Cx in channel_dipole(self.states)]
        self.identity = eye(self.n * self.na, format='csr')

    def final_states(self, labels):
        """Columns: the hydrogenic He+ states (n, l, m) on the ECS contour, in
        DVR normalisation (r R_nl(r) sqrt(w)), one per label."""
        radial = self.h.r * np.sqrt(self.h.grid.weights)
        out = np.zeros((self.na * self.n, len(labels)), complex)
        for col, (n, l, m) in enumerate(labels):
            at = self.index[(l, m)] * self.n
            out[at:at + self.n, col] = radial * radial_hydrogen(n, l, self.h.r)
        return out

    def dual_final_states(self, labels):
        """Start of the backward (adjoint) propagation.  The radial functions are
        continued analytically onto the contour and paired without conjugating
        the complex radius (b = phi_theta^T psi_theta), so the dual columns are
        the complex conjugates of the kets."""
        return self.final_states(labels).conj()

    def matrix(self, A):
        A = np.asarray(A, float)
        H = self.h0 + (0.5 * float(A @ A)) * self.identity
        for a, P in zip(A, self.P):
            if a:
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
