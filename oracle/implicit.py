"""ORACLE -- CF4 / [2/2] Pade / GMRES on the CPU (test infrastructure only).

NumPy restatement of python/implicit.py:13-109 on host arrays:
* separable ionic resolvent: per l, eig of T + diag(-2 cut/r + l(l+1)/2r^2),
  applied per channel as S2 [(I2 u I1^T)/(1 + i alpha (E + shift))] S1^T on the
  torch-layout slab u[c] = psi(:, :, c)^T (python/implicit.py:18-39);
* GMRES: x0 = b, restart cadence, CGS2, host least squares every 4th (2nd if
  flexible) iteration, acceptance on the true residual (python/implicit.py:46-74);
* Pade: both roots 3 +- i sqrt3, x <- 2 s - x (python/implicit.py:77-89);
* CF4: fields at t + C1 dt, t + C2 dt, weights (A2, A1) then (A1, A2), scalar
  A^2 and reference-energy phase, recursive dt/2 on GMRES failure
  (python/implicit.py:92-109).
"""
import numpy as np
from scipy.linalg import eig, inv

import oracle

SQRT3 = np.sqrt(3.)
C1 = .5 - SQRT3 / 6
C2 = .5 + SQRT3 / 6
A1 = (3 - 2 * SQRT3) / 12
A2 = (3 + 2 * SQRT3) / 12


class OracleEngine:
    """Engine protocol (``h``, ``apply``, ``diagonal``) on the C restatement."""

    def __init__(self, h):
        self.h = h

    def apply(self, y, field, velocity=True):
        return oracle.tensor_apply(self.h, y, field, velocity)

    @property
    def diagonal(self):
        h = self.h
        d = np.asarray(h.diag).copy()
        td = h.grid.kinetic.diagonal()
        d += td[:, None, None] + td[None, :, None]
        for a in range(h.nc):
            for k in range(h.vptr[a], h.vptr[a + 1]):
                if h.vc[k] == a:
                    d[:, :, a] += h.vcoef[k] * h.rad[:, :, h.vl[k]]
        return d.ravel(order='F')


class SeparablePreconditioner:
    def __init__(self, engine):
        h = engine.h
        self.h = h
        labels = [tuple(b[:2]) for b in h.basis] if hasattr(h, 'basis') else [(a[0], b[0]) for a, b in h.ch]
        maxl = max(max(p) for p in labels)
        T = h.grid.kinetic.toarray()
        r = h.r
        values, vectors, inverses = [], [], []
        for l in range(maxl + 1):
            e, S = eig(T + np.diag(-2 * h.cut / r + l * (l + 1) / (2 * r * r)))
            values.append(e)
            vectors.append(S)
            inverses.append(inv(S))
        self.S1 = [vectors[a] for a, b in labels]
        self.S2 = [vectors[b] for a, b in labels]
        self.I1 = [inverses[a] for a, b in labels]
        self.I2 = [inverses[b] for a, b in labels]
        self.energy = [values[b][:, None] + values[a][None, :] for a, b in labels]
        self.flexible = False

    def apply(self, x, alpha, shift):
        h = self.h
        u = np.asarray(x).reshape(h.nc, h.n, h.n)   # u[c, j, i]
        out = np.empty_like(u)
        for c in range(h.nc):
            z = self.I2[c] @ u[c] @ self.I1[c].T
            z = z / (1 + 1j * alpha * (self.energy[c] + shift))
            out[c] = self.S2[c] @ z @ self.S1[c].T
        return out.reshape(-1)


def gmres(apply, b, preconditioner, x0=None, tol=1e-11, restart=32, max_restarts=10, flexible=False):
    M = preconditioner if callable(preconditioner) else (lambda y: preconditioner * y)
    x = b.copy() if x0 is None else x0.copy()
    bnorm = np.linalg.norm(b)
    if bnorm == 0:
        return np.zeros_like(b), {'iterations': 0, 'relative_residual': 0.}
    iterations = 0
    true = float('nan')
    for cycle in range(max_restarts):
        r = b - apply(x)
        beta = np.linalg.norm(r)
        if beta <= tol * bnorm:
            return x, {'iterations': iterations, 'relative_residual': beta / bnorm}
        Q = np.empty((b.size, restart + 1), complex, order='F')
        Z = np.empty((b.size, restart), complex, order='F') if flexible else None
        H = np.zeros((restart + 1, restart), complex)
        Q[:, 0] = r / beta
        for j in range(restart):
            pre = M(Q[:, j].copy())
            if flexible:
                Z[:, j] = pre
            v = apply(pre)
            iterations += 1
            for _ in range(2):
                z = Q[:, :j + 1].conj().T @ v
                H[:j + 1, j] += z
                v = v - Q[:, :j + 1] @ z
            norm = np.linalg.norm(v)
            H[j + 1, j] = norm
            if norm < 1e-14 or (j + 1) % (2 if flexible else 4) == 0 or j + 1 == restart:
                m = j + 1
                small = H[:m + 1, :m]
                rhs = np.zeros(m + 1, complex)
                rhs[0] = beta
                y = np.linalg.lstsq(small, rhs, rcond=None)[0]
                estimate = np.linalg.norm(rhs - small @ y)
                if estimate < tol * bnorm or norm < 1e-14 or m == restart:
                    candidate = x + (Z[:, :m] @ y if flexible else M(Q[:, :m] @ y))
                    true = np.linalg.norm(b - apply(candidate)) / bnorm
                    if true <= tol:
                        return candidate, {'iterations': iterations, 'relative_residual': true}
                    if norm < 1e-14 or m == restart:
                        x = candidate
                        break
            Q[:, j + 1] = v / norm
    raise RuntimeError(f'GMRES did not converge: residual={true:.3e}, iterations={iterations}')


def pade_step(engine, x, field, dt, reference_energy=0., tol=1e-11, restart=None):
    def K(y):
        return engine.apply(y, field, True) - reference_energy * y
    a2 = getattr(engine, 'a2_coefficient', 1.)
    a2sum = a2 * sum(float(a) ** 2 for a in field)
    iterations = 0
    residual = 0.
    separable = getattr(engine, 'separable', None)
    for root in (3 + 1j * SQRT3, 3 - 1j * SQRT3):
        alpha = dt / root

        def A(y, alpha=alpha):
            return y + 1j * alpha * K(y)
        if separable is not None:
            def M(y, alpha=alpha):
                return separable.apply(y, alpha, a2sum - reference_energy)
        else:
            M = 1 / (1 + 1j * alpha * (engine.diagonal + a2sum - reference_energy))
        flexible = getattr(separable, 'flexible', False)
        nrestart = restart if restart is not None else (16 if flexible else 32)
        s, info = gmres(A, x, M, tol=tol, flexible=flexible, restart=nrestart)
        x = 2 * s - x
        iterations += info['iterations']
        residual = max(residual, info['relative_residual'])
    return x, {'linear_iterations': iterations, 'linear_residual': residual}


def cf4_step(engine, x, avec, t, dt, reference_energy=0., tol=1e-11, depth=0, restart=None):
    fields = np.array([avec(t + C1 * dt), avec(t + C2 * dt)], float)
    iterations = 0
    residual = 0.
    original = x
    try:
        for w in ((A2, A1), (A1, A2)):
            effective = 2 * np.einsum('a,aj->j', w, fields)
            correction = getattr(engine, 'a2_coefficient', 1.) * (np.dot(w, np.sum(fields ** 2, axis=1))
                                                                   - .5 * np.dot(effective, effective))
            x, info = pade_step(engine, x, effective, dt / 2, reference_energy, tol, restart)
            x = x * np.exp(-1j * dt * (correction + reference_energy / 2))
            iterations += info['linear_iterations']
            residual = max(residual, info['linear_residual'])
        return x, {'linear_iterations': iterations, 'linear_residual': residual, 'dimension': 0,
                   'residual_estimate': None}
    except RuntimeError as error:
        if 'GMRES did not converge' not in str(error) or depth >= 6:
            raise
        x, a = cf4_step(engine, original, avec, t, dt / 2, reference_energy, tol / 2, depth + 1, restart)
        x, b = cf4_step(engine, x, avec, t + dt / 2, dt / 2, reference_energy, tol / 2, depth + 1, restart)
        return x, {'linear_iterations': a['linear_iterations'] + b['linear_iterations'],
                   'linear_residual': max(a['linear_residual'], b['linear_residual']),
                   'dimension': 0, 'residual_estimate': None, 'subdivided': True}
