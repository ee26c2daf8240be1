"""ORACLE -- CPU Arnoldi exponential (test infrastructure only).

Restates python/propagate.py:6-50: two passes of classical Gram-Schmidt via
BLAS zgemv (conjugate transpose), h_{j+1,j} = ||v||, happy breakdown below
1e-13, convergence checked at j >= 3 when (j+1) % 4 == 0 or at maxdim with
c = expm(-i dt H_m)[:, 0] beta and err = |dt h_{m+1,m} c_m|; on failure the
step is split into two half steps with tol/2, to depth 8.
"""
import numpy as np
from scipy.linalg import expm
from scipy.linalg.blas import zgemv


def exponential_action(apply, x, dt, tol=1e-11, maxdim=60, trace=None):
    beta = np.linalg.norm(x)
    if beta == 0:
        return x.copy(), {'dimension': 0, 'residual_estimate': 0.}
    Q = np.empty((len(x), maxdim + 1), complex, order='F')
    H = np.zeros((maxdim + 1, maxdim), complex)
    Q[:, 0] = x / beta
    err = float('nan')
    for j in range(maxdim):
        v = apply(Q[:, j])
        if trace is not None:
            trace(f'Hq{j:03d}', v.copy())
        for _ in range(2):
            z = zgemv(1., Q[:, :j + 1], v, trans=2)
            H[:j + 1, j] += z
            v = zgemv(-1., Q[:, :j + 1], z, beta=1., y=v, overwrite_y=1)
        H[j + 1, j] = np.linalg.norm(v)
        happy = abs(H[j + 1, j]) < 1e-13
        if happy or j >= 3 and ((j + 1) % 4 == 0 or j + 1 == maxdim):
            m = j + 1
            c = expm(-1j * dt * H[:m, :m])[:, 0] * beta
            err = abs(dt * H[m, m - 1] * c[-1])
            if happy or err < tol:
                out = Q[:, :m] @ c
                if trace is not None:
                    trace('arnoldi_Q', Q[:, :m].copy())
                    trace('arnoldi_H', H[:m, :m].copy())
                    trace('output', out.copy())
                return out, {'dimension': m, 'residual_estimate': float(err)}
        if j + 1 < maxdim:
            Q[:, j + 1] = v / H[j + 1, j]
    raise RuntimeError(f'Arnoldi did not converge: dim={maxdim}, dt={dt}, estimate={err:.3e}; reduce dt')


def step(apply_at, x, t, dt, tol=1e-11, maxdim=60, depth=0, trace=None):
    try:
        return exponential_action(lambda y: apply_at(t + dt / 2, y), x, dt, tol, maxdim, trace)
    except RuntimeError:
        if depth >= 8:
            raise
        y, a = step(apply_at, x, t, dt / 2, tol / 2, maxdim, depth + 1, trace)
        y, b = step(apply_at, y, t + dt / 2, dt / 2, tol / 2, maxdim, depth + 1, trace)
        return y, {'dimension': max(a['dimension'], b['dimension']),
                   'residual_estimate': a['residual_estimate'] + b['residual_estimate'], 'subdivided': True}


def exponential_action_fixed(apply, x, dt, m):
    """The same Arnoldi/CGS2 exponential run to exactly ``m`` columns (or an
    earlier happy breakdown) without the convergence test: the smooth discrete
    map that finite differences of a replayed propagation are taken of."""
    beta = np.linalg.norm(x)
    if beta == 0:
        return x.copy()
    Q = np.empty((len(x), m + 1), complex, order='F')
    H = np.zeros((m + 1, m), complex)
    Q[:, 0] = x / beta
    for j in range(m):
        v = apply(Q[:, j])
        for _ in range(2):
            z = zgemv(1., Q[:, :j + 1], v, trans=2)
            H[:j + 1, j] += z
            v = zgemv(-1., Q[:, :j + 1], z, beta=1., y=v, overwrite_y=1)
        H[j + 1, j] = np.linalg.norm(v)
        if abs(H[j + 1, j]) < 1e-13 or j + 1 == m:
            k = j + 1
            return Q[:, :k] @ (expm(-1j * dt * H[:k, :k])[:, 0] * beta)
        Q[:, j + 1] = v / H[j + 1, j]
