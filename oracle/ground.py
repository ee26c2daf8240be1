"""ORACLE -- CPU ground-state solve (test infrastructure only).

Restates python/ground_s.py:55-89: SciPy's LOBPCG (largest=False, maxiter
250) on the real part of the compact L = 0 operator, applied by the C
restatement of the reference kernel, started from the screened (1s)^2 product
in the l = 0 slab, preconditioned by the exact inverse of H1 (+) H2 - shift in
dense per-l eigenbases (numpy/SciPy ``eigh``), shift = 2 min e_0 - 1/2.
Returns (E, compact state, residual ||H x - E x||) like ``GroundS.solve`` with
``expand=False``; the state's sign is fixed so that its largest-magnitude entry
is positive (the eigenvector's sign is arbitrary in the reference).

The compact operator tables are the product code's ``GroundS`` tables.
"""
import numpy as np
from scipy.linalg import eigh
from scipy.sparse.linalg import LinearOperator, lobpcg

from . import tensor_apply


def preconditioner(g):
    """The exact separable resolvent of python/ground_s.py:67-76."""
    T = g.grid.kinetic.toarray().real
    r = np.asarray(g.r).real
    spectra, vectors = [], []
    for l in range(g.nc):
        e, U = eigh(T + np.diag(-2 * g.cut / r + l * (l + 1) / (2 * r * r)))
        spectra.append(e)
        vectors.append(U)
    shift = 2 * min(spectra[0]) - .5

    def precondition(y):
        y = np.asarray(y).reshape(g.shape, order='F')
        out = np.empty(g.shape)
        for l in range(g.nc):
            U = vectors[l]
            e = spectra[l]
            z = U.T @ y[:, :, l] @ U
            out[:, :, l] = U @ (z / (e[:, None] + e[None, :] - shift)) @ U.T
        return out.ravel(order='F')
    return precondition, shift


def solve(g, tol=1e-10, maxiter=250):
    op = LinearOperator((g.size, g.size), matvec=lambda v: tensor_apply(g, v).real, dtype=float)
    precondition, _ = preconditioner(g)
    M = LinearOperator((g.size, g.size), matvec=precondition, dtype=float)
    x = g.starting_guess()
    E, v, history = lobpcg(op, x.reshape(-1, 1), M=M, largest=False, tol=tol, maxiter=maxiter,
                           retResidualNormsHistory=True)
    small = v[:, 0]
    k = int(np.argmax(abs(small)))
    if small[k] < 0:
        small = -small
    small = small.astype(complex)
    res = float(np.linalg.norm(tensor_apply(g, small) - E[0] * small))
    return float(E[0]), small, res, len(history)
