"""ORACLE -- independent Kronecker assembly of H (test infrastructure only).

Restates ``Helium.reference_sparse`` (python/helium3d.py:105-123): the full
sparse matrix for SMALL grids, built from Kronecker products of the radial
kinetic matrix, the diagonal, multipole diagonals and the dipole operators,
without any of the edge-walking logic of the kernels.  F-order flat index
``i + n j + n^2 c`` means electron 1 acts through ``kron(I, T)``.
"""
import numpy as np
from scipy.sparse import csr_matrix, diags, eye, kron


def reference_sparse(h, field=(0., 0., 0.), velocity=False):
    n = h.n
    T = h.grid.kinetic
    I = eye(n, format='csr')
    radial = kron(I, T) + kron(T, I)
    H = kron(eye(h.nc), radial) + diags(np.asarray(h.diag).ravel(order='F'))
    rad = h.rad
    for a in range(h.nc):
        for k in range(h.vptr[a], h.vptr[a + 1]):
            A = csr_matrix(([h.vcoef[k]], ([a], [h.vc[k]])), shape=(h.nc, h.nc))
            H = H + kron(A, diags(rad[:, :, h.vl[k]].ravel(order='F')))
        for k in range(h.dptr[a], h.dptr[a + 1]):
            coef = np.dot(field, h.dcoef[:, k])
            if abs(coef) < 1e-15:
                continue
            R = diags(h.r)
            if velocity:
                R = 1j * (T @ R - R @ T + diags(h.ldiff[k] / (2 * h.r)))
            op = kron(I, R) if h.de[k] == 1 else kron(R, I)
            A = csr_matrix(([coef], ([a], [h.dc[k]])), shape=(h.nc, h.nc))
            H = H + kron(A, op)
    if velocity:
        H = H + eye(h.size) * np.dot(field, field)
    return H.tocsr()
