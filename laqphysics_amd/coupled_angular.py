"""Total-L coupled (l1, l2, L, M=0) natural-parity channel basis.

Mirror of python/coupled_angular.py:13-82.  The coupled tables are an exact
Clebsch-Gordan change of basis ``U^T A U`` of the product-basis angular
matrices; only z polarisation is representable (python/coupled_angular.py:49,62).
"""
import numpy as np
from sympy.physics.wigner import wigner_3j

from .helium3d import OperatorTables


def edge_matrices(h):
    """Dense angular matrices of a product-basis operator.

    Returns ``A[lam, a, b]`` (Coulomb multipole factors) and ``Cz[e, a, b]``
    (z dipole component for electron e+1), summing duplicate edges.
    """
    rows_v = np.repeat(np.arange(h.nc), np.diff(h.vptr))
    A = np.zeros((h.lam_count, h.nc, h.nc), complex)
    np.add.at(A, (h.vl, rows_v, h.vc), h.vcoef)
    rows_d = np.repeat(np.arange(h.nc), np.diff(h.dptr))
    Cz = np.zeros((2, h.nc, h.nc), complex)
    np.add.at(Cz, (h.de - 1, rows_d, h.dc), h.dcoef[2])
    return A, Cz


class CoupledHelium(OperatorTables):
    def __init__(self, uncoupled, total_Lmax):
        h = uncoupled
        self.uncoupled = h
        if any(a[1] + b[1] != 0 for a, b in h.ch):
            raise ValueError('coupled backend currently requires M=0')
        self.lmax = max(a[0] for a, b in h.ch)
        self.total_Lmax = total_Lmax
        self.basis = [(l, j, L) for l in range(self.lmax + 1) for j in range(self.lmax + 1)
                      for L in range(abs(l - j), min(l + j, total_Lmax) + 1) if (l + j - L) % 2 == 0]
        self.n = h.n
        self.nc = len(self.basis)
        self.shape = (self.n, self.n, self.nc)
        self.size = int(np.prod(self.shape))
        self.grid = h.grid
        self.r = h.r
        self.cut = h.cut
        self.cutoff_radii = getattr(h, 'cutoff_radii', None)
        self.lam_count = h.lam_count
        for name in ('ptr', 'col', 'tv'):
            setattr(self, name, getattr(h, name))
        U = np.zeros((h.nc, self.nc))
        for c, (l, j, L) in enumerate(self.basis):
            for a, ((ll, m), (jj, n)) in enumerate(h.ch):
                if l == ll and j == jj:
                    U[a, c] = (-1) ** (l - j) * np.sqrt(2 * L + 1) * float(wigner_3j(l, j, L, m, n, 0))
        self.transform = U
        if np.linalg.norm(U.T @ U - np.eye(self.nc)) > 1e-11:
            raise RuntimeError('CG transformation is not isometric')
        source = [h.ch.index(((l, 0), (j, 0))) for l, j, L in self.basis]
        self.d1 = h.d1[:, source].copy()
        self.d2 = h.d2[:, source].copy()
        self._diag = None
        A, Cz = edge_matrices(h)
        V = [U.T @ A[lam] @ U for lam in range(h.lam_count)]
        Cc = [U.T @ Cz[e] @ U for e in range(2)]
        vp, vc, vl, vv, dp, dc, de, dv, dl = [0], [], [], [], [0], [], [], [], []
        for a, (l, j, L) in enumerate(self.basis):
            for b, (ll, jj, LL) in enumerate(self.basis):
                for lam in range(h.lam_count):
                    if abs(V[lam][a, b]) > 1e-13:
                        vc.append(b)
                        vl.append(lam)
                        vv.append(V[lam][a, b])
                for e in range(2):
                    if abs(Cc[e][a, b]) > 1e-13:
                        dc.append(b)
                        de.append(e + 1)
                        dv.append([0, 0, Cc[e][a, b]])
                        bra, ket = (l, j)[e], (ll, jj)[e]
                        dl.append(bra * (bra + 1) - ket * (ket + 1))
            vp.append(len(vc))
            dp.append(len(dc))
        self._set_edges(vp, vc, vl, vv, dp, dc, de, dv, dl)
        self.exchange = np.array([self.basis.index((j, l, L)) for l, j, L in self.basis])
        self.exchange_sign = np.array([(-1) ** (l + j - L) for l, j, L in self.basis])

    @property
    def rad(self):
        return self.uncoupled.rad

    def apply(self, x, field=(0., 0., 0.), velocity=False):
        if abs(field[0]) + abs(field[1]) > 1e-15:
            raise ValueError('M=0 coupled basis only supports z polarization')
        return OperatorTables.apply(self, x, field, velocity)

    def expand(self, x):
        u = np.asarray(x).reshape(self.n * self.n, self.nc, order='F')
        return (u @ self.transform.T).ravel(order='F')

    def project(self, x):
        u = np.asarray(x).reshape(self.n * self.n, self.uncoupled.nc, order='F')
        return (u @ self.transform).ravel(order='F')

    def ground(self, tol=1e-10, cache_dir=None, device='cuda:0'):
        """Ground state in the coupled basis (python/coupled_angular.py:73-78)."""
        from .ground import coupled_ground
        return coupled_ground(self, tol, cache_dir, device)
