"""Total-L coupled (l1, l2, L, M=0) natural-parity channel basis.

The channels and their order are those of the reference's ``CoupledHelium``
(python/coupled_angular.py:14-63).  Its operator is the product-basis one seen
through the Clebsch-Gordan isometry ``U[(l1 m, l2 -m), (l1 l2 L)] =
<l1 m l2 -m | L 0>``: every angular operator of the product basis
(``helium3d.product_algebra``) becomes ``U^T A U``, and ``helium3d.edge_tables``
turns those matrices into the kernel's edge lists.  An M = 0 basis carries the
z component of the dipole only, so x/y fields are refused
(python/coupled_angular.py:49,62).
"""
import numpy as np

from .angular import clebsch
from .helium3d import OperatorTables, edge_tables, product_algebra


def coupled_channels(lmax, total_Lmax):
    """(l1, l2, L) with |l1 - l2| <= L <= min(l1 + l2, total_Lmax) and natural parity."""
    return [(l1, l2, L) for l1 in range(lmax + 1) for l2 in range(lmax + 1)
            for L in range(abs(l1 - l2), min(l1 + l2, total_Lmax) + 1) if (l1 + l2 + L) % 2 == 0]


def isometry(product, coupled):
    """Dense CG change of basis from product channels ((l1,m1),(l2,m2)) to (l1, l2, L)."""
    U = np.zeros((len(product), len(coupled)))
    column = {c: i for i, c in enumerate(coupled)}
    for a, ((l1, m1), (l2, m2)) in enumerate(product):
        for L in range(abs(l1 - l2), l1 + l2 + 1):
            c = column.get((l1, l2, L))
            if c is not None:
                U[a, c] = clebsch(l1, m1, l2, m2, L, m1 + m2)
    return U


class CoupledHelium(OperatorTables):
    def __init__(self, uncoupled, total_Lmax):
        h = uncoupled
        if any(a[1] + b[1] != 0 for a, b in h.ch):
            raise ValueError('the coupled basis is built on M = 0 product channels only')
        self.uncoupled = h
        self.lmax = max(a[0] for a, _ in h.ch)
        self.total_Lmax = total_Lmax
        self.basis = coupled_channels(self.lmax, total_Lmax)
        self.n, self.nc = h.n, len(self.basis)
        self.shape = (self.n, self.n, self.nc)
        self.size = self.n * self.n * self.nc
        self.grid, self.r, self.cut = h.grid, h.r, h.cut
        self.cutoff_radii = getattr(h, 'cutoff_radii', None)
        self.lam_count = h.lam_count
        self.ptr, self.col, self.tv = h.ptr, h.col, h.tv
        U = isometry(h.ch, self.basis)
        if np.linalg.norm(U.T @ U - np.eye(self.nc)) > 1e-11:
            raise RuntimeError('Clebsch-Gordan change of basis lost orthonormality')
        self.transform = U
        # the one-electron diagonal depends on l1 and l2 only: take it from
        # the product channel (l1 0, l2 0)
        where = {c: i for i, c in enumerate(h.ch)}
        pick = [where[((l1, 0), (l2, 0))] for l1, l2, _ in self.basis]
        self.d1, self.d2 = h.d1[:, pick].copy(), h.d2[:, pick].copy()
        self._diag = None
        _, coulomb, dipole = product_algebra(*h.angular_key)

        def seen(A):
            return U.T @ (A @ U)

        tables = edge_tables([seen(V) for V in coulomb],
                             [(None, None, seen(dipole[0][2])), (None, None, seen(dipole[1][2]))],
                             [c[0] for c in self.basis], [c[1] for c in self.basis])
        self._set_edges(*tables)
        self.exchange = np.array([self.basis.index((l2, l1, L)) for l1, l2, L in self.basis])
        self.exchange_sign = np.array([(-1) ** (l1 + l2 - L) for l1, l2, L in self.basis])

    @property
    def rad(self):
        return self.uncoupled.rad

    def apply(self, x, field=(0., 0., 0.), velocity=False):
        if abs(field[0]) + abs(field[1]) > 1e-15:
            raise ValueError('M=0 coupled basis only supports z polarization')
        return OperatorTables.apply(self, x, field, velocity)

    def expand(self, x):
        """Coupled -> product channels."""
        u = np.asarray(x).reshape(-1, self.nc, order='F')
        return (u @ self.transform.T).ravel(order='F')

    def project(self, x):
        """Product -> coupled channels."""
        u = np.asarray(x).reshape(-1, self.uncoupled.nc, order='F')
        return (u @ self.transform).ravel(order='F')

    def ground(self, tol=1e-10, cache_dir=None, device='cuda:0'):
        """Ground state in the coupled basis (python/coupled_angular.py:73-78)."""
        from .ground import coupled_ground
        return coupled_ground(self, tol, cache_dir, device)
