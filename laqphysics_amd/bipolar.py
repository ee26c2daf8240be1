"""Direct bipolar (l1, l2, L, M) basis: all M, both parities.

The channels and their order are those of the reference's ``BipolarHelium``
(python/bipolar.py:34-99).  The angular operators are formed in the coupled
basis itself by Racah algebra (``angular.scalar_coupled`` for the multipoles of
1/r12, ``angular.tensor_coupled`` for one-electron tensors), as sparse channel
matrices, and ``helium3d.edge_tables`` turns them into the kernel's edge lists.
The product basis is reached through the Clebsch-Gordan matrix ``transform``.
"""
from types import SimpleNamespace

import numpy as np
from scipy.sparse import csr_matrix

from .angular import channels, clebsch, scalar_coupled, spherical_to_cartesian, tensor_coupled
from .helium3d import OperatorTables, cutoff, edge_tables, multipoles, one_electron_diagonal


def bipolar_channels(lmax, total_Lmax, M=None, natural_parity=False):
    """(l1, l2, L, M) channels in the reference's order (python/bipolar.py:38-40)."""
    return [(l1, l2, L, m) for l1 in range(lmax + 1) for l2 in range(lmax + 1)
            for L in range(abs(l1 - l2), min(l1 + l2, total_Lmax) + 1)
            if not natural_parity or (l1 + l2 + L) % 2 == 0
            for m in range(-L, L + 1) if M is None or m == M]


class BipolarHelium(OperatorTables):
    def __init__(self, grid, lmax=3, total_Lmax=3, M=None, cutoff_radii=None, natural_parity=False):
        self.grid = grid
        self.r = np.asfortranarray(grid.r)
        self.n = len(self.r)
        self.lmax, self.total_Lmax = lmax, total_Lmax
        self.basis = bipolar_channels(lmax, total_Lmax, M, natural_parity)
        self.index = {c: i for i, c in enumerate(self.basis)}
        self.nc = len(self.basis)
        self.shape = (self.n, self.n, self.nc)
        self.size = self.n * self.n * self.nc
        self.cutoff_radii = cutoff_radii
        self.cut = np.ones(self.n) if cutoff_radii is None else cutoff(self.r.real, *cutoff_radii)
        self._finish_radial(grid)
        l1 = [c[0] for c in self.basis]
        l2 = [c[1] for c in self.basis]
        self.d1 = np.stack([one_electron_diagonal(self.r, self.cut, l) for l in l1], axis=1)
        self.d2 = np.stack([one_electron_diagonal(self.r, self.cut, l) for l in l2], axis=1)
        self._diag = None
        self.lam_count = 2 * lmax + 1
        self._rad = None
        dipole = []
        for electron in (1, 2):
            dipole.append(spherical_to_cartesian(*(self.tensor_matrix(electron, 1, q) for q in (-1, 0, 1))))
        self._set_edges(*edge_tables(self._multipoles(), dipole, l1, l2))
        self.exchange = np.array([self.index[(b, a, L, m)] for a, b, L, m in self.basis])
        self.exchange_sign = np.array([(-1) ** (a + b - L) for a, b, L, m in self.basis])
        product = channels(lmax, M)
        self.transform = self._clebsch_matrix(product)
        gap = (self.transform.T @ self.transform).toarray() - np.eye(self.nc)
        if np.abs(gap).max(initial=0.0) > 1e-12:
            raise RuntimeError('Clebsch-Gordan change of basis lost orthonormality')
        self.uncoupled = SimpleNamespace(grid=grid, r=self.r, n=self.n, nc=len(product), ch=product, cut=self.cut,
                                         lmax=lmax, cutoff_radii=cutoff_radii, angular_key=(lmax, M),
                                         shape=(self.n, self.n, len(product)),
                                         size=self.n * self.n * len(product))
        self.uncoupled.prepare_surface = lambda radius: OperatorTables.prepare_surface(self.uncoupled, radius)

    def _multipoles(self):
        """Order-k multipole factors of 1/r12: diagonal in (L, M), one block per (L, M)."""
        blocks = {}
        for a, c in enumerate(self.basis):
            blocks.setdefault(c[2:], []).append(a)
        entries = {k: ([], [], []) for k in range(self.lam_count)}
        for members in blocks.values():
            for a in members:
                l1, l2, L, _ = self.basis[a]
                for b in members:
                    k1, k2 = self.basis[b][:2]
                    for k in range(max(abs(l1 - k1), abs(l2 - k2)), min(l1 + k1, l2 + k2) + 1):
                        value = scalar_coupled(l1, l2, k1, k2, L, k)
                        if value:
                            rows, cols, vals = entries[k]
                            rows.append(a)
                            cols.append(b)
                            vals.append(value)
        return [csr_matrix((entries[k][2], (entries[k][0], entries[k][1])), shape=(self.nc, self.nc))
                for k in range(self.lam_count)]

    def tensor_matrix(self, electron, k, q):
        """Channel matrix of the one-electron tensor C^k_q on ``electron`` (1 or 2)."""
        rows, cols, vals = [], [], []
        for a, bra in enumerate(self.basis):
            l1, l2, L, M = bra
            own = bra[electron - 1]
            N = M - q
            # the tensor changes the electron's l by at most k with parity (-1)^k,
            # and L by at most k
            for new in range(abs(own - k), min(own + k, self.lmax) + 1, 2):
                k1, k2 = (new, l2) if electron == 1 else (l1, new)
                for J in range(max(abs(L - k), abs(k1 - k2)), min(L + k, k1 + k2, self.total_Lmax) + 1):
                    b = self.index.get((k1, k2, J, N))
                    if b is None:
                        continue
                    value = tensor_coupled(bra, self.basis[b], k, q, electron)
                    if abs(value) > 1e-13:
                        rows.append(a)
                        cols.append(b)
                        vals.append(value)
        return csr_matrix((np.asarray(vals, complex), (rows, cols)), shape=(self.nc, self.nc))

    def _clebsch_matrix(self, product):
        where = {c: i for i, c in enumerate(product)}
        rows, cols, vals = [], [], []
        for c, (l1, l2, L, M) in enumerate(self.basis):
            for m1 in range(max(-l1, M - l2), min(l1, M + l2) + 1):
                value = clebsch(l1, m1, l2, M - m1, L, M)
                if abs(value) > 1e-14:
                    rows.append(where[((l1, m1), (l2, M - m1))])
                    cols.append(c)
                    vals.append(value)
        return csr_matrix((vals, (rows, cols)), shape=(len(product), self.nc))

    @property
    def rad(self):
        if self._rad is None:
            self._rad = multipoles(self.r, self.cut, self.lam_count)
        return self._rad

    def expand(self, x):
        """Bipolar -> product channels."""
        a = np.asarray(x).reshape(-1, self.nc, order='F')
        return (self.transform @ a.T).T.ravel(order='F')

    def project(self, x):
        """Product -> bipolar channels."""
        a = np.asarray(x).reshape(-1, self.uncoupled.nc, order='F')
        return (self.transform.T @ a.T).T.ravel(order='F')

    def ground(self, tol=1e-10, cache_dir=None, device='cuda:0'):
        """Ground state in the bipolar basis (python/bipolar.py:110-117)."""
        from .ground import bipolar_ground
        return bipolar_ground(self, tol, cache_dir, device)
