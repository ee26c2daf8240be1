"""Direct bipolar (l1, l2, L, M) basis: all M, both parities.

Mirror of python/bipolar.py:15-122.  Coulomb couples channels of equal (L, M)
through 6j symbols; the rank-1 dipole tensor of either electron is reduced with
the Wigner-Eckart theorem, giving x/y/z components for circular pulses.
"""
from functools import lru_cache
from types import SimpleNamespace

import numpy as np
from scipy.sparse import csr_matrix
from sympy.physics.wigner import wigner_3j, wigner_6j

from .angular import channels
from .helium3d import OperatorTables, cutoff, multipoles, one_electron_diagonal


@lru_cache(None)
def coupled_coulomb(l, j, ll, jj, L, k):
    """Order-k multipole between (l, j, L) and (ll, jj, L) (python/bipolar.py:15-21)."""
    if not (abs(l - ll) <= k <= l + ll and abs(j - jj) <= k <= j + jj) or (l + ll + k) % 2 or (j + jj + k) % 2:
        return 0.0
    radial = np.sqrt((2 * l + 1) * (2 * j + 1) * (2 * ll + 1) * (2 * jj + 1))
    return float((-1) ** (l + ll + L) * radial * wigner_3j(l, k, ll, 0, 0, 0)
                 * wigner_3j(j, k, jj, 0, 0, 0) * wigner_6j(l, j, L, jj, ll, k))


@lru_cache(None)
def coupled_tensor(bra, ket, k, q, electron):
    """<bra| C^k_q(electron) |ket> in the bipolar basis (python/bipolar.py:23-32)."""
    l, j, L, M = bra
    ll, jj, J, N = ket
    if M != N + q or abs(L - J) > k or L + J < k:
        return 0j
    if electron == 2:
        sign = (-1) ** (l + j - L + ll + jj - J)
        return sign * coupled_tensor((j, l, L, M), (jj, ll, J, N), k, q, 1)
    if j != jj or abs(l - ll) > k or l + ll < k or (l + ll + k) % 2:
        return 0j
    value = ((-1) ** (L - M + j + J + k) * np.sqrt((2 * L + 1) * (2 * J + 1) * (2 * l + 1) * (2 * ll + 1))
             * wigner_3j(l, k, ll, 0, 0, 0) * wigner_3j(L, k, J, -M, q, N) * wigner_6j(l, L, j, J, ll, k))
    return complex(float(value))


class BipolarHelium(OperatorTables):
    def __init__(self, grid, lmax=3, total_Lmax=3, M=None, cutoff_radii=None, natural_parity=False):
        self.grid = grid
        self.r = np.asfortranarray(grid.r)
        self.n = len(self.r)
        self.lmax = lmax
        self.total_Lmax = total_Lmax
        self.basis = [(l, j, L, m) for l in range(lmax + 1) for j in range(lmax + 1)
                      for L in range(abs(l - j), min(l + j, total_Lmax) + 1)
                      if not natural_parity or (l + j - L) % 2 == 0
                      for m in range(-L, L + 1) if M is None or m == M]
        self.index = {x: i for i, x in enumerate(self.basis)}
        self.nc = len(self.basis)
        self.shape = (self.n, self.n, self.nc)
        self.size = int(np.prod(self.shape))
        self.cutoff_radii = cutoff_radii
        self.cut = np.ones(self.n) if cutoff_radii is None else cutoff(self.r.real, *cutoff_radii)
        self._finish_radial(grid)
        r = self.r
        self.d1 = np.stack([one_electron_diagonal(r, self.cut, l) for l, j, L, m in self.basis], axis=1)
        self.d2 = np.stack([one_electron_diagonal(r, self.cut, j) for l, j, L, m in self.basis], axis=1)
        self._diag = None
        self.lam_count = 2 * lmax + 1
        self._rad = None
        by_LM = {}
        for b, (_, _, L, m) in enumerate(self.basis):
            by_LM.setdefault((L, m), []).append(b)
        vp, vc, vl, vv, dp, dc, de, dv, dl = [0], [], [], [], [0], [], [], [], []
        root2 = np.sqrt(2)
        for a, (l, j, L, m) in enumerate(self.basis):
            for b in by_LM[(L, m)]:
                ll, jj, _, _ = self.basis[b]
                for k in range(max(abs(l - ll), abs(j - jj)), min(l + ll, j + jj) + 1):
                    value = coupled_coulomb(l, j, ll, jj, L, k)
                    if abs(value) > 1e-13:
                        vc.append(b)
                        vl.append(k)
                        vv.append(value)
            for e in (1, 2):
                for b in self._tensor_candidates(a, 1, e):
                    minus = coupled_tensor(self.basis[a], self.basis[b], 1, -1, e)
                    plus = coupled_tensor(self.basis[a], self.basis[b], 1, 1, e)
                    value = np.array([(minus - plus) / root2, 1j * (minus + plus) / root2,
                                      coupled_tensor(self.basis[a], self.basis[b], 1, 0, e)])
                    if np.max(abs(value)) > 1e-13:
                        dc.append(b)
                        de.append(e)
                        dv.append(value)
                        left, right = self.basis[a][e - 1], self.basis[b][e - 1]
                        dl.append(left * (left + 1) - right * (right + 1))
            vp.append(len(vc))
            dp.append(len(dc))
        self._set_edges(vp, vc, vl, vv, dp, dc, de, dv, dl)
        self.exchange = np.array([self.index[(j, l, L, m)] for l, j, L, m in self.basis])
        self.exchange_sign = np.array([(-1) ** (l + j - L) for l, j, L, m in self.basis])
        product = channels(lmax, M)
        lookup = {x: i for i, x in enumerate(product)}
        rows, cols, values = [], [], []
        for c, (l, j, L, m) in enumerate(self.basis):
            for m1 in range(-l, l + 1):
                m2 = m - m1
                if abs(m2) > j:
                    continue
                value = (-1) ** (l - j + m) * np.sqrt(2 * L + 1) * float(wigner_3j(l, j, L, m1, m2, -m))
                if abs(value) > 1e-14:
                    rows.append(lookup[((l, m1), (j, m2))])
                    cols.append(c)
                    values.append(value)
        self.transform = csr_matrix((values, (rows, cols)), shape=(len(product), self.nc))
        gap = self.transform.T @ self.transform - csr_matrix(np.eye(self.nc))
        if np.max(abs(gap.data), initial=0) > 1e-12:
            raise RuntimeError('bipolar transformation is not isometric')
        self.uncoupled = SimpleNamespace(grid=grid, r=self.r, n=self.n, nc=len(product), ch=product, cut=self.cut)
        self.uncoupled.prepare_surface = lambda radius: OperatorTables.prepare_surface(self.uncoupled, radius)

    @property
    def rad(self):
        if self._rad is None:
            self._rad = multipoles(self.r, self.cut, self.lam_count)
        return self._rad

    def _tensor_candidates(self, a, k, electron):
        """Channels reachable from ``a`` by a rank-k tensor on one electron (python/bipolar.py:79-87)."""
        l, j, L, m = self.basis[a]
        found = set()
        own = self.basis[a][electron - 1]
        for other in range(max(0, own - k), min(self.lmax, own + k) + 1):
            ll, jj = (other, j) if electron == 1 else (l, other)
            for J in range(max(abs(ll - jj), abs(L - k)), min(ll + jj, L + k, self.total_Lmax) + 1):
                for N in range(max(-J, m - k), min(J, m + k) + 1):
                    b = self.index.get((ll, jj, J, N))
                    if b is not None:
                        found.add(b)
        return sorted(found)

    def expand(self, x):
        a = np.asarray(x).reshape(self.n * self.n, self.nc, order='F')
        return (self.transform @ a.T).T.ravel(order='F')

    def project(self, x):
        a = np.asarray(x).reshape(self.n * self.n, self.uncoupled.nc, order='F')
        return (self.transform.T @ a.T).T.ravel(order='F')

    def ground(self, tol=1e-10, cache_dir=None, device='cuda:0'):
        """Ground state in the bipolar basis (python/bipolar.py:110-117)."""
        from .ground import bipolar_ground
        return bipolar_ground(self, tol, cache_dir, device)
