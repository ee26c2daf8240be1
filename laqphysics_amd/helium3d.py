"""Two-electron 3D helium operator tables in the product Y_lm channel basis.

Mirror of the reference operator object (python/helium3d.py:12-103).  It
assembles the same host tables the reference hands to its native kernel:

* ``ptr, col, tv``  -- CSR of the normalised radial kinetic matrix T;
* ``diag``          -- ``-2 cut/r + l(l+1)/2r^2`` for each electron, summed
                       (python/helium3d.py:25-27), kept F-ordered (n, n, nc);
* ``rad``           -- multipoles ``r_<^lam / r_>^(lam+1) cut_i cut_j`` with
                       r_< / r_> chosen by *index order* on the complex contour
                       (python/helium3d.py:29-30);
* Coulomb edges ``vptr, vc, vl, vcoef`` and dipole edges
  ``dptr, dc, de, dcoef(3,nd), ldiff`` (python/helium3d.py:31-48).

In addition it keeps the separable pieces ``d1[:, c]`` / ``d2[:, c]`` whose sum
is ``diag`` exactly (the same numpy expression), so the device engine can
regenerate ``diag`` from two n-vectors per channel instead of streaming n^2*nc.

``apply`` goes through the drop-in ``native.tensor_apply`` (HIP, host buffers),
as the reference's goes through its Fortran library (python/helium3d.py:79-80).
"""
from functools import lru_cache

import numpy as np
from scipy.sparse import coo_matrix, csr_matrix, identity, kron
from scipy.sparse.linalg import LinearOperator

from .angular import C, channels, spherical_to_cartesian


def cutoff(x, inner, outer):
    """Smoothstep switch 1 -> 0 between ``inner`` and ``outer`` (python/tdse1d.py:18-21)."""
    if not 0 < inner < outer:
        raise ValueError('0<cutoff_inner<cutoff_outer')
    u = np.clip((abs(x) - inner) / (outer - inner), 0, 1)
    return 1 - 10 * u ** 3 + 15 * u ** 4 - 6 * u ** 5


def one_electron_diagonal(r, cut, l):
    """``-2 cut / r + l(l+1) / (2 r^2)`` on the (complex) radial grid."""
    return -2 * cut / r + l * (l + 1) / (2 * r * r)


def multipoles(r, cut, lam_count):
    """``rad[i, j, lam] = r_<^lam / r_>^(lam+1) * cut_i * cut_j`` (F-ordered).

    ``r_<`` and ``r_>`` are chosen by index order, which on the ECS contour is
    the order of the real contour parameter (python/helium3d.py:29-30).
    """
    n = len(r)
    i, j = np.indices((n, n))
    lo = r[np.minimum(i, j)]
    hi = r[np.maximum(i, j)]
    weight = cut[:, None] * cut[None, :]
    return np.asfortranarray(np.stack([lo ** lam / hi ** (lam + 1) * weight
                                       for lam in range(lam_count)], axis=2))


class OperatorTables:
    """Shared behaviour of all channel bases (product, coupled, bipolar)."""

    def _finish_radial(self, grid):
        t = grid.kinetic.tocsr()
        self.ptr = t.indptr.astype(np.int32)
        self.col = t.indices.astype(np.int32)
        self.tv = t.data.astype(complex)

    def _set_edges(self, vptr, vc, vl, vcoef, dptr, dc, de, dcoef, ldiff):
        self.vptr, self.vc, self.vl, self.vcoef = vptr, vc, vl, np.asarray(vcoef, complex)
        self.dptr, self.dc, self.de = dptr, dc, de
        self.dcoef = np.asfortranarray(np.asarray(dcoef, complex).reshape(3, -1))
        self.ldiff = np.asarray(ldiff, float)

    @property
    def diag(self):
        """Full F-ordered (n, n, nc) diagonal, built lazily from ``d1 + d2``."""
        if getattr(self, '_diag', None) is None:
            self._diag = np.asfortranarray(self.d1[:, None, :] + self.d2[None, :, :])
        return self._diag

    @diag.setter
    def diag(self, value):
        self._diag = value

    def apply(self, x, field=(0., 0., 0.), velocity=False):
        from .native import tensor_apply
        return tensor_apply(self, x, field, velocity)

    def operator(self, field=(0., 0., 0.), velocity=False):
        return LinearOperator((self.size, self.size), dtype=complex,
                              matvec=lambda x: self.apply(x, field, velocity))

    def exchange_error(self, x):
        u = np.asarray(x).reshape(self.shape, order='F')
        swapped = u.transpose(1, 0, 2)[:, :, self.exchange]
        sign = getattr(self, 'exchange_sign', None)
        if sign is not None:
            swapped = swapped * sign[None, None, :]
        return float(np.linalg.norm(u - swapped))

    def prepare_surface(self, radius):
        """Discrete commutator ``[T, Theta(r - R)]`` rows touching the surface.

        Same checks and errors as python/helium3d.py:52-67: the stencil must be
        on the real grid, off the ECS bridge and outside the cut-off potential.
        """
        theta = (self.r.real > radius).astype(float)
        T = self.grid.kinetic.toarray()
        comm = T * (theta[None, :] - theta[:, None])
        idx = np.flatnonzero(np.max(abs(comm), axis=1) > 1e-14)
        if not len(idx):
            raise ValueError('empty surface commutator')
        if np.any(abs(self.r[idx].imag) > 1e-12):
            raise ValueError('surface must be inside real grid')
        if np.any(abs(self.grid.weights[idx].imag) > 1e-12):
            raise ValueError('surface stencil touches the ECS bridge; extend the real domain')
        if np.any(self.cut[idx] > 1e-12):
            raise ValueError('Volkov surface stencil overlaps nonzero potential; increase R')
        self.surface_indices = idx
        self.surface_T = comm[idx, :]
        self.surface_P = (1j * comm * (self.r[None, :] - self.r[:, None]))[idx, :]
        return idx

    def flux(self, psi, field):
        """tSURFF source ``(n, n_s, nc)`` on the device (python/helium3d.py:69-77)."""
        from .native import surface_flux
        return surface_flux(self, psi, field)


EDGE_DROP = 1e-13   # angular couplings at or below this magnitude are not edges


def edge_tables(coulomb, dipole, l1, l2):
    """Edge lists of a channel basis from its angular operators.

    ``coulomb[lam]``: (nc, nc) matrix of the order-``lam`` multipole factor;
    ``dipole[e]`` for electrons e = 1, 2: the (x, y, z) components of C^1 on
    that electron, each an (nc, nc) matrix or None; ``l1``, ``l2``: the one-
    electron l of each channel.  Edges of output channel a are listed source by
    source (ascending), multipole orders ascending within a source and
    electron 1 before electron 2, which is the edge order of the reference's
    tables (python/helium3d.py:31-48).  Returns
    ``(vptr, vc, vl, vcoef, dptr, dc, de, dcoef(3, nd), ldiff)``.
    """
    nc = len(l1)
    parts = []
    for lam, V in enumerate(coulomb):
        V = coo_matrix(V)
        keep = np.abs(V.data) > EDGE_DROP
        parts.append((V.row[keep], V.col[keep], np.full(keep.sum(), lam), V.data[keep].astype(complex)))
    row, col, lam, val = (np.concatenate(x) for x in zip(*parts)) if parts else ([],) * 4
    order = np.lexsort((lam, col, row))
    vptr = np.concatenate(([0], np.cumsum(np.bincount(row, minlength=nc)))).astype(np.int32)
    vc, vl, vcoef = col[order].astype(np.int32), lam[order].astype(np.int32), val[order]

    ls = (np.asarray(l1), np.asarray(l2))
    rows, cols, eles, comps = [], [], [], []
    for e, parts3 in ((1, dipole[0]), (2, dipole[1])):
        dense = np.stack([np.zeros((nc, nc), complex) if m is None else np.asarray(
            m.toarray() if hasattr(m, 'toarray') else m, complex) for m in parts3])
        a, b = np.nonzero(np.abs(dense).max(axis=0) > EDGE_DROP)
        rows.append(a)
        cols.append(b)
        eles.append(np.full(len(a), e))
        comps.append(dense[:, a, b])
    row, col, ele = np.concatenate(rows), np.concatenate(cols), np.concatenate(eles)
    comp = np.concatenate(comps, axis=1)
    order = np.lexsort((ele, col, row))
    row, col, ele, comp = row[order], col[order], ele[order], comp[:, order]
    lrow = np.where(ele == 1, ls[0][row], ls[1][row])
    lcol = np.where(ele == 1, ls[0][col], ls[1][col])
    dptr = np.concatenate(([0], np.cumsum(np.bincount(row, minlength=nc)))).astype(np.int32)
    return (vptr, vc, vl, vcoef, dptr, col.astype(np.int32), ele.astype(np.int32),
            np.asfortranarray(comp), (lrow * (lrow + 1) - lcol * (lcol + 1)).astype(float))


@lru_cache(None)
def _one_electron_tensor(lmax, k, q):
    """C^k_q on the one-electron states (l, m), l <= lmax, m = -l..l (sparse)."""
    single = [(l, m) for l in range(lmax + 1) for m in range(-l, l + 1)]
    rows, cols, vals = [], [], []
    for a, bra in enumerate(single):
        for b, ket in enumerate(single):
            v = C(bra, k, q, ket)
            if v:
                rows.append(a)
                cols.append(b)
                vals.append(v)
    return csr_matrix((vals, (rows, cols)), shape=(len(single), len(single)))


@lru_cache(None)
def product_algebra(lmax, M):
    """Angular operators of the product channels ((l1,m1),(l2,m2)), m1+m2 = M
    (all channels when M is None), as channel-space matrices.

    The two-electron space is the Kronecker product of the one-electron one,
    restricted to the listed channels: the multipole factor of 1/r12 is
    ``sum_q (-1)^q C^lam_q (x) C^lam_-q`` and the dipole of electron 1 (2) is
    ``C^1 (x) 1`` (``1 (x) C^1``).  Returns (channels, coulomb, dipole)."""
    ch = channels(lmax, M)
    ns = (lmax + 1) ** 2
    pos = {s: i for i, s in enumerate((l, m) for l in range(lmax + 1) for m in range(-l, l + 1))}
    take = np.array([pos[a] * ns + pos[b] for a, b in ch])
    eye1 = identity(ns, format='csr')

    def restrict(op):
        return csr_matrix(op.tocsr()[take][:, take])

    coulomb = []
    for lam in range(2 * lmax + 1):
        V = csr_matrix((ns * ns, ns * ns))
        for q in range(-lam, lam + 1):
            V = V + (-1.0) ** q * kron(_one_electron_tensor(lmax, lam, q), _one_electron_tensor(lmax, lam, -q))
        coulomb.append(restrict(V))
    minus, zero, plus = (_one_electron_tensor(lmax, 1, q) for q in (-1, 0, 1))
    cart = spherical_to_cartesian(minus, zero, plus)
    dipole = ([restrict(kron(c, eye1)) for c in cart], [restrict(kron(eye1, c)) for c in cart])
    return ch, coulomb, dipole


class Helium(OperatorTables):
    """Product-basis operator (python/helium3d.py:12-50)."""

    def __init__(self, grid, lmax=2, M=0, cutoff_radii=None):
        self.grid = grid
        self.r = np.asfortranarray(grid.r, dtype=complex)
        self.lmax = lmax
        self.ch = channels(lmax, M)
        self.n = len(self.r)
        self.nc = len(self.ch)
        self.shape = (self.n, self.n, self.nc)
        self.size = int(np.prod(self.shape))
        self._finish_radial(grid)
        r = self.r
        self.cutoff_radii = cutoff_radii
        self.cut = np.ones(self.n) if cutoff_radii is None else cutoff(r.real, *cutoff_radii)
        self.d1 = np.stack([one_electron_diagonal(r, self.cut, a[0]) for a, b in self.ch], axis=1)
        self.d2 = np.stack([one_electron_diagonal(r, self.cut, b[0]) for a, b in self.ch], axis=1)
        self._diag = None
        self.lam_count = 2 * lmax + 1
        self._rad = None
        self.angular_key = (lmax, M)
        _, coulomb_ops, dipole_ops = product_algebra(lmax, M)
        self._set_edges(*edge_tables(coulomb_ops, dipole_ops, [a[0] for a, b in self.ch], [b[0] for a, b in self.ch]))
        self.exchange = np.array([self.ch.index((b, a)) for a, b in self.ch])

    @property
    def rad(self):
        if self._rad is None:
            self._rad = multipoles(self.r, self.cut, self.lam_count)
        return self._rad

    def ground(self, tol=1e-10, method='coupled', cache_dir=None, device='cuda:0'):
        """(E, psi_0, residual) of the field-free ground state (python/helium3d.py:85-99)."""
        from .ground import product_ground
        return product_ground(self, tol, method, cache_dir, device)
