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
from scipy.sparse.linalg import LinearOperator

from .angular import cartesian, channels, coulomb


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

    def _set_edges(self, vp, vc, vl, vv, dp, dc, de, dv, dl):
        self.vptr = np.array(vp, np.int32)
        self.vc = np.array(vc, np.int32)
        self.vl = np.array(vl, np.int32)
        self.vcoef = np.array(vv, complex)
        self.dptr = np.array(dp, np.int32)
        self.dc = np.array(dc, np.int32)
        self.de = np.array(de, np.int32)
        self.dcoef = np.asfortranarray(np.array(dv, complex).reshape(-1, 3).T)
        self.ldiff = np.array(dl, dtype=float)

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


@lru_cache(None)
def _product_edges(lmax, M):
    """Coulomb and field-coupling edge lists of the product channels
    (python/helium3d.py:31-48); they depend on (lmax, M) only, so the real
    and the complex-scaled operator of one run share them."""
    ch = channels(lmax, M)
    lam_count = 2 * lmax + 1
    vp, vc, vl, vv, dp, dc, de, dv, dl = [0], [], [], [], [0], [], [], [], []
    for a, bra in enumerate(ch):
        for b, ket in enumerate(ch):
            for lam in range(lam_count):
                value = coulomb(bra, ket, lam)
                if abs(value) > 1e-13:
                    vc.append(b)
                    vl.append(lam)
                    vv.append(value)
            for electron in (0, 1):
                if bra[1 - electron] != ket[1 - electron]:
                    continue
                value = cartesian(bra[electron], ket[electron])
                if np.max(np.abs(value)) < 1e-13:
                    continue
                dc.append(b)
                de.append(electron + 1)
                dv.append(value)
                l, ll = bra[electron][0], ket[electron][0]
                dl.append(l * (l + 1) - ll * (ll + 1))
        vp.append(len(vc))
        dp.append(len(dc))
    return tuple(tuple(x) for x in (vp, vc, vl, vv, dp, dc, de, dv, dl))


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
        self._set_edges(*_product_edges(lmax, M))
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
