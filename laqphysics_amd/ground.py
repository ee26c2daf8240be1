"""Field-free 1S ground state of helium, solved on the device.

Mirror of python/ground_s.py:15-89 (``GroundS``).  A total-L = 0 singlet is
sum_m (-1)^(l-m) Y_lm Y_l,-m / sqrt(2l+1) in every (l, l) pair, so the ground
solve needs only lmax+1 radial slabs u_l(r1, r2) instead of the full channel
set; the radial discretisation is the propagated one on the real grid.

Compact operator (python/ground_s.py:33-45): the diagonal of the (l, l) pair
channels, the radial kinetic CSR and multipoles of the parent, and Coulomb
edges l <- j of order k in |l-j| .. l+j step 2 with coefficient
(-1)^(l+j) sqrt((2l+1)(2j+1)) (l j k; 0 0 0)^2, kept where |c| > 1e-14.  No
field couplings.

Solver (python/ground_s.py:55-89): LOBPCG for the smallest eigenvalue of the
real symmetric operator, started from the screened (1s)^2 product
u = r exp(-27/16 r) sqrt(w) in the l = 0 slab and preconditioned by the exact
inverse of H1 (+) H2 - shift in the per-l one-electron eigenbases
(shift = 2 min e_0 - 1/2).  Here every H.psi runs in the HIP engine, the
preconditioner's dense transforms in the separable-resolvent GEMMs
(``hex_separable_apply`` with alpha = -i, which turns its 1 + i alpha (e_i +
e_j + s) denominator into e_i + e_j - shift), and the block-size-1 LOBPCG's
vector work in the engine's dot/combine/axpby kernels; only the 3 x 3
Rayleigh-Ritz problem is solved on the host.  The eigenvector's sign is fixed
so that its largest-magnitude entry is positive.

Results, the on-disk cache (``ground_<sha256>.npz`` with ``state``,
``energy``, ``residual``; the key hashes the same arrays as the reference, so
the two share cache files) and errors follow the reference.  ``method='eigsh'``
runs SciPy's ARPACK on the host with the device H.psi as its matvec.
"""
import ctypes as ct
import hashlib
import os
from pathlib import Path
from types import SimpleNamespace

import numpy as np
import torch
from scipy.linalg import cholesky, eigh, solve_triangular
from sympy.physics.wigner import wigner_3j

from . import native
from .helium3d import cutoff, multipoles, one_electron_diagonal


def compact_coulomb(nc):
    """Coulomb edges of the L = 0 compact basis (python/ground_s.py:37-43)."""
    ptr, cols, orders, values = [0], [], [], []
    for l in range(nc):
        for j in range(nc):
            for k in range(abs(l - j), l + j + 1, 2):
                c = (-1) ** (l + j) * np.sqrt((2 * l + 1) * (2 * j + 1)) * float(wigner_3j(l, j, k, 0, 0, 0)) ** 2
                if abs(c) > 1e-14:
                    cols.append(j)
                    orders.append(k)
                    values.append(c)
        ptr.append(len(cols))
    return (np.array(ptr, np.int32), np.array(cols, np.int32), np.array(orders, np.int32),
            np.array(values, complex))


class GroundS:
    """Compact L = 0 singlet operator and its ground-state solve."""

    @classmethod
    def from_grid(cls, grid, lmax, cutoff_radii=None):
        """Ground-only tables straight from a grid (python/ground_s.py:16-30)."""
        r = np.asfortranarray(grid.r, dtype=complex)
        n = len(r)
        T = grid.kinetic.tocsr()
        cut = np.ones(n) if cutoff_radii is None else cutoff(r.real, *cutoff_radii)
        d = np.stack([one_electron_diagonal(r, cut, l) for l in range(lmax + 1)], axis=1)
        parent = SimpleNamespace(n=n, nc=lmax + 1, shape=(n, n, lmax + 1), grid=grid, r=r, cut=cut,
                                 ptr=T.indptr.astype(np.int32), col=T.indices.astype(np.int32),
                                 tv=T.data.astype(complex), lam_count=2 * lmax + 1,
                                 ch=[((l, 0), (l, 0)) for l in range(lmax + 1)], d1=d, d2=d.copy())
        result = cls(parent)
        result.compact_only = True
        return result

    def __init__(self, h):
        self.parent = h
        self.grid = h.grid
        self.n = h.n
        self.lmax = max(a[0] for a, b in h.ch)
        self.nc = self.lmax + 1
        self.shape = (self.n, self.n, self.nc)
        self.size = int(np.prod(self.shape))
        self.compact_only = False
        for name in ('r', 'cut', 'ptr', 'col', 'tv'):
            setattr(self, name, getattr(h, name))
        index = [h.ch.index(((l, 0), (l, 0))) for l in range(self.nc)]
        self.d1 = np.ascontiguousarray(np.asarray(h.d1)[:, index])
        self.d2 = np.ascontiguousarray(np.asarray(h.d2)[:, index])
        self.lam_count = int(getattr(h, 'lam_count', 2 * self.lmax + 1))
        self.vptr, self.vc, self.vl, self.vcoef = compact_coulomb(self.nc)
        self.dptr = np.zeros(self.nc + 1, np.int32)
        self.dc = np.zeros(0, np.int32)
        self.de = np.zeros(0, np.int32)
        self.dcoef = np.empty((3, 0), complex, order='F')
        self.ldiff = np.zeros(0)
        self._rad = None
        self._engine = None

    @property
    def diag(self):
        return np.asfortranarray(self.d1[:, None, :] + self.d2[None, :, :])

    @property
    def rad(self):
        if self._rad is None:
            self._rad = multipoles(self.r, self.cut, self.lam_count)
        return self._rad

    def apply(self, x):
        """Field-free compact H.psi with host buffers (python/ground_s.py:47)."""
        return native.tensor_apply(self, x)

    def expand(self, x):
        """Compact slabs -> parent product channels (python/ground_s.py:49-54)."""
        if self.compact_only:
            return np.asarray(x).copy()
        u = np.asarray(x).reshape(self.shape, order='F')
        out = np.zeros(self.parent.shape, complex, order='F')
        for c, ((l, m), (j, n)) in enumerate(self.parent.ch):
            if l == j and m == -n:
                out[:, :, c] = (-1) ** (l - m) / np.sqrt(2 * l + 1) * u[:, :, l]
        return out.ravel(order='F')

    def starting_guess(self):
        """Screened (1s)^2 product in the l = 0 slab (python/ground_s.py:64-65)."""
        r = np.asarray(self.r).real
        u = r * np.exp(-1.6875 * r) * np.sqrt(self.grid.weights.real)
        x = np.zeros(self.shape)
        x[:, :, 0] = u[:, None] * u[None, :]
        return x.ravel(order='F')

    def cache_key(self, tol):
        digest = hashlib.sha256(b'helium-coupled-S-v1')
        for a in (self.r, self.grid.weights, self.tv, self.col, self.ptr, self.cut):
            digest.update(np.ascontiguousarray(a).tobytes())
        digest.update(str((self.lmax, tol)).encode())
        return digest.hexdigest()

    def engine(self, device='cuda:0'):
        if self._engine is None:
            from .hip_backend import HipHamiltonian
            self._engine = HipHamiltonian(self, device)
        return self._engine

    def solve(self, tol=1e-10, cache_dir=None, expand=True, method='lobpcg', device='cuda:0', maxiter=250):
        """(E, state, residual) of the ground state (python/ground_s.py:55-89)."""
        if self.grid.ecs_angle:
            raise ValueError('ground solve requires real grid')
        path = Path(cache_dir) / ('ground_' + self.cache_key(tol) + '.npz') if cache_dir else None
        if path is not None and path.exists():
            data = np.load(path)
            small = data['state']
            return float(data['energy']), self.expand(small) if expand else small, float(data['residual'])
        engine = self.engine(device)
        if method == 'lobpcg':
            E, small = _lobpcg(self, engine, tol, maxiter)
        elif method == 'eigsh':
            from scipy.sparse.linalg import LinearOperator, eigsh

            def matvec(v):
                y = engine.apply(engine.state(np.asarray(v, complex)), (0., 0., 0.), False)
                return engine.host(y).real
            op = LinearOperator((self.size, self.size), matvec=matvec, dtype=float)
            e, v = eigsh(op, k=1, which='SA', v0=self.starting_guess(), tol=tol, ncv=min(40, self.size - 1),
                         maxiter=20000)
            E, small = float(e[0]), _fix_sign(v[:, 0]).astype(complex)
            self.last_iterations = None
        else:
            raise ValueError('unknown ground eigensolver')
        x = engine.state(small)
        hx = engine.apply(x, (0., 0., 0.), False)
        res = engine.norm(_axpby(1., hx, -E, x, torch.empty_like(x)))
        if path is not None:
            path.parent.mkdir(parents=True, exist_ok=True)
            tmp = path.with_name(path.stem + f'.{os.getpid()}.tmp.npz')
            np.savez(tmp, state=small, energy=E, residual=res)
            os.replace(tmp, path)
        return float(E), self.expand(small) if expand else small, float(res)


def _fix_sign(v):
    v = np.asarray(v)
    k = int(np.argmax(abs(v)))
    return v if v[k].real >= 0 else -v


def _axpby(a, x, b, y, out):
    from .hip_backend import _stream, _vp
    ca = (ct.c_double * 2)(complex(a).real, complex(a).imag)
    cb = (ct.c_double * 2)(complex(b).real, complex(b).imag)
    native.check(native.lib().hex_axpby(x.numel(), ca, _vp(x), cb, _vp(y), _vp(out), _stream(out.device)), 'axpby')
    return out


class _Resolvent:
    """Exact inverse of H1 (+) H2 - shift in the per-l real eigenbases
    (python/ground_s.py:68-76), on the separable-resolvent GEMMs."""

    def __init__(self, g, engine):
        T = g.grid.kinetic.toarray().real
        r = np.asarray(g.r).real
        values, vectors = [], []
        for l in range(g.nc):
            e, U = eigh(T + np.diag(-2 * g.cut / r + l * (l + 1) / (2 * r * r)))
            values.append(e)
            vectors.append(U)
        self.shift = 2 * min(values[0]) - .5
        n, nl = g.n, g.nc
        self._S = np.ascontiguousarray(np.stack([np.asfortranarray(U).ravel(order='F') for U in vectors]),
                                       dtype=complex)
        self._I = np.ascontiguousarray(np.stack([np.asfortranarray(U.T).ravel(order='F') for U in vectors]),
                                       dtype=complex)
        self._E = np.ascontiguousarray(np.stack(values), dtype=complex)
        self._l = np.arange(nl, dtype=np.int32)
        handle = ct.c_void_p()
        native.check(native.lib().hex_separable_create(
            engine.device.index or 0, n, g.nc, nl, self._S.ctypes.data_as(ct.c_void_p),
            self._I.ctypes.data_as(ct.c_void_p), self._E.ctypes.data_as(ct.c_void_p),
            self._l.ctypes.data_as(ct.c_void_p), self._l.ctypes.data_as(ct.c_void_p), ct.byref(handle)),
            'ground preconditioner')
        self._handle = handle

    def __del__(self):
        if getattr(self, '_handle', None) is not None and getattr(native, '_lib', None) is not None:
            native.lib().hex_separable_destroy(self._handle)
            self._handle = None

    def __call__(self, x, out):
        from .hip_backend import _stream, _vp
        # 1 + i(-i)(e_i + e_j + s) = e_i + e_j - shift for s = -shift - 1
        native.check(native.lib().hex_separable_apply(self._handle, 0.0, -1.0, -self.shift - 1.0, _vp(x), _vp(out),
                                                      _stream(x.device)), 'ground preconditioner')
        return out


@torch.no_grad()
def _lobpcg(g, engine, tol, maxiter):
    """Block-size-1 LOBPCG (Knyazev) for the smallest eigenpair of the real
    symmetric compact H, with the ground resolvent as preconditioner.

    Each iteration: residual r = Hx - lam x; w = M r with x projected out and
    normalised; Rayleigh-Ritz on span{x, w, p} (p the previous search
    direction) through a Cholesky-orthonormalised 3 x 3 Gram problem, p dropped
    when its Gram matrix is numerically singular; stop when ||r|| <= tol."""
    from .hip_backend import _combine, _engine_for, _norm, _scale_into, _vp, _stream
    M = _Resolvent(g, engine)
    N = g.size
    dev = engine.device
    S = torch.zeros((3, N), dtype=torch.complex128, device=dev)    # x, w, p
    e = _engine_for(S)
    AS = torch.zeros((3, N), dtype=torch.complex128, device=dev)   # H x, H w, H p
    tmp = torch.empty((3, N), dtype=torch.complex128, device=dev)
    r = torch.empty(N, dtype=torch.complex128, device=dev)

    def apply_into(src, dst):
        engine.apply(src, (0., 0., 0.), False, out=dst)

    def dots(k, v):
        z = np.zeros(k, complex)
        native.check(native.lib().hex_dots(e._engine, _vp(S), N, k, _vp(v), N, z.ctypes.data_as(ct.c_void_p),
                                           _stream(dev)), 'dots')
        return z.real

    x0 = engine.state(g.starting_guess())
    _scale_into(x0, _norm(e, x0), S[0])
    apply_into(S[0], AS[0])
    lam = float(dots(1, AS[0])[0])
    have_p = False
    history = []
    for it in range(maxiter):
        _axpby(1., AS[0], -lam, S[0], r)
        rn = _norm(e, r)
        history.append(rn)
        if rn <= tol:
            break
        M(r, S[1])
        xw = float(dots(1, S[1])[0])
        _axpby(1., S[1], -xw, S[0], S[1])
        _scale_into(S[1], _norm(e, S[1]), S[1])
        apply_into(S[1], AS[1])
        k = 3 if have_p else 2
        while True:
            GB = np.empty((k, k))
            GA = np.empty((k, k))
            for b in range(k):
                GB[:, b] = dots(k, S[b])
                GA[:, b] = dots(k, AS[b])
            GB = (GB + GB.T) / 2
            GA = (GA + GA.T) / 2
            try:
                L = cholesky(GB, lower=True)
                if np.linalg.cond(GB) > 1e12:
                    raise np.linalg.LinAlgError('near-singular search space')
            except np.linalg.LinAlgError:
                if k == 3:
                    k = 2
                    continue
                raise RuntimeError('LOBPCG search space collapsed')
            break
        # generalized symmetric eigenproblem GA c = mu GB c via the Cholesky factor
        Li = solve_triangular(L, np.eye(k), lower=True)
        mu, Y = eigh(Li @ GA @ Li.T)
        c = Li.T @ Y[:, 0]
        lam_new = float(mu[0])
        # new x and p (the part of the update outside x), with their H images
        _combine(e, S, k, c, N, tmp[0])
        _combine(e, AS, k, c, N, tmp[1])
        cp = c.copy()
        cp[0] = 0.
        _combine(e, S, k, cp, N, tmp[2])
        _combine(e, AS, k, cp, N, r)
        nx = _norm(e, tmp[0])
        _scale_into(tmp[0], nx, S[0])
        _scale_into(tmp[1], nx, AS[0])
        npn = _norm(e, tmp[2])
        if npn > 0:
            _scale_into(tmp[2], npn, S[2])
            _scale_into(r, npn, AS[2])
            have_p = True
        else:
            have_p = False
        lam = lam_new
    else:
        g.last_iterations = len(history)
        raise RuntimeError(f'LOBPCG did not converge: residual {history[-1]:.3e} after {maxiter} iterations')
    g.last_iterations = len(history)
    g.residual_history = history
    x = engine.host(S[0]).real
    return lam, _fix_sign(x).astype(complex)


def _cache(cache_dir):
    return cache_dir or os.environ.get('HELIUM_GROUND_CACHE')


def product_ground(h, tol=1e-10, method='coupled', cache_dir=None, device='cuda:0'):
    """``Helium.ground`` (python/helium3d.py:85-99)."""
    if method == 'coupled':
        return GroundS(h).solve(tol, _cache(cache_dir), device=device)
    if method != 'uncoupled':
        raise ValueError('ground method must be coupled or uncoupled')
    if h.grid.ecs_angle:
        raise ValueError('ground eigsh requires a real grid')
    from scipy.sparse.linalg import LinearOperator, eigsh
    from .hip_backend import HipHamiltonian
    engine = HipHamiltonian(h, device)
    r = np.asarray(h.r).real
    u = r * np.exp(-1.6875 * r) * np.sqrt(h.grid.weights.real)
    x = np.zeros(h.shape, order='F')
    x[:, :, h.ch.index(((0, 0), (0, 0)))] = u[:, None] * u[None, :]

    def matvec(v):
        return engine.host(engine.apply(engine.state(np.asarray(v, complex)), (0., 0., 0.), False)).real
    op = LinearOperator((h.size, h.size), matvec=matvec, dtype=float)
    e, v = eigsh(op, k=1, which='SA', v0=x.ravel(order='F'), tol=tol, ncv=35, maxiter=10000)
    psi = v[:, 0].astype(complex)
    hpsi = engine.host(engine.apply(engine.state(psi), (0., 0., 0.), False))
    return float(e[0]), psi, float(np.linalg.norm(hpsi - e[0] * psi))


def coupled_ground(h, tol=1e-10, cache_dir=None, device='cuda:0'):
    """``CoupledHelium.ground`` (python/coupled_angular.py:73-78): the compact
    slabs land in the (l, l, L = 0) coupled channels."""
    s = GroundS(h.uncoupled)
    E, g, res = s.solve(tol, _cache(cache_dir), expand=False, device=device)
    small = np.asarray(g).reshape(s.shape, order='F')
    out = np.zeros(h.shape, complex, order='F')
    for l in range(h.lmax + 1):
        out[:, :, h.basis.index((l, l, 0))] = small[:, :, l]
    return E, out.ravel(order='F'), res


def bipolar_ground(h, tol=1e-10, cache_dir=None, device='cuda:0'):
    """``BipolarHelium.ground`` (python/bipolar.py:110-117), with this operator's
    own cut-off and multipoles."""
    index = [h.index[(l, l, 0, 0)] for l in range(h.lmax + 1)]
    parent = SimpleNamespace(n=h.n, nc=h.lmax + 1, shape=(h.n, h.n, h.lmax + 1), grid=h.grid, r=h.r, cut=h.cut,
                             ptr=h.ptr, col=h.col, tv=h.tv, lam_count=h.lam_count,
                             ch=[((l, 0), (l, 0)) for l in range(h.lmax + 1)],
                             d1=np.asarray(h.d1)[:, index], d2=np.asarray(h.d2)[:, index])
    s = GroundS(parent)
    s.compact_only = True
    E, g, res = s.solve(tol=tol, cache_dir=cache_dir, expand=False, device=device)
    g = np.asarray(g).reshape(s.shape, order='F')
    out = np.zeros(h.shape, complex, order='F')
    for l in range(h.lmax + 1):
        out[:, :, index[l]] = g[:, :, l]
    return E, out.ravel(order='F'), res
