"""Device operator and device Arnoldi on the HIP engine.

Drop-in for the reference's GPU operator and propagator
(python/torch_backend.py:11-167): ``HipHamiltonian`` has the constructor,
attributes and methods of ``TorchHamiltonian``; ``exponential_action`` and
``step`` have the signatures, return values, adaptive control flow and error
messages of the reference's device Arnoldi.  PyTorch owns the device buffers
and supplies the stream; all arithmetic on them is done by the HIP kernels of
``libhexhip.so``.
"""
import ctypes as ct

import numpy as np
import torch

from . import native


def _stream(device):
    return ct.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def _vp(t):
    return ct.c_void_p(t.data_ptr())


class HipHamiltonian:
    """Device-resident H (python/torch_backend.py:11-138)."""

    def __init__(self, h, device='cuda:0', angular_sparse=True, coulomb_backend='sparse', bq=None):
        if coulomb_backend not in ('sparse', 'blocks'):
            raise ValueError('unknown Coulomb contraction')
        self.h = h
        self.device = torch.device(device)
        if self.device.type != 'cuda':
            raise ValueError('HipHamiltonian runs on a HIP device (cuda:N); there is no CPU fallback')
        self.n, self.nc, self.size = h.n, h.nc, int(h.size)
        self.dtype = torch.complex128
        self.angular_sparse = angular_sparse
        # Both contraction names select the same coupling pass, which
        # regenerates the multipoles per tile and stores no W blocks.
        self.coulomb_backend = coulomb_backend
        self.coulomb_block_bytes = 0
        self._engine = None
        self._keep = self._tables(h)
        handle = ct.c_void_p()
        if bq is None:
            native.check(native.lib().hex_engine_create(ct.byref(self._keep['struct']), self.device.index or 0,
                                                        ct.byref(handle)), 'engine')
        else:
            # radial field-coupling walks over q in place of r (mixed gauge)
            self._keep['bq'] = np.ascontiguousarray(bq, dtype=np.complex128)
            native.check(native.lib().hex_engine_create_mixed(ct.byref(self._keep['struct']),
                                                              self._keep['bq'].ctypes.data_as(ct.c_void_p),
                                                              self.device.index or 0, ct.byref(handle)), 'engine')
        self._engine = handle
        self._diagonal = None
        self._surface_ns = 0
        if hasattr(h, 'surface_indices'):
            self._set_surface(h)

    @staticmethod
    def _tables(h):
        def arr(x, dtype):
            return np.ascontiguousarray(np.asarray(x), dtype=dtype)
        keep = {
            'ptr': arr(h.ptr, np.int32), 'col': arr(h.col, np.int32), 'tv': arr(h.tv, np.complex128),
            'r': arr(h.r, np.complex128), 'cut': arr(h.cut, np.float64),
            'd1': np.asfortranarray(h.d1, dtype=np.complex128), 'd2': np.asfortranarray(h.d2, dtype=np.complex128),
            'vptr': arr(h.vptr, np.int32), 'vc': arr(h.vc, np.int32), 'vl': arr(h.vl, np.int32),
            'vcoef': arr(h.vcoef, np.complex128),
            'dptr': arr(h.dptr, np.int32), 'dc': arr(h.dc, np.int32), 'de': arr(h.de, np.int32),
            'dcoef': np.asfortranarray(h.dcoef, dtype=np.complex128), 'ldiff': arr(h.ldiff, np.float64),
        }
        t = native.HexTables()
        t.n, t.nc, t.lam_count = h.n, h.nc, int(getattr(h, 'lam_count', 2 * h.lmax + 1))
        t.nt, t.nv, t.nd = len(keep['tv']), len(keep['vc']), len(keep['dc'])
        for name in ('ptr', 'col', 'tv', 'r', 'cut', 'd1', 'd2', 'vptr', 'vc', 'vl', 'vcoef',
                     'dptr', 'dc', 'de', 'dcoef', 'ldiff'):
            setattr(t, name, keep[name].ctypes.data_as(ct.c_void_p))
        keep['struct'] = t
        return keep

    def __del__(self):
        if getattr(self, '_engine', None) is not None and getattr(native, '_lib', None) is not None:
            native.lib().hex_engine_destroy(self._engine)
            self._engine = None

    @property
    def engine_bytes(self):
        return int(native.lib().hex_engine_bytes(self._engine))

    # -- tensor plumbing (python/torch_backend.py:77-80) ---------------------
    def tensor(self, x):
        return torch.as_tensor(np.array(x, copy=True), dtype=self.dtype, device=self.device)

    def state(self, x):
        return self.tensor(np.asarray(x).ravel(order='F') if np.ndim(x) > 1 else np.asarray(x).ravel())

    def host(self, x):
        return x.detach().cpu().numpy()

    def synchronize(self):
        torch.cuda.synchronize(self.device)

    def fields(self, field):
        return tuple(float(a) for a in field)

    def _check_state(self, x):
        if x.dtype != self.dtype or x.device != self.device or x.numel() != self.size:
            raise ValueError(f'state must be a complex128 tensor of {self.size} entries on {self.device}')
        if not x.is_contiguous():
            raise ValueError('state must be contiguous')

    # -- H.psi (python/torch_backend.py:109-131) ------------------------------
    @torch.no_grad()
    def apply(self, x, field=(0., 0., 0.), velocity=True, out=None, scale=1.0):
        x = x.reshape(-1)
        self._check_state(x)
        if out is None:
            out = torch.empty_like(x)
        f = (ct.c_double * 3)(*self.fields(field))
        native.check(native.lib().hex_apply(self._engine, f, int(bool(velocity)), float(scale), _vp(x), _vp(out),
                                            _stream(self.device)), 'apply')
        return out

    @torch.no_grad()
    def apply_adjoint(self, x, field=(0., 0., 0.), velocity=True, out=None):
        """H(field)^dagger x (conjugated irECS contour); used by the costate."""
        x = x.reshape(-1)
        self._check_state(x)
        if out is None:
            out = torch.empty_like(x)
        f = (ct.c_double * 3)(*self.fields(field))
        native.check(native.lib().hex_apply_adjoint(self._engine, f, int(bool(velocity)), _vp(x), _vp(out),
                                                    _stream(self.device)), 'apply_adjoint')
        return out

    @torch.no_grad()
    def apply_terms(self, x, field, velocity=True, terms=7, a=1., b=0., adjoint=False, out=None):
        """out = a * H_terms x + b * x; terms: 1 base (diag+T+Coulomb), 2 dipole, 4 |A|^2."""
        x = x.reshape(-1)
        self._check_state(x)
        if out is None:
            out = torch.empty_like(x)
        f = (ct.c_double * 3)(*self.fields(field))
        ca = (ct.c_double * 2)(complex(a).real, complex(a).imag)
        cb = (ct.c_double * 2)(complex(b).real, complex(b).imag)
        native.check(native.lib().hex_apply_terms(self._engine, f, int(bool(velocity)), int(bool(adjoint)),
                                                  int(terms), ca, cb, _vp(x), _vp(out), _stream(self.device)),
                     'apply_terms')
        return out

    @torch.no_grad()
    def apply_batched(self, x, fields, velocity=True, out=None):
        """H(fields[b]) x[b] for a (B, size) stack of states in one launch pair
        (the batched pattern's apply; no reference counterpart)."""
        x = x.reshape(-1, self.size)
        if x.dtype != self.dtype or x.device != self.device or not x.is_contiguous():
            raise ValueError('states must be a contiguous complex128 (B, size) tensor on the engine device')
        B = x.shape[0]
        f = torch.as_tensor(np.asarray(fields, float).reshape(B, 3), dtype=torch.float64, device=self.device)
        if out is None:
            out = torch.empty_like(x)
        native.check(native.lib().hex_apply_batched(self._engine, B, _vp(f), int(bool(velocity)), _vp(x), _vp(out),
                                                    _stream(self.device)), 'apply_batched')
        return out

    def set_launch(self, rows_per_block):
        native.check(native.lib().hex_engine_set_launch(self._engine, int(rows_per_block)), 'set_launch')

    def autotune(self, members=1):
        """Set the fastest radial launch shape of this operator's apply for
        ``members`` states at a time (velocity gauge with a field; timed once
        per element order, n, channel count and member count, then cached,
        as the batched engine does) and return it."""
        from .batched import autotune
        return autotune(self, members)

    @torch.no_grad()
    def apply_affine(self, x, field, velocity, a, b, out=None):
        """out = a * H(field) x + b * x in one fused pass."""
        x = x.reshape(-1)
        self._check_state(x)
        if out is None:
            out = torch.empty_like(x)
        f = (ct.c_double * 3)(*self.fields(field))
        ca = (ct.c_double * 2)(complex(a).real, complex(a).imag)
        cb = (ct.c_double * 2)(complex(b).real, complex(b).imag)
        native.check(native.lib().hex_apply_affine(self._engine, f, int(bool(velocity)), ca, cb, _vp(x), _vp(out),
                                                   _stream(self.device)), 'apply')
        return out

    @property
    def diagonal(self):
        """Jacobi diagonal: diag + T_ii + T_jj + Coulomb self terms (torch_backend.py:24-31)."""
        if self._diagonal is None:
            h = self.h
            d = np.asarray(h.diag).copy()
            td = h.grid.kinetic.diagonal()
            d += td[:, None, None] + td[None, :, None]
            rad = h.rad
            for a in range(h.nc):
                for k in range(h.vptr[a], h.vptr[a + 1]):
                    if h.vc[k] == a:
                        d[:, :, a] += h.vcoef[k] * rad[:, :, h.vl[k]]
            self._diagonal = self.tensor(d.ravel(order='F'))
        return self._diagonal

    # -- tSURFF source (python/torch_backend.py:133-138) ----------------------
    def _set_surface(self, h):
        J = np.asarray(h.surface_T, complex)
        P = np.asarray(h.surface_P, complex)
        cols = np.flatnonzero(np.any(J != 0, axis=0) | np.any(P != 0, axis=0))
        k0, k1 = int(cols.min()), int(cols.max()) + 1
        Jw = np.ascontiguousarray(J[:, k0:k1])
        Pw = np.ascontiguousarray(P[:, k0:k1])
        native.check(native.lib().hex_engine_set_surface(self._engine, J.shape[0], k0, k1 - k0,
                                                         Jw.ctypes.data_as(ct.c_void_p),
                                                         Pw.ctypes.data_as(ct.c_void_p)), 'surface')
        self._surface_ns = J.shape[0]
        self.surface_indices = np.asarray(h.surface_indices)

    def _need_surface(self):
        if not self._surface_ns:
            if not hasattr(self.h, 'surface_indices'):
                raise ValueError('prepare_surface must be called before flux')
            self._set_surface(self.h)

    @torch.no_grad()
    def flux_device(self, x, field, out=None, terms=3):
        """Surface source left on the device as a (nc, ns, n) tensor; ``terms``
        1 keeps only the J rows, 2 only the field-coupling rows."""
        self._need_surface()
        x = x.reshape(-1)
        self._check_state(x)
        if out is None:
            out = torch.empty((self.nc, self._surface_ns, self.n), dtype=self.dtype, device=self.device)
        f = (ct.c_double * 3)(*self.fields(field))
        native.check(native.lib().hex_flux_terms(self._engine, f, int(terms), _vp(x), _vp(out),
                                                 _stream(self.device)), 'flux')
        return out

    @torch.no_grad()
    def flux_adjoint(self, g0, gd, out=None):
        """Adjoint of psi -> (J part, unit-z-field part) of the surface source:
        the state J^H g0 + D_z^H gd for (nc, ns, n) device tensors g0, gd."""
        self._need_surface()
        shape = (self.nc, self._surface_ns, self.n)
        for g in (g0, gd):
            if tuple(g.shape) != shape or g.dtype != self.dtype or g.device != self.device or not g.is_contiguous():
                raise ValueError(f'frame gradients must be contiguous complex128 {shape} tensors on {self.device}')
        if out is None:
            out = torch.empty(self.size, dtype=self.dtype, device=self.device)
        native.check(native.lib().hex_flux_adjoint(self._engine, _vp(g0), _vp(gd), _vp(out), _stream(self.device)),
                     'flux adjoint')
        return out

    @torch.no_grad()
    def flux(self, x, field):
        """Host (n, ns, nc) surface source, as the reference returns it."""
        return self.host(self.flux_device(x, field).permute(2, 1, 0))

    # -- Krylov helpers ------------------------------------------------------
    def norm(self, x):
        value = ct.c_double()
        native.check(native.lib().hex_norm(self._engine, _vp(x), x.numel(), ct.byref(value), _stream(self.device)),
                     'norm')
        return value.value


class Workspace:
    """Reduction scratch of one device for the Krylov kernels."""

    def __init__(self, device):
        handle = ct.c_void_p()
        native.check(native.lib().hex_workspace_create(device, ct.byref(handle)), 'workspace')
        self._engine = handle

    def __del__(self):
        if getattr(self, '_engine', None) is not None and getattr(native, '_lib', None) is not None:
            native.lib().hex_engine_destroy(self._engine)
            self._engine = None


_WORKSPACES = {}


def _engine_for(x):
    key = x.device.index or 0
    if key not in _WORKSPACES:
        _WORKSPACES[key] = Workspace(key)
    return _WORKSPACES[key]


def _norm(e, x):
    value = ct.c_double()
    native.check(native.lib().hex_norm(e._engine, _vp(x), x.numel(), ct.byref(value), _stream(x.device)), 'norm')
    return value.value


def _scale_into(x, s, y):
    native.check(native.lib().hex_scale(_vp(x), float(s), _vp(y), x.numel(), _stream(x.device)), 'scale')


def _combine(e, Q, m, c, N, out):
    cc = np.ascontiguousarray(c, dtype=np.complex128)
    native.check(native.lib().hex_combine(e._engine, _vp(Q), N, m, cc.ctypes.data_as(ct.c_void_p), _vp(out), N,
                                          _stream(out.device)), 'combine')
    return out


_CHECK_WORK = {}


def _check_buffers(device, order):
    """Device Hessenberg (order + 1) x order, expm scratch and the check result
    record of the Krylov control kernels (kept per device and order)."""
    key = (device.index or 0, order)
    if key not in _CHECK_WORK:
        _CHECK_WORK[key] = (torch.zeros((order + 1) * order, dtype=torch.complex128, device=device),
                            torch.empty(8 * order * order, dtype=torch.complex128, device=device),
                            torch.empty(2 + 64, dtype=torch.complex128, device=device))
    return _CHECK_WORK[key]


@torch.no_grad()
def exponential_action(apply, x, dt, tol=1e-10, maxdim=48):
    """exp(-i dt H) x by Arnoldi with CGS2 (python/torch_backend.py:140-158).

    Same check cadence (j >= 3 and (j+1) % 4 == 0, or j+1 == maxdim), happy
    breakdown below 1e-13, error estimate |dt h_{m+1,m} c_m| and error message.
    The Krylov control runs on the device: hex_cgs2_commit leaves every
    Hessenberg column in device memory, and at the check points (and at the
    last column) hex_arnoldi_check replays the per-column tests on the columns
    formed since the last one -- a happy breakdown included, so a breakdown
    between check points returns what the reference returns when it meets it
    -- with the small exponential by the one-wave scaling-and-squaring Pade;
    the host reads back only its verdict, dimension, estimate and
    coefficients.
    """
    if maxdim > 62:
        raise ValueError('maxdim above 62 is not supported by the device Krylov kernels')
    x = x.reshape(-1).contiguous()
    e = _engine_for(x)
    N = x.numel()
    beta = _norm(e, x)
    if beta == 0:
        return x.clone(), {'dimension': 0, 'residual_estimate': 0.}
    Q = torch.empty((maxdim + 1, N), dtype=x.dtype, device=x.device)
    _scale_into(x, beta, Q[0])
    err = float('nan')
    Hdev = torch.empty((maxdim, 64), dtype=x.dtype, device=x.device)
    Hfull, work, result = _check_buffers(x.device, maxdim)
    stream = _stream(x.device)
    lib = native.lib()
    synced = 0
    for j in range(maxdim):
        v = apply(Q[j])
        native.check(lib.hex_cgs2_commit(e._engine, _vp(Q), N, j + 1, _vp(v), N, _vp(Hdev[j]), _vp(Q[j + 1]),
                                         stream), 'cgs2')
        if not (j >= 3 and (j + 1) % 4 == 0 or j + 1 == maxdim):
            continue
        native.check(lib.hex_arnoldi_check(_vp(Hdev), synced, j, maxdim, float(dt), float(beta), float(tol),
                                           _vp(Hfull), _vp(work), _vp(result), stream), 'arnoldi check')
        r = result[:2 + j + 1].cpu().numpy()
        synced = j + 1
        if r[1].real >= 0:
            err = float(r[1].real)
        if r[0].real == 1:
            m = int(r[0].imag)
            y = torch.empty_like(x)
            _combine(e, Q, m, r[2:2 + m], N, y)
            return y, {'dimension': m, 'residual_estimate': err}
    raise RuntimeError(f'GPU Arnoldi did not converge: dim={maxdim}, dt={dt}, estimate={err:.3e}')


def step(apply_at, x, t, dt, tol=1e-10, maxdim=48, depth=0):
    """Midpoint-field Arnoldi step, halved recursively on non-convergence
    (python/torch_backend.py:160-167)."""
    try:
        return exponential_action(lambda y: apply_at(t + dt / 2, y), x, dt, tol, maxdim)
    except RuntimeError as error:
        if 'Arnoldi did not converge' not in str(error) or depth >= 8:
            raise
        y, a = step(apply_at, x, t, dt / 2, tol / 2, maxdim, depth + 1)
        y, b = step(apply_at, y, t + dt / 2, dt / 2, tol / 2, maxdim, depth + 1)
        return y, {'dimension': max(a['dimension'], b['dimension']),
                   'residual_estimate': a['residual_estimate'] + b['residual_estimate'], 'subdivided': True}
