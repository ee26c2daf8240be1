"""CF4 Magnus + [2/2] Pade time step with right-preconditioned GMRES, on the device.

Drop-in for python/implicit.py:13-109: ``SeparablePreconditioner``, ``gmres``,
``pade_step`` and ``cf4_step`` keep the reference's signatures, return values,
restart/true-residual control flow and error messages.  The vector work runs
in the HIP engine (fused affine H.psi, CGS2 and norm reductions, basis
combinations); the preconditioner's dense transforms are complex128 GEMMs on
the device (complex64 GEMMs for precision 'complex64', with flexible GMRES as
in the reference); the (restart+1) x restart least-squares problem is solved
on the device at the check points (hex_gmres_check).

``pade_step`` and ``cf4_step`` take an optional ``restart`` (default: the
reference's 32, or 16 for a flexible solve).
"""
import ctypes as ct
import inspect

import numpy as np
import torch
from scipy.linalg import eig, inv

from . import native
from .hip_backend import _combine, _engine_for, _norm, _scale_into, _stream, _vp

SQRT3 = np.sqrt(3.)
C1 = .5 - SQRT3 / 6
C2 = .5 + SQRT3 / 6
A1 = (3 - 2 * SQRT3) / 12
A2 = (3 + 2 * SQRT3) / 12


def _axpby(a, x, b, y, out):
    ca = (ct.c_double * 2)(complex(a).real, complex(a).imag)
    cb = (ct.c_double * 2)(complex(b).real, complex(b).imag)
    native.check(native.lib().hex_axpby(x.numel(), ca, _vp(x), cb, _vp(y), _vp(out), _stream(out.device)), 'axpby')
    return out


class SeparablePreconditioner:
    """Exact resolvent of H1 (+) H2 in per-l eigenbases (python/implicit.py:13-39)."""

    def __init__(self, engine, precision='complex128'):
        h = engine.h
        self.engine = engine
        labels = [tuple(b[:2]) for b in h.basis] if hasattr(h, 'basis') else [(a[0], b[0]) for a, b in h.ch]
        maxl = max(max(p) for p in labels)
        T = h.grid.kinetic.toarray()
        r = h.r
        values, vectors, inverses, conditions = [], [], [], []
        for l in range(maxl + 1):
            H = T + np.diag(-2 * h.cut / r + l * (l + 1) / (2 * r * r))
            e, S = eig(H)
            conditions.append(float(np.linalg.cond(S)))
            values.append(e)
            vectors.append(S)
            inverses.append(inv(S))
        if max(conditions) > 1e12:
            raise RuntimeError('ionic eigenbasis too ill-conditioned for separable preconditioner')
        if precision not in ('complex128', 'complex64'):
            raise ValueError('invalid preconditioner precision')
        # 'complex64' (python/implicit.py:27-29, the reference's production
        # setting) runs the four transforms as complex64 GEMMs on complex64
        # copies of the eigenbases, and the GMRES that uses it is the flexible
        # variant (restart 16, least squares every second column), as in the
        # reference; input and output stay complex128.
        self.condition_numbers = conditions
        self.requested_precision = precision
        self.precision = precision
        self.flexible = precision == 'complex64'
        self.dtype = torch.complex64 if self.flexible else torch.complex128
        n, nl = h.n, maxl + 1
        self._S = np.ascontiguousarray(np.stack([np.asfortranarray(S).ravel(order='F') for S in vectors]))
        self._I = np.ascontiguousarray(np.stack([np.asfortranarray(S).ravel(order='F') for S in inverses]))
        self._E = np.ascontiguousarray(np.stack(values).astype(complex))
        self._l1 = np.array([a for a, b in labels], np.int32)
        self._l2 = np.array([b for a, b in labels], np.int32)
        handle = ct.c_void_p()
        create = native.lib().hex_separable_create_c64 if self.flexible else native.lib().hex_separable_create
        native.check(create(
            engine.device.index or 0, n, h.nc, nl, self._S.ctypes.data_as(ct.c_void_p),
            self._I.ctypes.data_as(ct.c_void_p), self._E.ctypes.data_as(ct.c_void_p),
            self._l1.ctypes.data_as(ct.c_void_p), self._l2.ctypes.data_as(ct.c_void_p), ct.byref(handle)),
            'separable')
        self._handle = handle

    def __del__(self):
        # at interpreter exit the module globals may already be gone
        if getattr(self, '_handle', None) is not None and getattr(native, '_lib', None) is not None:
            native.lib().hex_separable_destroy(self._handle)
            self._handle = None

    @property
    def device_bytes(self):
        return int(native.lib().hex_separable_bytes(self._handle))

    @property
    def lanes(self):
        """Streams the GEMMs of one application are spread over (chosen at creation)."""
        return int(native.lib().hex_separable_lanes(self._handle))

    @lanes.setter
    def lanes(self, value):
        native.check(native.lib().hex_separable_set_lanes(self._handle, int(value)), 'separable lanes')

    def apply(self, x, alpha, shift, out=None):
        x = x.reshape(-1)
        if out is None:
            out = torch.empty_like(x)
        alpha = complex(alpha)
        native.check(native.lib().hex_separable_apply(self._handle, alpha.real, alpha.imag, float(shift), _vp(x),
                                                      _vp(out), _stream(x.device)), 'separable apply')
        return out


@torch.no_grad()
def gmres(apply, b, preconditioner, x0=None, tol=1e-11, restart=32, max_restarts=10, flexible=False,
          reuse_preconditioned=True):
    """Right-preconditioned (F)GMRES with CGS2 and true-residual acceptance
    (python/implicit.py:45-74).

    The reference forms a candidate of the non-flexible solve as
    x + M(V y).  M is linear, so with ``reuse_preconditioned`` (the default)
    the products M v_j of the Arnoldi loop are kept, as the flexible variant
    keeps them, and the candidate is x + sum_j y_j (M v_j): one basis
    combination instead of one more preconditioner application per candidate,
    at the price of ``restart`` more vectors of storage.  The iterates differ
    from the reference's by rounding only.

    The Hessenberg columns and the next basis vectors are formed on the
    device (``hex_cgs2_commit``), and so is the Krylov control: at every
    4th (2nd, flexible) iteration and at the restart length,
    ``hex_gmres_check`` replays the reference's per-iteration test on the
    columns formed since, solving each least-squares problem by Givens
    rotations in one wave; the host reads back only its verdict, dimension,
    estimate and coefficients, and forms the candidate with its true residual
    as the reference does.  Iteration counts are the reference's: columns
    computed past a breakdown found on such a check are discarded and not
    counted."""
    if restart > 62:
        raise ValueError('restart above 62 is not supported by the device Krylov kernels')
    M = preconditioner if callable(preconditioner) else (lambda y: preconditioner * y)
    keep = flexible or reuse_preconditioned
    try:
        writes_out = callable(preconditioner) and 'out' in inspect.signature(preconditioner).parameters
    except (TypeError, ValueError):
        writes_out = False
    b = b.reshape(-1)
    e = _engine_for(b)
    N = b.numel()
    x = b.clone() if x0 is None else x0.reshape(-1).clone()
    bnorm = _norm(e, b)
    if bnorm == 0:
        return torch.zeros_like(b), {'iterations': 0, 'relative_residual': 0.}
    iterations = 0
    true = float('nan')
    cadence = 2 if flexible else 4
    r = torch.empty_like(b)
    # Hessenberg columns stay on the device between the check points: the
    # check kernel reads them every ``cadence``-th iteration and at the restart
    # length and replays the reference's per-iteration test on each column, so
    # a breakdown between check points (columns formed past it are discarded
    # and not counted) gives the reference's result
    Hdev = torch.empty((restart, 64), dtype=b.dtype, device=b.device)
    Hfull = torch.zeros((restart + 1) * restart, dtype=b.dtype, device=b.device)
    result = torch.empty(2 + 64, dtype=b.dtype, device=b.device)
    stream = _stream(b.device)
    lib = native.lib()
    for cycle in range(max_restarts):
        _axpby(1., b, -1., apply(x), r)
        beta = _norm(e, r)
        if beta <= tol * bnorm:
            return x, {'iterations': iterations, 'relative_residual': beta / bnorm}
        Q = torch.empty((restart + 1, N), dtype=b.dtype, device=b.device)
        Z = torch.empty((restart, N), dtype=b.dtype, device=b.device) if keep else None
        _scale_into(r, beta, Q[0])

        def candidate_of(m, y, norm):
            """The reference's candidate test after column m - 1 (the device
            check found the least-squares estimate below tol |b|, a breakdown or
            the restart length): (outcome, candidate), outcome 'done',
            'restart' or None (go on)."""
            nonlocal true
            step = torch.empty_like(b)
            if keep:
                _combine(e, Z, m, y, N, step)
            else:
                _combine(e, Q, m, y, N, step)
                step = M(step)
            candidate = _axpby(1., x, 1., step, torch.empty_like(b))
            _axpby(1., b, -1., apply(candidate), r)
            true = _norm(e, r) / bnorm
            if true <= tol:
                return 'done', candidate
            if norm < 1e-14 or m == restart:
                return 'restart', candidate
            return None, None

        synced = 0
        outcome = None
        for j in range(restart):
            if keep and writes_out:
                pre = M(Q[j], out=Z[j])
            else:
                pre = M(Q[j])
                if keep:
                    Z[j].copy_(pre)
            v = apply(pre)
            iterations += 1
            native.check(lib.hex_cgs2_commit(e._engine, _vp(Q), N, j + 1, _vp(v), N, _vp(Hdev[j]), _vp(Q[j + 1]),
                                             stream), 'cgs2')
            if not ((j + 1) % cadence == 0 or j + 1 == restart):
                continue
            first = synced
            while first <= j:
                native.check(lib.hex_gmres_check(_vp(Hdev), first, j, restart, cadence, float(beta),
                                                 float(tol * bnorm), _vp(Hfull), _vp(result), stream), 'gmres check')
                rec = result[:2 + j + 1].cpu().numpy()
                if rec[0].real != 1:
                    break
                m = int(rec[0].imag)
                outcome, candidate = candidate_of(m, rec[2:2 + m], float(rec[1].imag))
                if outcome is not None:
                    # columns past m - 1 were computed ahead of this test
                    iterations -= j - (m - 1)
                    break
                first = m
            synced = j + 1
            if outcome == 'done':
                return candidate, {'iterations': iterations, 'relative_residual': true}
            if outcome == 'restart':
                x = candidate
                break
    raise RuntimeError(f'GMRES did not converge: residual={true:.3e}, iterations={iterations}')


@torch.no_grad()
def pade_step(engine, x, field, dt, reference_energy=0., tol=1e-11, restart=None):
    """[2/2] Pade factor pair (python/implicit.py:76-89): for both roots
    3 +- i sqrt3 solve (I + i alpha (H - E_ref)) s = x and set x = 2 s - x."""
    a2 = getattr(engine, 'a2_coefficient', 1.)
    field = tuple(float(a) for a in field)
    A2sum = a2 * sum(a ** 2 for a in field)
    iterations = 0
    residual = 0.
    separable = getattr(engine, 'separable', None)
    for root in (3 + 1j * SQRT3, 3 - 1j * SQRT3):
        alpha = dt / root
        ia = 1j * alpha

        if hasattr(engine, 'apply_affine'):
            def A(y, ia=ia):
                return engine.apply_affine(y, field, True, ia, 1. - ia * reference_energy)
        else:
            # engines that only follow the reference protocol (apply, diagonal)
            def A(y, ia=ia):
                return _axpby(ia, engine.apply(y, field, True), 1. - ia * reference_energy, y, torch.empty_like(y))
        if separable is not None:
            def M(y, alpha=alpha, out=None):
                return separable.apply(y, alpha, A2sum - reference_energy, out=out)
        else:
            diagonal = engine.diagonal + A2sum - reference_energy
            M = 1 / (1 + 1j * alpha * diagonal)
        flexible = getattr(separable, 'flexible', False)
        nrestart = restart if restart is not None else (16 if flexible else 32)
        s, info = gmres(A, x, M, tol=tol, flexible=flexible, restart=nrestart)
        x = _axpby(2., s, -1., x, torch.empty_like(s))
        iterations += info['iterations']
        residual = max(residual, info['relative_residual'])
    return x, {'linear_iterations': iterations, 'linear_residual': residual}


@torch.no_grad()
def cf4_step(engine, x, avec, t, dt, reference_energy=0., tol=1e-11, depth=0, restart=None):
    """Commutator-free 4th-order Magnus step of two Pade exponentials with the
    analytic A^2 phase and reference-energy phase (python/implicit.py:91-109)."""
    fields = np.array([avec(t + C1 * dt), avec(t + C2 * dt)], float)
    iterations = 0
    residual = 0.
    original = x
    try:
        for w in ((A2, A1), (A1, A2)):
            effective = 2 * np.einsum('a,aj->j', w, fields)
            correction = getattr(engine, 'a2_coefficient', 1.) * (np.dot(w, np.sum(fields ** 2, axis=1))
                                                                   - .5 * np.dot(effective, effective))
            x, info = pade_step(engine, x, effective, dt / 2, reference_energy, tol, restart)
            phase = np.exp(-1j * dt * (correction + reference_energy / 2))
            x = _axpby(phase, x, 0., x, x)
            iterations += info['linear_iterations']
            residual = max(residual, info['linear_residual'])
        return x, {'linear_iterations': iterations, 'linear_residual': residual, 'dimension': 0,
                   'residual_estimate': None}
    except RuntimeError as error:
        if 'GMRES did not converge' not in str(error) or depth >= 6:
            raise
        x, a = cf4_step(engine, original, avec, t, dt / 2, reference_energy, tol / 2, depth + 1, restart)
        x, b = cf4_step(engine, x, avec, t + dt / 2, dt / 2, reference_energy, tol / 2, depth + 1, restart)
        return x, {'linear_iterations': a['linear_iterations'] + b['linear_iterations'],
                   'linear_residual': max(a['linear_residual'], b['linear_residual']),
                   'dimension': 0, 'residual_estimate': None, 'subdivided': True}
