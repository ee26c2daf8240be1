"""Adjoint (costate) gradient of a propagation objective with respect to a
sampled control field.  No reference counterpart: the reference has no
gradients.  Written beside the plain propagator, which it reuses unchanged.

Control.  E(t) is piecewise linear through K samples theta_k * w_k at
t_k = k T/(K-1) (w_k = sin^2(pi t_k/T) when windowed), z-polarised, zero after
T; velocity gauge A(t) = -int_0^t E exactly (piecewise quadratic).

Forward.  psi_{n+1} = U_n psi_n, U_n = exp(-i dt H(A_n)), A_n = A(t_n + dt/2),
by the device Arnoldi of ``hip_backend`` (same tolerance and check cadence).
A step whose estimate stays above tol is split in two half steps at tol/2,
each with its own midpoint field, recursively (python/torch_backend.py:160-167);
every unsplit sub-step (a "leaf") is one factor U of the discrete map, and the
backward pass below runs over the leaves.  Checkpoints every
``checkpoint_every`` steps.

Backward.  Per leaf, the costate obeys lambda_n = U_n^dagger lambda_{n+1}, started from
the objective's gradient lambda_N = dJ/dpsi_N^*, and propagated with the
adjoint apply (H^dagger: the conjugated irECS contour).  The exact derivative
of the discrete step is the Frechet derivative of the exponential,
    dJ/dA_n = 2 Re <lambda_{n+1}, L(X_n, E_n) psi_n>,
    L(X, E) = int_0^1 e^{(1-s)X} E e^{sX} ds,  X = -i dt H, E = -i dt dH/dA,
evaluated in the two Krylov spaces already built by the recomputed forward step
(basis V, Hessenberg H_m) and by the costate step (basis W, Hessenberg G):
    <lambda, L psi> = conj(gamma) beta (-i dt) R_{00},
    R = upper-right block of expm([[-i dt G^H, W^H dH/dA V], [0, -i dt H_m]]).
dH/dA_z = D_z + 2 A_z with D_z the field-coupling part for a unit field
(``apply_terms`` with only the dipole edges).  Forward segments are recomputed
from their checkpoint on the way back, keeping their Krylov bases.

Gradient w.r.t. the samples: dJ/dtheta_k = sum over leaves dJ/dA_n dA_n/dtheta_k with
dA(t)/dtheta_k = -w_k int_0^t phi_k (phi_k the hat function of node k).
"""
import ctypes as ct

import numpy as np
import torch

from . import native
from .hip_backend import _check_buffers, _combine, _engine_for, _norm, _scale_into, _stream, _vp
from .ionic import C1, C2, Ionic
from .surface import SurfaceContractor, VolkovAccumulator, directions


class PiecewiseLinearControl:
    def __init__(self, theta, T, window=True):
        theta = np.asarray(theta, float)
        self.K = len(theta)
        if self.K < 2 or not T > 0:
            raise ValueError('need at least two samples and T > 0')
        self.T = float(T)
        self.tk = np.linspace(0., self.T, self.K)
        self.h = self.T / (self.K - 1)
        self.w = np.sin(np.pi * self.tk / self.T) ** 2 if window else np.ones(self.K)
        self.theta = theta

    def hat_integrals(self, t):
        """g_k(t) = int_0^t phi_k(s) ds for every node k (vector)."""
        h = self.h
        a = np.clip(t - (self.tk - h), 0., h)
        b = np.clip(t - self.tk, 0., h)
        left = a * a / (2 * h)
        right = b - b * b / (2 * h)
        left[0] = 0.
        right[-1] = 0.
        return left + right

    def electric(self, t):
        if t < 0 or t > self.T:
            return 0.
        return float(np.interp(t, self.tk, self.theta * self.w))

    def vector(self, t):
        return -float(np.dot(self.theta * self.w, self.hat_integrals(min(max(t, 0.), self.T))))

    def avec(self, t):
        return np.array([0., 0., self.vector(t)])

    def dA_dtheta(self, t):
        return -self.w * self.hat_integrals(min(max(t, 0.), self.T))


def _arnoldi(apply, x, dt, tol, maxdim, fixed=None):
    """Arnoldi exponential that also returns its basis and Hessenberg matrix.

    Same algorithm, decisions and device Krylov control as
    ``hip_backend.exponential_action`` (columns committed on the device, the
    small exponential and the tests replayed there at the check points, a
    breakdown between them included); returns (y, Q (maxdim+1, N) tensor, H_m (m x m, host), beta, m),
    or None when the estimate stays above ``tol`` up to ``maxdim``.  ``fixed``
    forces the Krylov dimension (finite-difference runs reuse the dimensions
    of the run they perturb, so the discrete map they difference is smooth):
    the only test besides a breakdown is then after column fixed - 1."""
    if maxdim > 62:
        raise ValueError('maxdim above 62 is not supported by the device Krylov kernels')
    e = _engine_for(x)
    N = x.numel()
    beta = _norm(e, x)
    Q = torch.empty((maxdim + 1, N), dtype=x.dtype, device=x.device)
    if beta == 0:
        return x.clone(), Q, np.zeros((0, 0), complex), 0., 0
    _scale_into(x, beta, Q[0])
    Hdev = torch.empty((maxdim, 64), dtype=x.dtype, device=x.device)
    Hfull, work, result = _check_buffers(x.device, maxdim)
    stream = _stream(x.device)
    lib = native.lib()
    synced = 0
    for j in range(maxdim):
        v = apply(Q[j])
        native.check(lib.hex_cgs2_commit(e._engine, _vp(Q), N, j + 1, _vp(v), N, _vp(Hdev[j]), _vp(Q[j + 1]),
                                         stream), 'cgs2')
        if fixed is not None:
            if not (j + 1 == fixed or j + 1 == maxdim):
                continue
            native.check(lib.hex_arnoldi_check_fixed(_vp(Hdev), synced, j, maxdim, int(fixed), float(dt),
                                                     float(beta), _vp(Hfull), _vp(work), _vp(result), stream),
                         'arnoldi check')
        else:
            if not (j >= 3 and (j + 1) % 4 == 0 or j + 1 == maxdim):
                continue
            native.check(lib.hex_arnoldi_check(_vp(Hdev), synced, j, maxdim, float(dt), float(beta), float(tol),
                                               _vp(Hfull), _vp(work), _vp(result), stream), 'arnoldi check')
        r = result[:2 + j + 1].cpu().numpy()
        synced = j + 1
        if r[0].real == 1:
            m = int(r[0].imag)
            y = torch.empty_like(x)
            _combine(e, Q, m, r[2:2 + m], N, y)
            # the device Hessenberg holds the gathered columns; below the
            # subdiagonal it is zero by definition
            Hm = np.triu(Hfull.view(maxdim + 1, maxdim)[:m, :m].cpu().numpy(), -1)
            return y, Q, Hm, beta, m
    return None


class SurfaceYield:
    """He+ channel yield from tSURFF as a function of the control samples
    theta and of the surface frames recorded in the forward run, with its
    exact gradient.

    Y = sum_{j, e, d} w_e k_e W_d |a_j(E_e, Omega_d)|^2 over the channels
    ``labels`` (He+(n=2) by default), with trapezoid weights w_e on the energy
    grid and the solid-angle weights W_d (k^2 dk = k dE).  The amplitudes are
    the ones the device tSURFF chain accumulates (``run.OnlineSurface``):
    surface frame F_n = flux(psi_n, A(t_n)) at every step, projection onto the
    ionic duals chi(t_n) (backward CF4 of the He+ channels, python/implicit.py:111-123
    with python/streaming_surface.py:198-214), Volkov phase and trapezoid in
    time (python/streaming_surface.py:39-64).  Value and gradient -- dY/dtheta
    through the duals, the frames' field and the Volkov phases, and dY/dF_n^*,
    which becomes a costate source at every step -- run on the device
    (hex_yield_value / hex_yield_gradient); the host forms only the control's
    A(t) at the samples and the chain rule to theta."""

    def __init__(self, problem, radius, labels=((2, 0, 0), (2, 1, 0)), energy=None, angles=None):
        h = problem.h
        base = getattr(h, 'uncoupled', h)
        h.prepare_surface(radius)
        if base is not h:
            base.prepare_surface(radius)
        self.problem = problem
        self.labels = [tuple(x) for x in labels]
        unique = list(dict.fromkeys((n, l, abs(m)) for n, l, m in self.labels))
        columns = [unique.index((n, l, abs(m))) for n, l, m in self.labels]
        self.ion = Ionic(base, {m for n, l, m in unique})
        con = SurfaceContractor(base, self.ion, self.labels, unique, columns, True, getattr(h, 'transform', None),
                                'cpu')
        energy = np.linspace(.2, 1.6, 15) if energy is None else np.asarray(energy, float)
        angles = directions(True, 8, 1) if angles is None else angles
        vol = VolkovAccumulator(con, energy, angles, None)
        self.shape = con.shape
        ent = con.entries
        i32 = lambda x: np.ascontiguousarray(x, dtype=np.int32)
        self.ej = i32([e[0] * con.shape[1] + e[1] for e in ent])
        self.ec = i32([e[2] for e in ent])
        self.ea = i32([e[3] for e in ent])
        self.eu = i32([e[4] for e in ent])
        self.ew = np.ascontiguousarray([e[5] for e in ent], dtype=complex)
        self.tests = np.ascontiguousarray(vol.tests, dtype=complex)
        self.harm = np.ascontiguousarray(vol.harmonics, dtype=complex)
        self.energy = np.ascontiguousarray(energy, dtype=float)
        self.kz = np.ascontiguousarray(vol.k[:, None] * vol.unit[None, :, 2], dtype=float)
        wE = np.gradient(energy) if len(energy) > 1 else np.ones(1)
        wE[0] = wE[-1] = (energy[1] - energy[0]) / 2 if len(energy) > 1 else 1.
        self.wy = np.ascontiguousarray(wE[:, None] * vol.k[:, None] * angles[2][None, :], dtype=float)
        # ionic matrices of the dual propagation (python/surface3d.py:8-48), dense
        self.h0 = np.ascontiguousarray(self.ion.h0.toarray().conj().T, dtype=complex)
        self.pz = np.ascontiguousarray(self.ion.P[2].toarray().conj().T, dtype=complex)
        self.chiN = np.ascontiguousarray(self.ion.dual_final_states(unique), dtype=complex)
        self.na, self.nr, self.nu = self.ion.na, self.ion.n, len(unique)
        # A(t) = G(t) . theta at the sample times and at the dual CF4 stages
        ctl = problem.control
        G = lambda t: -ctl.w * ctl.hat_integrals(min(max(t, 0.), ctl.T))
        N, dt, t0 = problem.nsteps, problem.dt, problem.t0
        self.times = t0 + dt * np.arange(N + 1)
        self.Gs = np.array([G(t) for t in self.times])                                  # (N+1, K)
        self.Gc = np.array([[G(t - C1 * dt), G(t - C2 * dt)] for t in self.times])    # (N+1, 2, K)
        self._dev = None
        self._valued = None

    def __del__(self):
        if getattr(self, '_dev', None) is not None and getattr(native, '_lib', None) is not None:
            native.lib().hex_yield_destroy(self._dev)
            self._dev = None

    def _device_yield(self, device):
        if self._dev is None:
            nlab, nout, ns = self.shape
            keep = (self.h0, self.pz, self.chiN, self.ej, self.ec, self.ea, self.eu, self.ew, self.tests, self.harm,
                    self.energy, self.kz, self.wy, self.times)
            ptr = lambda x: x.ctypes.data_as(ct.c_void_p)
            d = native.HexYieldDesc()
            d.M, d.nu, d.na, d.nr, d.N = self.h0.shape[0], self.nu, self.na, self.nr, len(self.times) - 1
            d.dt = float(self.problem.dt)
            d.h0, d.pz, d.chiN = ptr(self.h0), ptr(self.pz), ptr(self.chiN)
            d.nent, d.nlab, d.nout, d.ns, d.nc = len(self.ej), nlab, nout, ns, self.problem.h.nc
            d.ej, d.ec, d.ea, d.eu, d.ew = ptr(self.ej), ptr(self.ec), ptr(self.ea), ptr(self.eu), ptr(self.ew)
            d.nE, d.nD = len(self.energy), self.kz.shape[1]
            d.tests, d.harm, d.energy = ptr(self.tests), ptr(self.harm), ptr(self.energy)
            d.kz, d.wy, d.times = ptr(self.kz), ptr(self.wy), ptr(self.times)
            handle = ct.c_void_p()
            native.check(native.lib().hex_yield_create(device.index or 0, ct.byref(d), ct.byref(handle)), 'yield')
            del keep
            self._dev = handle
        return self._dev

    def _frames_ok(self, F0, Fd):
        S = len(self.times)
        shape = (S, self.problem.h.nc, self.shape[2], self.nr)
        for F in (F0, Fd):
            if tuple(F.shape) != shape or F.dtype != torch.complex128 or F.device.type != 'cuda' or \
                    not F.is_contiguous():
                raise ValueError(f'frames must be contiguous complex128 {shape} device tensors')

    def value(self, theta, F0, Fd):
        """Y from theta (K,) and the frames F0_n = the J part of flux(psi_n),
        Fd_n = its field-coupling part for the unit field e_z, both
        (N+1, nc, ns, n) device tensors."""
        self._frames_ok(F0, Fd)
        theta = np.asarray(theta, dtype=float)
        A = np.ascontiguousarray(self.Gs @ theta)
        a = np.ascontiguousarray(self.Gc @ theta)   # (N+1, 2)
        dev = self._device_yield(F0.device)
        Y = ct.c_double()
        native.check(native.lib().hex_yield_value(dev, A.ctypes.data_as(ct.c_void_p), a.ctypes.data_as(ct.c_void_p),
                                                  _vp(F0), _vp(Fd), ct.byref(Y), _stream(F0.device)), 'yield value')
        self._valued = (theta.copy(), F0.data_ptr(), Fd.data_ptr(), Y.value)
        return Y.value

    def gradient(self, theta, F0, Fd):
        """(Y, dY/dtheta, dY/dconj(F0), dY/dconj(Fd)); the last two are device
        tensors in the frames' layout."""
        theta = np.asarray(theta, dtype=float)
        v = self._valued
        if v is None or not np.array_equal(v[0], theta) or v[1] != F0.data_ptr() or v[2] != Fd.data_ptr():
            self.value(theta, F0, Fd)
        S = len(self.times)
        dA = np.empty(S)
        da = np.empty((S, 2))
        G0 = torch.empty_like(F0)
        Gd = torch.empty_like(Fd)
        native.check(native.lib().hex_yield_gradient(self._dev, _vp(F0), _vp(Fd), dA.ctypes.data_as(ct.c_void_p),
                                                     da.ctypes.data_as(ct.c_void_p), _vp(G0), _vp(Gd),
                                                     _stream(F0.device)), 'yield gradient')
        g = self.Gs.T @ dA + self.Gc[:, 0].T @ da[:, 0] + self.Gc[:, 1].T @ da[:, 1]
        return self._valued[3], g, G0, Gd

    def frames(self, engine, psi):
        """(F0, Fd) of one state on the device: the J rows of the surface
        source and its field-coupling rows for the unit field e_z."""
        return engine.flux_device(psi, (0., 0., 0.), terms=1), engine.flux_device(psi, (0., 0., 1.), terms=2)

    def source(self, g0, gd, n_state):
        """dY/dpsi^* (device state layout) from dY/dF0^*, dY/dFd^* of one frame
        (each (nc, ns, n) on the device): the adjoint of the frame map, on the
        device (hex_flux_adjoint)."""
        return self.problem.engine.flux_adjoint(g0.contiguous(), gd.contiguous())


class Problem:
    """A propagation from psi0 over nsteps under a control, with the objective
    J = 1 - <psi_N, W psi_N> (ionisation probability: W is the inner-region
    weight f_i f_j with f = interior weight / |weight| inside R, as the
    reference's real-region norm, python/run3d.py:216-219), plus optionally
    the tSURFF He+ channel yield (``add_yield``)."""

    def __init__(self, engine, psi0, control, dt, nsteps, t0=0., tol=1e-10, maxdim=10, checkpoint_every=64,
                 radius=None, ionization_weight=1.):
        self.engine = engine
        self.h = engine.h
        self.psi0 = psi0 if isinstance(psi0, torch.Tensor) else engine.state(psi0)
        self.control = control
        self.dt, self.nsteps, self.t0 = float(dt), int(nsteps), float(t0)
        self.tol, self.maxdim = float(tol), int(maxdim)
        self.every = int(checkpoint_every)
        g = self.h.grid
        f = np.asarray(g.interior_weights) / np.abs(g.weights)
        if radius is not None:
            f = np.where(self.h.r.real <= radius, f, 0.)
        n, nc = self.h.n, self.h.nc
        weight = (f[:, None] * f[None, :])[:, :, None] * np.ones(nc)[None, None, :]
        self.weight = torch.as_tensor(weight.ravel(order='F'), dtype=torch.complex128, device=engine.device)
        self.costate_maxdim = 2 * self.maxdim
        self.dims = [[] for _ in range(self.nsteps)]   # sub-steps (t, dt, Krylov dim) of the last forward run
        self.forced_dims = None                        # replay these sub-steps instead of adapting
        self.ionization_weight = float(ionization_weight)
        self.surface = None                            # optional SurfaceYield term of the objective
        self.yield_weight = 1.
        self.frames = None                             # (F0, Fd) of the last forward run, device tensors

    def add_yield(self, radius, labels=((2, 0, 0), (2, 1, 0)), energy=None, angles=None, weight=1.):
        """Add ``weight`` times the tSURFF He+ channel yield of ``labels`` to the objective."""
        self.surface = SurfaceYield(self, radius, labels, energy, angles)
        self.yield_weight = float(weight)
        return self.surface

    # objective and its gradient dJ/dpsi^*
    def objective(self, psi):
        """J = ionisation probability (1 - inner-region probability) plus, when
        a yield term is set, the He+ channel yield of the last forward run."""
        J = self.ionization_weight * float(1. - torch.vdot(psi, self.weight * psi).real)
        if self.surface is not None:
            J += self.yield_weight * self.surface.value(self.control.theta, *self.frames)
        return J

    def terminal_costate(self, psi):
        return -self.ionization_weight * (self.weight * psi)

    def _apply_at(self, t):
        A = self.control.avec(t)
        return lambda v: self.engine.apply(v, A, True)

    def _advance(self, psi, t, dt, tol, depth, leaves):
        """The reference's recursive step (python/torch_backend.py:160-167):
        on non-convergence, two half steps at tol/2, at most 8 levels deep."""
        out = _arnoldi(self._apply_at(t + dt / 2), psi, dt, tol, self.maxdim)
        if out is None:
            if depth >= 8:
                raise RuntimeError(f'GPU Arnoldi did not converge in the adjoint propagation '
                                   f'(subdivision deeper than 8): dim={self.maxdim}, dt={dt}')
            psi = self._advance(psi, t, dt / 2, tol / 2, depth + 1, leaves)
            return self._advance(psi, t + dt / 2, dt / 2, tol / 2, depth + 1, leaves)
        leaves.append((t, dt, out))
        return out[0]

    def step(self, n, psi):
        """One step; returns the new state and its leaves [(t, dt, arnoldi output)],
        one per unsplit sub-step, in time order."""
        leaves = []
        if self.forced_dims is not None:
            for t, dt, m in self.forced_dims[n]:
                out = _arnoldi(self._apply_at(t + dt / 2), psi, dt, self.tol, self.maxdim, fixed=m)
                leaves.append((t, dt, out))
                psi = out[0]
        else:
            psi = self._advance(psi, self.t0 + n * self.dt, self.dt, self.tol, 0, leaves)
        self.dims[n] = [(t, dt, out[4]) for t, dt, out in leaves]
        return psi, leaves

    def forward(self, keep=True):
        psi = self.psi0.clone()
        checkpoints = {0: psi.clone()} if keep else None
        frames = [self.surface.frames(self.engine, psi)] if self.surface is not None else None
        for n in range(self.nsteps):
            psi = self.step(n, psi)[0]
            if frames is not None:
                frames.append(self.surface.frames(self.engine, psi))
            if keep and (n + 1) % self.every == 0 and n + 1 < self.nsteps:
                checkpoints[n + 1] = psi.clone()
        if frames is not None:
            self.frames = tuple(torch.stack([f[k] for f in frames]) for k in (0, 1))
        return psi, checkpoints

    def gradient_A(self, psiN, checkpoints, sources=None):
        """dJ/dA at the midpoint of every sub-step, by the costate recursion
        (see module doc); returns [(t_mid, dJ/dA_z)].  ``sources(n)`` is the
        direct dJ/dpsi_n^* of a frame-dependent term, added to the costate at
        every step boundary."""
        eng = self.engine
        e = _engine_for(psiN)
        N = psiN.numel()
        grads = []
        lam = self.terminal_costate(psiN)
        if sources is not None:
            lam = lam + sources(self.nsteps)
        dv = torch.empty(N, dtype=torch.complex128, device=psiN.device)
        pending = []
        for s0 in reversed(sorted(checkpoints)):
            s1 = min(s0 + self.every, self.nsteps)
            # recompute the segment forward, keeping each sub-step's Krylov data
            psi = checkpoints[s0]
            saved = []
            for n in range(s0, s1):
                psi, leaves = self.step(n, psi)
                saved.append(leaves)
            for n in range(s1 - 1, s0 - 1, -1):
                lam = self._leaves_backward(saved[n - s0], lam, eng, e, N, dv, grads, pending)
                if sources is not None:
                    lam = lam + sources(n)
            self._flush_leaves(pending, grads, psiN.device)
            del saved
        return grads[::-1]

    def _leaves_backward(self, leaves, lam, eng, e, N, dv, grads, pending):
        """Costate through one step's leaves (latest first), appending a slot
        for each leaf's dJ/dA (filled by ``_flush_leaves``); returns the
        costate at the start of the step."""
        for t, dt, (_, Q, Hm, beta, m) in reversed(leaves):
            A = self.control.avec(t + dt / 2)
            # costate step lambda <- exp(+i dt H^dagger) lambda, in its own Krylov space
            out = _arnoldi(lambda v: eng.apply_adjoint(v, A, True), lam, -dt, self.tol, self.costate_maxdim)
            if out is None:
                raise RuntimeError(f'costate Arnoldi did not converge: dim={self.costate_maxdim}, dt={dt}')
            lam, W, G, gamma, mw = out
            if m == 0 or mw == 0:
                grads.append((t + dt / 2, 0.))
                continue
            # M = W^H (D + 2 A_z) V: one dot vector per column, left in device
            # memory until the segment's leaves are flushed together
            stream = _stream(dv.device)
            lib = native.lib()
            Mdev = self._leaf_dots(dv.device, len(pending))
            for j in range(m):
                eng.apply_terms(Q[j], (0., 0., 1.), True, terms=2, a=1., b=2. * A[2], out=dv)
                native.check(lib.hex_dots_device(e._engine, _vp(W), N, mw, _vp(dv), N, _vp(Mdev[j]), stream),
                             'dots')
            pending.append((len(grads), t + dt / 2, dt, G, Hm, np.conj(gamma) * beta * (-1j * dt), m, mw))
            grads.append(None)
            if len(pending) == self._LEAF_SLOTS:
                self._flush_leaves(pending, grads, dv.device)
        return lam

    _LEAF_SLOTS = 128

    def _leaf_dots(self, device, slot):
        """Device rows for the per-column dot vectors W^H dH/dA V of leaf ``slot``."""
        buf = getattr(self, '_mdots', None)
        if buf is None or buf.device != device:
            buf = self._mdots = torch.empty((self._LEAF_SLOTS, self.maxdim, self.costate_maxdim),
                                            dtype=torch.complex128, device=device)
        return buf[slot]

    def _flush_leaves(self, pending, grads, device):
        """dJ/dA of the pending leaves: <lambda, L(X, E) psi> = conj(gamma)
        beta (-i dt) R_00, R the upper-right block of expm([[-i dt G^H, M],
        [0, -i dt H_m]]); the augmented matrices, zero-padded to one order
        (the padding leaves their leading block's exponential unchanged), are
        exponentiated together on the device (hex_expm_batched)."""
        if not pending:
            return
        count = len(pending)
        P = max(m + mw for *_, m, mw in pending)
        Mh = self._mdots[:count].cpu().numpy()
        bigs = np.zeros((count, P, P), complex)
        for k, (_, _, dt, G, Hm, _, m, mw) in enumerate(pending):
            bigs[k, :mw, :mw] = -1j * dt * G.conj().T
            bigs[k, :mw, mw:mw + m] = Mh[k, :m, :mw].T
            bigs[k, mw:mw + m, mw:mw + m] = -1j * dt * Hm
        A = torch.as_tensor(bigs, device=device)
        E = torch.empty_like(A)
        work = torch.empty(6 * count * P * P, dtype=torch.complex128, device=device)
        native.check(native.lib().hex_expm_batched(count, P, _vp(A), _vp(E), _vp(work), _stream(device)), 'expm')
        Eh = E.cpu().numpy()
        for k, (slot, tmid, _, _, _, scale, m, mw) in enumerate(pending):
            grads[slot] = (tmid, 2. * (scale * Eh[k, 0, mw]).real)
        pending.clear()

    def gradient(self, psiN=None, checkpoints=None):
        """dJ/dtheta (length K) and J."""
        if psiN is None:
            psiN, checkpoints = self.forward()
        g = np.zeros(self.control.K)
        sources = None
        J = self.ionization_weight * float(1. - torch.vdot(psiN, self.weight * psiN).real)
        if self.surface is not None:
            # explicit field dependence and dY/dF^* on the device
            Y, gY, g0, gd = self.surface.gradient(self.control.theta, *self.frames)
            w = self.yield_weight
            J += w * Y
            g += w * gY
            if w != 1.:
                g0, gd = g0 * w, gd * w
            sources = lambda n: self.surface.source(g0[n], gd[n], n)
        for t, gA in self.gradient_A(psiN, checkpoints, sources):
            g += gA * self.control.dA_dtheta(t)
        return g, J


class Objective(torch.autograd.Function):
    """J(theta) as a differentiable torch function: forward propagation, and
    the costate gradient in ``backward``."""

    @staticmethod
    def forward(ctx, theta, problem):
        problem.control.theta = theta.detach().cpu().numpy().astype(float)
        psiN, checkpoints = problem.forward()
        ctx.problem = problem
        ctx.saved = (psiN, checkpoints)
        return torch.tensor(problem.objective(psiN), dtype=torch.float64)

    @staticmethod
    def backward(ctx, grad_output):
        problem = ctx.problem
        psiN, checkpoints = ctx.saved
        g, _ = problem.gradient(psiN, checkpoints)
        return torch.as_tensor(g * float(grad_output), dtype=torch.float64), None
