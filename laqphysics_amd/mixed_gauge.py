"""Mixed length/velocity gauge with a compact radial gauge function.

Mirror of python/mixed_gauge.py.  With psi_g = exp(i A.r f(r)) psi_V and
q = r (1 - f), the one-electron terms are

    H0 + A.i[H0, q rhat] + E.r f rhat + 1/2 [A^2 (q/r)^2 + (q'^2 - (q/r)^2)(A.rhat)^2],

f the smoothstep switch from 1 (inside ``inner``) to 0 (beyond ``outer``);
the electron-electron potential is unchanged, and beyond the switch the
velocity-gauge surface flux holds.  The field enters through 12 coefficients
(``parameters``): A, E and the six A_a A_b products (off-diagonal ones doubled),
each an independent affine coefficient of the operator (so CF4 combines them
linearly and adds no separate A^2 phase: ``a2_coefficient = 0``).

Per electron e the two-electron operator adds to the field-free H
(python/mixed_gauge.py:96-117):
  * i (sum_a A_a C_a) (T q - q T)_e psi          -- the velocity dipole edges with
    q in place of r in the radial walk,
  * (i/2) (sum_a A_a C_a ldiff) q/r^2 psi        -- and q/r^2 in place of 1/r,
  * (sum_a E_a C_a) (r f)_e psi                   -- a length-gauge coupling over the
    same edges with r f in place of r,
  * 1/2 |A|^2 c0(r_e) psi, c0 = (1 - f)^2,
  * 1/2 (sum_k Q_k S_k) c2(r_e) psi, c2 = q'^2 - (1 - f)^2, with S_k the
    projected rhat_a rhat_b channel matrices of that electron.
On the device the field-free passes, the q walks and all channel couplings
run in the engine's two H.psi passes (``hex_apply_mixed``).
"""
import ctypes as ct

import numpy as np
import torch
from scipy.sparse import csr_matrix, diags, eye, identity, kron

from . import native
from .angular import C
from .ionic import Ionic, channel_dipole

PAIRS = ((0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2))   # order of the six quadratic coefficients


def profile(r, inner, outer):
    """(q, r f, c0, c2) on the grid, for the gauge switch f (python/mixed_gauge.py:19-27).

    f = 1 - u^3 (10 - 15 u + 6 u^2), u = clip((Re r - inner)/(outer - inner)),
    falls from 1 to 0 with f' = -30 u^2 (1 - u)^2 / (outer - inner);
    q = r (1 - f), c0 = (1 - f)^2 and c2 = q'^2 - c0 with q' = 1 - f - r f'.
    ``inner == outer == 0`` is the pure velocity gauge (f = 0)."""
    r = np.asarray(r)
    if inner == 0 and outer == 0:
        f = np.zeros(len(r))
        df = np.zeros(len(r))
    else:
        if not 0 <= inner < outer:
            raise ValueError('the gauge switch needs 0 <= inner < outer')
        width = outer - inner
        u = np.clip((r.real - inner) / width, 0.0, 1.0)
        f = 1.0 - u ** 3 * (10.0 - 15.0 * u + 6.0 * u * u)
        df = -30.0 * (u * (1.0 - u)) ** 2 / width
    keep = 1.0 - f
    dq = keep - r * df
    return r * keep, r * f, keep ** 2, dq ** 2 - keep ** 2


def parameters(avec, efunc, t):
    """The 12 affine field coefficients at t: A (3), E (3) and the six products
    A_a A_b of ``PAIRS``, off-diagonal ones counted twice (python/mixed_gauge.py:29-32)."""
    A = np.asarray(avec(t), float)
    E = np.asarray(efunc(t), float)
    a, b = np.array(PAIRS).T
    return np.concatenate([A, E, A[a] * A[b] * np.where(a == b, 1.0, 2.0)])


def _rank2_cartesian(C2):
    """The six rhat_a rhat_b of ``PAIRS`` from the rank-2 tensor C^2_q (q = -2..2)
    and the identity: rhat_a rhat_b = delta_ab / 3 + (traceless part in C^2)."""
    m2, m1, z0, p1, p2 = (C2[q] for q in (-2, -1, 0, 1, 2))
    one = C2['identity']
    s6 = np.sqrt(6.0)
    return [one / 3 - z0 / 3 + (p2 + m2) / s6,
            one / 3 - z0 / 3 - (p2 + m2) / s6,
            one / 3 + 2 * z0 / 3,
            -1j * (p2 - m2) / s6,
            (m1 - p1) / s6,
            1j * (m1 + p1) / s6]


def angular_squares(states):
    """rhat_a rhat_b (``PAIRS`` order) between one-electron states (l, m): (6, na, na)."""
    na = len(states)
    C2 = {q: np.array([[C(a, 2, q, b) for b in states] for a in states], complex) for q in range(-2, 3)}
    C2['identity'] = np.eye(na)
    squares = np.array(_rank2_cartesian(C2))
    if np.abs(squares[0] + squares[1] + squares[2] - np.eye(na)).max() > 2e-12:
        raise RuntimeError('rhat_x^2 + rhat_y^2 + rhat_z^2 is not the identity on these states')
    return squares


def two_electron_squares(h):
    """[electron 1, electron 2] lists of the six channel matrices of rhat_a rhat_b
    on that electron (python/mixed_gauge.py:43-71).  Product basis: Kronecker
    factors restricted to the channels; coupled basis: the same seen through its
    Clebsch-Gordan isometry; bipolar basis: from its rank-2 channel tensors."""
    result = []
    if hasattr(h, 'tensor_matrix'):
        for electron in (1, 2):
            C2 = {q: h.tensor_matrix(electron, 2, q) for q in range(-2, 3)}
            C2['identity'] = identity(h.nc, format='csr')
            result.append([csr_matrix(S) for S in _rank2_cartesian(C2)])
        return result
    base = getattr(h, 'uncoupled', h)
    lmax, M = base.angular_key
    single = [(l, m) for l in range(lmax + 1) for m in range(-l, l + 1)]
    pos = {s: i for i, s in enumerate(single)}
    take = np.array([pos[a] * len(single) + pos[b] for a, b in base.ch])
    one = angular_squares(single)
    eye1 = identity(len(single), format='csr')
    U = getattr(h, 'transform', None)
    for electron in (1, 2):
        matrices = []
        for S in one:
            S = csr_matrix(S)
            full = kron(S, eye1) if electron == 1 else kron(eye1, S)
            Q = csr_matrix(full.tocsr()[take][:, take])
            if U is not None:
                Q = csr_matrix(U.T @ (Q @ U))
            Q.data[np.abs(Q.data) < 1e-13] = 0.0
            Q.eliminate_zeros()
            matrices.append(Q)
        result.append(matrices)
    return result


def square_edges(h):
    """The angular-square couplings as edge lists: per output channel, (source,
    electron, six coefficients), in row order of the union pattern."""
    ptr, src, ele, coef = [0], [], [], []
    per = two_electron_squares(h)
    patterns = []
    for matrices in per:
        pattern = sum(abs(Q) for Q in matrices).tocsr()
        pattern.eliminate_zeros()
        pattern.sort_indices()
        dense = [Q.toarray() for Q in matrices]
        patterns.append((pattern, dense))
    for a in range(h.nc):
        for e, (pattern, dense) in enumerate(patterns):
            for b in pattern.indices[pattern.indptr[a]:pattern.indptr[a + 1]]:
                src.append(int(b))
                ele.append(e + 1)
                coef.append([complex(D[a, b]) for D in dense])
        ptr.append(len(src))
    return (np.array(ptr, np.int32), np.array(src, np.int32), np.array(ele, np.int32),
            np.array(coef, complex).reshape(-1, 6))


class MixedIonic(Ionic):
    """He+ Hamiltonian in the mixed gauge (python/mixed_gauge.py:119-140):
    ``matrix(c) = h0 + sum_k c_k G_k`` over the 12 coefficients, with, for
    ``Bq = [T, diag(q)]`` and the channel dipole ``C_x``,

        G_x (A)  = i [C_x (x) Bq + (C_x o Lam) (x) diag(q / 2r^2)],
        G_x (E)  = C_x (x) diag(r f),
        G_ab     = 1/2 [S_ab (x) diag(c2) + delta_ab 1 (x) diag(c0)].
    """

    def __init__(self, h, magnetic_numbers=None, inner=2., outer=4.):
        super().__init__(h, magnetic_numbers)
        r = np.asarray(h.r)
        q, F, c0, c2 = profile(r, inner, outer)
        T = h.grid.kinetic.tocsr()
        Bq = csr_matrix(T @ diags(q) - diags(q) @ T)
        ls = np.array([l for l, _ in self.states])
        lam = (ls * (ls + 1))[:, None] - (ls * (ls + 1))[None, :]
        hq = diags(q / (2 * r * r))
        dip = channel_dipole(self.states)
        self.Pq = [csr_matrix(1j * (kron(Cx, Bq) + kron(Cx.multiply(lam), hq))) for Cx in dip]
        self.Df = [csr_matrix(kron(Cx, diags(F))) for Cx in dip]
        squares = angular_squares(self.states)
        radial0 = kron(eye(self.na), diags(c0), format='csr')
        self.Q = [0.5 * (kron(csr_matrix(S), diags(c2), format='csr') + (radial0 if k < 3 else 0 * radial0))
                  for k, S in enumerate(squares)]

    def matrix(self, field):
        if len(field) != 12:
            raise ValueError('the mixed-gauge ionic operator takes 12 field coefficients')
        H = self.h0.copy()
        for c, G in zip(field, (*self.Pq, *self.Df, *self.Q)):
            if c:
                H = H + c * G
        return H.tocsr()


class MixedHamiltonian:
    """Device operator in the mixed gauge (python/mixed_gauge.py:73-117): the
    engine of ``HipHamiltonian`` built with q = r (1 - f) for its radial
    field-coupling walks, plus the gauge's tables."""
    a2_coefficient = 0.  # A_a A_b are independent affine coefficients in CF4

    def __init__(self, h, device='cuda:0', inner=2., outer=4., **kwargs):
        from .hip_backend import HipHamiltonian
        q, F, c0, c2 = profile(h.r, inner, outer)
        self.inner, self.outer = inner, outer
        self.base = HipHamiltonian(h, device, bq=q, **kwargs)
        self.h = h
        self.device = self.base.device
        self.n, self.nc, self.size = h.n, h.nc, int(h.size)
        self.coulomb_backend = self.base.coulomb_backend
        self.coulomb_block_bytes = 0
        r = np.asarray(h.r, complex)
        self._keep = {
            'hq': np.ascontiguousarray(q / (2 * r * r), complex), 'F': np.ascontiguousarray(F, complex),
            'c0': np.ascontiguousarray(c0, float), 'c2': np.ascontiguousarray(c2, complex),
        }
        sptr, ssrc, sele, scoef = square_edges(h)
        self._keep.update(sptr=sptr, ssrc=ssrc, sele=sele, scoef=np.ascontiguousarray(scoef))
        k = self._keep
        native.check(native.lib().hex_engine_set_mixed(
            self.base._engine, k['hq'].ctypes.data_as(ct.c_void_p), k['F'].ctypes.data_as(ct.c_void_p),
            k['c0'].ctypes.data_as(ct.c_void_p), k['c2'].ctypes.data_as(ct.c_void_p), len(ssrc),
            sptr.ctypes.data_as(ct.c_void_p), ssrc.ctypes.data_as(ct.c_void_p), sele.ctypes.data_as(ct.c_void_p),
            k['scoef'].ctypes.data_as(ct.c_void_p)), 'mixed gauge tables')
        self.separable = None

    # tensor plumbing of the engine it wraps
    def state(self, x):
        return self.base.state(x)

    def host(self, x):
        return self.base.host(x)

    def tensor(self, x):
        return self.base.tensor(x)

    def synchronize(self):
        self.base.synchronize()

    def norm(self, x):
        return self.base.norm(x)

    @staticmethod
    def fields(field):
        if len(field) != 12:
            raise ValueError('mixed gauge requires A, E and six quadratic coefficients')
        return tuple(float(v) for v in field)

    @torch.no_grad()
    def apply_affine(self, x, field, velocity, a, b, out=None):
        """out = a H(field) x + b x for the 12 coefficients ``field``."""
        from .hip_backend import _stream, _vp
        x = x.reshape(-1)
        self.base._check_state(x)
        if out is None:
            out = torch.empty_like(x)
        f = (ct.c_double * 12)(*self.fields(field))
        ca = (ct.c_double * 2)(complex(a).real, complex(a).imag)
        cb = (ct.c_double * 2)(complex(b).real, complex(b).imag)
        native.check(native.lib().hex_apply_mixed(self.base._engine, f, ca, cb, _vp(x), _vp(out),
                                                  _stream(self.device)), 'apply_mixed')
        return out

    def apply(self, x, field, velocity=True, out=None):
        return self.apply_affine(x, field, velocity, 1., 0., out)

    def flux_device(self, x, field, out=None):
        """Surface source: beyond the switch the velocity-gauge flux holds, with A = field[:3]."""
        return self.base.flux_device(x, tuple(field[:3]), out)

    def flux(self, x, field):
        return self.base.flux(x, tuple(field[:3]))
