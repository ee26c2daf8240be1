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
from scipy.sparse import bmat, csr_matrix, diags, eye, kron

from . import native
from .angular import cartesian
from .ionic import Ionic

PAIRS = ((0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2))


def profile(r, inner, outer):
    """(q, r f, c0, c2) on the grid (python/mixed_gauge.py:19-27)."""
    r = np.asarray(r)
    x = r.real
    if inner == outer == 0:
        f = np.zeros(len(r))
        fp = f.copy()
    else:
        if not 0 <= inner < outer:
            raise ValueError('mixed gauge needs 0 <= inner < outer')
        s = np.clip((x - inner) / (outer - inner), 0, 1)
        f = 1 - 10 * s ** 3 + 15 * s ** 4 - 6 * s ** 5
        fp = (-30 * s * s + 60 * s ** 3 - 30 * s ** 4) / (outer - inner)
    q = r * (1 - f)
    qp = 1 - f - r * fp
    return q, r * f, (1 - f) ** 2, qp ** 2 - (1 - f) ** 2


def parameters(avec, efunc, t):
    """The 12 field coefficients [A, E, A_a A_b (doubled off the diagonal)] at t
    (python/mixed_gauge.py:29-32)."""
    A = np.asarray(avec(t))
    E = np.asarray(efunc(t))
    quadratic = [A[a] * A[b] * (1 if a == b else 2) for a, b in PAIRS]
    return np.r_[A, E, quadratic]


def angular_squares(states):
    """Projected rhat_a rhat_b on the one-electron states, with the lmax+1
    intermediate shell included (python/mixed_gauge.py:34-41)."""
    top = max(l for l, m in states) + 1
    extended = [(l, m) for l in range(top + 1) for m in range(-l, l + 1)]
    N = np.array([[cartesian(a, b) for b in extended] for a in states]).transpose(2, 0, 1)
    squares = np.array([.5 * (N[a] @ N[b].conj().T + N[b] @ N[a].conj().T) for a, b in PAIRS])
    if np.max(abs(sum(squares[:3]) - np.eye(len(states)))) > 2e-12:
        raise RuntimeError('incomplete angular quadratic operator')
    return squares


def two_electron_squares(h):
    """[electron 1, electron 2] lists of the six channel matrices S_k
    (python/mixed_gauge.py:43-71); product and coupled channel bases."""
    if hasattr(h, 'tensor_matrix') or (hasattr(h, 'basis') and len(h.basis[0]) == 4):
        raise ValueError('the mixed gauge covers the product and coupled channel bases')
    base = getattr(h, 'uncoupled', h)
    states = sorted({a for a, b in base.ch} | {b for a, b in base.ch})
    lookup = {a: i for i, a in enumerate(states)}
    one = angular_squares(states)
    result = []
    for electron in (0, 1):
        rows, cols, values = [], [], []
        for a, bra in enumerate(base.ch):
            for b, ket in enumerate(base.ch):
                if bra[1 - electron] != ket[1 - electron]:
                    continue
                value = one[:, lookup[bra[electron]], lookup[ket[electron]]]
                if np.max(abs(value)) > 1e-14:
                    rows.append(a)
                    cols.append(b)
                    values.append(value)
        values = np.array(values)
        matrices = []
        for k in range(6):
            Q = csr_matrix((values[:, k], (rows, cols)), shape=(base.nc, base.nc))
            Q.eliminate_zeros()
            if hasattr(h, 'transform'):
                Q = csr_matrix(h.transform.T @ (Q @ h.transform))
            Q.data[abs(Q.data) < 1e-13] = 0.
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
    """He+ Hamiltonian in the mixed gauge (python/mixed_gauge.py:119-140)."""

    def __init__(self, h, magnetic_numbers=None, inner=2., outer=4.):
        super().__init__(h, magnetic_numbers)
        q, F, c0, c2 = profile(h.r, inner, outer)
        T = h.grid.kinetic.tocsr()
        B = T @ diags(q) - diags(q) @ T
        blocksP = [[[csr_matrix((self.n, self.n), dtype=complex) if a == b else None for b in range(self.na)]
                    for a in range(self.na)] for _ in range(3)]
        blocksD = [[[csr_matrix((self.n, self.n), dtype=complex) if a == b else None for b in range(self.na)]
                    for a in range(self.na)] for _ in range(3)]
        for a, (l, m) in enumerate(self.states):
            for b, (ll, mm) in enumerate(self.states):
                coeff = cartesian((l, m), (ll, mm))
                P = 1j * (B + diags((l * (l + 1) - ll * (ll + 1)) * q / (2 * h.r * h.r)))
                for k in range(3):
                    if abs(coeff[k]) > 1e-14:
                        blocksP[k][a][b] = coeff[k] * P
                        blocksD[k][a][b] = coeff[k] * diags(F)
        self.Pq = [bmat(x, format='csr') for x in blocksP]
        self.Df = [bmat(x, format='csr') for x in blocksD]
        square = angular_squares(self.states)
        self.Q = [.5 * (kron(csr_matrix(S), diags(c2), format='csr')
                        + (kron(eye(self.na), diags(c0), format='csr') if k < 3
                           else csr_matrix(self.h0.shape, dtype=complex)))
                  for k, S in enumerate(square)]

    def matrix(self, field):
        if len(field) != 12:
            raise ValueError('mixed ionic gauge requires 12 coefficients')
        H = self.h0.copy()
        for coefficient, operator in zip(field, [*self.Pq, *self.Df, *self.Q]):
            if coefficient:
                H = H + coefficient * operator
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
