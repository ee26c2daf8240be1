"""ORACLE -- tSURFF surface source and Volkov accumulation (test infrastructure only).

``flux`` restates python/helium3d.py:69-77: out[i, s, c] = sum_k J[s, k] psi(i, k, c)
plus, for electron-2 dipole edges, coef * sum_k JP[s, k] psi(i, k, dc) with the
surface rows J = [T, Theta] and JP = i [T, Theta] o (r_k - r_s) prepared by
``prepare_surface`` (python/helium3d.py:52-67).
``Contractor`` restates python/surface_contract.py:5-56, ``Volkov`` the online
trapezoid of python/streaming_surface.py:39-64 and ``integrate`` the closed-form
reverse trapezoid of python/surface_projection.py:126-144.

The ionic duals chi(t) fed to ``Contractor`` come from
laqphysics_amd.ionic.backward_duals on both sides of every comparison.  Both
of its variants are pinned through the reference's recorded spectra of whole
runs: the exact-exponential one (python/surface3d.py:101-110) by
tests/golden/3d_dt_refined, the CF4 one (python/implicit.py:111-123) by
tests/golden/res_l8_L2_p6.
"""
import numpy as np


def flux(h, psi, field):
    u = np.asarray(psi).reshape(h.shape, order='F')
    out = np.einsum('ij,caj->cia', h.surface_T, u.transpose(2, 0, 1)).transpose(2, 1, 0)
    for a in range(h.nc):
        for k in range(h.dptr[a], h.dptr[a + 1]):
            if h.de[k] != 2:
                continue
            coef = np.dot(field, h.dcoef[:, k])
            if abs(coef) > 1e-15:
                out[:, :, a] += coef * (u[:, :, h.dc[k]] @ h.surface_P.T)
    return out


class Contractor:
    """Restates python/surface_contract.py:5-56 (all three contraction paths)."""

    def __init__(self, h, ion, labels, unique, columns, folded, angular_transform=None):
        from scipy.sparse import issparse
        self.h, self.ion, self.labels, self.unique, self.columns, self.folded = h, ion, labels, unique, columns, folded
        key = lambda lm: (lm[0], abs(lm[1]) if folded else lm[1])
        active = [c for c, (inner, outer) in enumerate(h.ch) if key(inner) in ion.index]
        self.outerstates = sorted({h.ch[c][1] for c in active})
        outerindex = {x: i for i, x in enumerate(self.outerstates)}
        self.shape = (len(labels), len(self.outerstates), len(h.surface_indices))
        self.sparse = angular_transform is not None and issparse(angular_transform)
        self.fast = bool(angular_transform is not None and folded and not self.sparse)
        self.transform = angular_transform
        if self.fast:
            U = angular_transform
            lookup = {pair: i for i, pair in enumerate(h.ch)}
            pairs = []
            for c in range(U.shape[1]):
                support = np.flatnonzero(abs(U[:, c]) > 1e-14)
                pairs.append(next(iter({(h.ch[i][0][0], h.ch[i][1][0]) for i in support})))
            self.coeff = np.zeros((len(labels), U.shape[1]))
            self.inner = np.zeros_like(self.coeff, dtype=int)
            self.groups = []
            for j, (n, l, m) in enumerate(labels):
                group = np.full(U.shape[1], -1, int)
                for c, (l1, l2) in enumerate(pairs):
                    pair = ((l1, m), (l2, -m))
                    if pair in lookup:
                        self.coeff[j, c] = U[lookup[pair], c]
                        self.inner[j, c] = ion.index[(l1, abs(m) if folded else m)]
                        group[c] = outerindex[(l2, -m)]
                self.groups.append(group)
        else:
            self.coefficients = None
            self.coupled_columns = None
            if self.sparse:
                coo = angular_transform.tocoo()
                valid = np.array([key(h.ch[row][0]) in ion.index for row in coo.row])
                active = coo.row[valid]
                self.coupled_columns = coo.col[valid]
                self.coefficients = coo.data[valid]
            self.active = active
            self.inner = np.array([ion.index[key(h.ch[c][0])] for c in active])
            self.out = np.array([outerindex[h.ch[c][1]] for c in active])
            self.order = np.argsort(self.out, kind='stable')
            self.starts = np.r_[0, np.flatnonzero(np.diff(self.out[self.order])) + 1]
            self.mask = np.array([[h.ch[c][0][1] == lab[2] for c in active] for lab in labels]) if folded else None

    def __call__(self, chi, frame):
        v = chi.reshape(self.ion.na, self.h.n, len(self.unique))
        result = np.zeros(self.shape, complex)
        if self.fast:
            weights = np.array([v[self.inner[j], :, self.columns[j]].T.conj() * self.coeff[j][None, :]
                                for j in range(len(self.labels))])
            projected = np.einsum('jrc,rsc->jsc', weights, frame)
            for j, group in enumerate(self.groups):
                for outer in np.unique(group[group >= 0]):
                    result[j, outer] = projected[j][:, group == outer].sum(axis=1)
        else:
            current = frame[:, :, self.coupled_columns] if self.sparse else frame[:, :, self.active]
            weights = v[self.inner][:, :, self.columns].conj().transpose(1, 2, 0)
            if self.mask is not None:
                weights = weights * self.mask[None, :, :]
            projected = np.einsum('rja,rsa->jsa', weights, current)
            if self.coefficients is not None:
                projected = projected * self.coefficients[None, None, :]
            reduced = np.add.reduceat(projected[:, :, self.order], self.starts, axis=2).transpose(0, 2, 1)
            result[:, self.out[self.order][self.starts], :] = reduced
        return result


def volkov_tables(h, outerstates, energy, theta, phi):
    from scipy.special import spherical_jn, sph_harm_y
    k = np.sqrt(2 * np.asarray(energy))
    r = h.r[h.surface_indices].real
    w = h.grid.weights[h.surface_indices].real
    tests = np.array([np.sqrt(2 / np.pi) * (-1j) ** l * np.sqrt(w)[None, :] * r[None, :]
                      * spherical_jn(l, k[:, None] * r[None, :]) for l, m in outerstates])
    harmonics = np.array([sph_harm_y(l, m, theta, phi) for l, m in outerstates])
    unit = np.array([np.sin(theta) * np.cos(phi), np.sin(theta) * np.sin(phi), np.cos(theta)]).T
    return k, tests, harmonics, unit


class Volkov:
    """Restates python/streaming_surface.py:39-64 (online trapezoid accumulation)."""

    def __init__(self, h, outerstates, nlabels, energy, theta, phi, avec):
        self.energy = np.asarray(energy)
        self.k, self.tests, self.harmonics, self.unit = volkov_tables(h, outerstates, energy, theta, phi)
        self.avec = avec
        self.amplitude = np.zeros((nlabels, len(energy), len(theta)), complex)
        self.previous = np.zeros_like(self.amplitude)
        self.intA = np.zeros(3)
        self.intA2 = 0.
        self.time = 0.
        self.index = -1

    def add(self, index, time, q):
        if self.index >= 0:
            d = time - self.time
            before, after = self.avec(self.time), self.avec(time)
            self.intA += d * (before + after) / 2
            self.intA2 += d * (np.dot(before, before) + np.dot(after, after)) / 2
        else:
            d = 0.
        radial = np.einsum('jls,les->jel', q, self.tests)
        phase = np.exp(1j * (self.energy[:, None] * time + self.k[:, None] * (self.unit @ self.intA)[None, :]
                             + self.intA2 / 2))
        current = 1j * np.sqrt(2) * (radial @ self.harmonics) * phase[None, :, :]
        if self.index >= 0:
            self.amplitude += d * (self.previous + current) / 2
        self.previous = current
        self.time = float(time)
        self.index = int(index)


def integrate(projected, h, outerstates, times, avec, energy, theta, phi):
    """Closed-form reverse trapezoid over a stored projected history
    (python/surface_projection.py:126-144)."""
    from scipy.integrate import cumulative_trapezoid
    k, tests, harmonics, unit = volkov_tables(h, outerstates, energy, theta, phi)
    av = np.array([avec(t) for t in times])
    intA = cumulative_trapezoid(av, times, axis=0, initial=0)
    intA2 = cumulative_trapezoid(np.sum(av * av, axis=1), times, initial=0)
    quad = np.empty(len(times))
    quad[0] = (times[1] - times[0]) / 2
    quad[-1] = (times[-1] - times[-2]) / 2
    quad[1:-1] = (times[2:] - times[:-2]) / 2
    b = np.zeros((projected.shape[1], len(energy), len(theta)), complex)
    for i in range(len(times) - 1, -1, -1):
        radial = np.einsum('jls,les->jel', projected[i], tests)
        phase = np.exp(1j * (np.asarray(energy)[:, None] * times[i] + k[:, None] * (unit @ intA[i])[None, :]
                             + intA2[i] / 2))
        b += 1j * np.sqrt(2) * quad[i] * (radial @ harmonics) * phase[None, :, :]
    return b
