"""Online tSURFF on the device: channel projection of each surface frame and
coherent Volkov accumulation of the channel amplitudes.

Mirrors python/surface_contract.py:5-56 (``SurfaceContractor``) and
python/streaming_surface.py:26-123 (``directions``, ``energy_grid``,
``VolkovAccumulator``/``TorchVolkovAccumulator``).  Surface frames come from
``HipHamiltonian.flux_device`` and never leave the device; the ionic duals
chi(t) come from ``laqphysics_amd.ionic`` (host sparse LU, as in the reference).

Every contraction path of the reference (product channels, the coupled M=0
fast path, the sparse bipolar transform) is compiled once on the host into an
entry list (label j, outer channel o, surface column c, inner ionic state a,
dual column u, weight w); the device kernel evaluates
q[j, o, s] = sum_entries w * sum_r conj(chi[a, r, u]) frame[c, s, r].
"""
import ctypes as ct

import numpy as np
import torch
from scipy.sparse import issparse
from scipy.special import spherical_jn, sph_harm_y

from . import native
from .hip_backend import _stream, _vp


def directions(axial=True, theta_points=None, phi_points=None):
    """Gauss-Legendre theta x uniform phi quadrature (python/streaming_surface.py:26-31)."""
    nt = int(theta_points if theta_points is not None else (32 if axial else 24))
    nf = int(phi_points if phi_points is not None else (1 if axial else 32))
    if nt < 2 or nf < 1 or (not axial and nf < 2):
        raise ValueError('invalid angular quadrature')
    z, w = np.polynomial.legendre.leggauss(nt)
    t, p = np.meshgrid(np.arccos(z), np.arange(nf) * 2 * np.pi / nf, indexing='ij')
    return t.ravel(), p.ravel(), np.repeat(w, nf) * 2 * np.pi / nf


def energy_grid(segments=None, spectrum_energy=(.01, 1., 397)):
    """Positive photo-electron energies (python/streaming_surface.py:33-37)."""
    if segments:
        result = np.unique(np.concatenate([np.linspace(float(a), float(b), int(n)) for a, b, n in segments]))
    else:
        result = np.linspace(*spectrum_energy)
    if len(result) < 3 or np.any(result <= 0) or not np.isfinite(result).all():
        raise ValueError('invalid positive energy grid')
    return result


class SurfaceContractor:
    def __init__(self, h, ion, labels, unique, columns, folded, angular_transform=None, device='cuda:0'):
        self.h = h
        self.ion = ion
        self.labels = [tuple(x) for x in labels]
        self.unique = list(unique)
        self.columns = list(columns)
        self.folded = folded
        self.device = torch.device(device)

        def inner_key(lm):
            return (lm[0], abs(lm[1]) if folded else lm[1])
        active = [c for c, (inner, outer) in enumerate(h.ch) if inner_key(inner) in ion.index]
        self.outerstates = sorted({h.ch[c][1] for c in active})
        outerindex = {x: i for i, x in enumerate(self.outerstates)}
        self.shape = (len(self.labels), len(self.outerstates), len(h.surface_indices))
        sparse = angular_transform is not None and issparse(angular_transform)
        fast = bool(angular_transform is not None and folded and not sparse)
        entries = []   # (j, o, c, a, u, w)
        if fast:
            U = np.asarray(angular_transform)
            lookup = {pair: i for i, pair in enumerate(h.ch)}
            pairs = []
            for c in range(U.shape[1]):
                support = np.flatnonzero(abs(U[:, c]) > 1e-14)
                pair = {(h.ch[i][0][0], h.ch[i][1][0]) for i in support}
                if len(pair) != 1:
                    raise ValueError('coupled column has no unique radial angular pair')
                pairs.append(next(iter(pair)))
            for j, (n, l, m) in enumerate(self.labels):
                for c, (l1, l2) in enumerate(pairs):
                    pair = ((l1, m), (l2, -m))
                    if pair in lookup:
                        entries.append((j, outerindex[(l2, -m)], c, ion.index[(l1, abs(m) if folded else m)],
                                        self.columns[j], complex(U[lookup[pair], c])))
        elif angular_transform is None or sparse:
            if sparse:
                coo = angular_transform.tocoo()
                keep = np.array([inner_key(h.ch[row][0]) in ion.index for row in coo.row], bool)
                rows, cols, coefs = coo.row[keep], coo.col[keep], coo.data[keep]
            else:
                rows = np.array(active, int)
                cols = rows
                coefs = np.ones(len(rows))
            for row, col, coef in zip(rows, cols, coefs):
                inner, outer = h.ch[row]
                for j, label in enumerate(self.labels):
                    w = complex(coef)
                    if folded and inner[1] != label[2]:
                        w = 0j
                    entries.append((j, outerindex[outer], int(col), ion.index[inner_key(inner)],
                                    self.columns[j], w))
        else:
            raise ValueError('dense angular transform requires a folded (z-polarised) projection')
        nlab, nout = len(self.labels), len(self.outerstates)
        entries.sort(key=lambda e: (e[0] * nout + e[1]))
        keys = np.array([e[0] * nout + e[1] for e in entries], int)
        ptr = np.searchsorted(keys, np.arange(nlab * nout + 1)).astype(np.int32)
        dev = self.device
        self._ptr = torch.as_tensor(ptr, device=dev)
        self._c = torch.as_tensor(np.array([e[2] for e in entries], np.int32), device=dev)
        self._a = torch.as_tensor(np.array([e[3] for e in entries], np.int32), device=dev)
        self._u = torch.as_tensor(np.array([e[4] for e in entries], np.int32), device=dev)
        self._w = torch.as_tensor(np.array([e[5] for e in entries], complex), device=dev)
        self.entries = entries

    def __call__(self, chi, frame):
        """q (labels, outer, ns) on the device from chi (na*n, len(unique)) and a
        device frame (nc, ns, n) or a host frame (n, ns, nc)."""
        if isinstance(frame, np.ndarray):
            frame = torch.as_tensor(np.ascontiguousarray(frame.transpose(2, 1, 0)), device=self.device)
        frame = frame.contiguous()
        if isinstance(chi, np.ndarray):
            chi = torch.as_tensor(np.ascontiguousarray(chi), dtype=torch.complex128, device=self.device)
        nc, ns, n = frame.shape
        q = torch.empty(self.shape, dtype=torch.complex128, device=self.device)
        native.check(native.lib().hex_surface_contract(
            n, ns, len(self.unique), self.shape[0], self.shape[1], _vp(self._ptr), _vp(self._c), _vp(self._a),
            _vp(self._u), _vp(self._w), _vp(chi), _vp(frame), _vp(q), _stream(self.device)), 'surface contract')
        return q


class VolkovAccumulator:
    """Trapezoid accumulation of the Volkov-projected channel amplitudes, on the device."""

    def __init__(self, contractor, energy, angles, avec):
        h = contractor.h
        self.labels = contractor.labels
        self.avec = avec
        self.device = contractor.device
        self.energy = np.asarray(energy, float)
        self.k = np.sqrt(2 * self.energy)
        self.theta, self.phi, self.solid_angle_weights = angles
        self.unit = np.array([np.sin(self.theta) * np.cos(self.phi), np.sin(self.theta) * np.sin(self.phi),
                              np.cos(self.theta)]).T
        r = h.r[h.surface_indices].real
        w = h.grid.weights[h.surface_indices].real
        self.tests = np.array([np.sqrt(2 / np.pi) * (-1j) ** l * np.sqrt(w)[None, :] * r[None, :]
                               * spherical_jn(l, self.k[:, None] * r[None, :]) for l, m in contractor.outerstates])
        self.harmonics = np.array([sph_harm_y(l, m, self.theta, self.phi) for l, m in contractor.outerstates])
        dev = self.device
        t = lambda x, dtype=torch.complex128: torch.as_tensor(np.ascontiguousarray(x), dtype=dtype, device=dev)
        self._tests = t(self.tests)
        self._harm = t(self.harmonics)
        self._energy = t(self.energy, torch.float64)
        self._k = t(self.k, torch.float64)
        self._unit = t(self.unit, torch.float64)
        shape = (len(self.labels), len(self.energy), len(self.theta))
        self.amplitude = torch.zeros(shape, dtype=torch.complex128, device=dev)
        self.previous = torch.zeros_like(self.amplitude)
        self.intA = np.zeros(3)
        self.intA2 = 0.
        self.time = 0.
        self.index = -1
        self._nout = len(contractor.outerstates)
        self._ns = len(h.surface_indices)

    def add(self, index, time, q):
        if index != self.index + 1:
            raise ValueError('nonsequential spectrum samples')
        if self.index >= 0:
            d = time - self.time
            before = self.avec(self.time)
            after = self.avec(time)
            self.intA += d * (before + after) / 2
            self.intA2 += d * (np.dot(before, before) + np.dot(after, after)) / 2
        else:
            d = 0.
        if isinstance(q, np.ndarray):
            q = torch.as_tensor(q, device=self.device)
        q = q.contiguous()
        A = (ct.c_double * 3)(*self.intA)
        native.check(native.lib().hex_volkov_add(
            len(self.labels), self._nout, self._ns, len(self.energy), len(self.theta), _vp(q), _vp(self._tests),
            _vp(self._harm), _vp(self._energy), _vp(self._k), _vp(self._unit), float(time), A, float(self.intA2),
            float(d), int(self.index < 0), _vp(self.previous), _vp(self.amplitude), _stream(self.device)), 'volkov')
        self.time = float(time)
        self.index = int(index)

    def host_amplitude(self):
        return self.amplitude.cpu().numpy()

    def payload(self):
        """Host arrays that restore the accumulation after ``index``, under the
        keys of the reference's accumulator banks (python/streaming_surface.py:66-68)."""
        return {'amplitude': self.amplitude.cpu().numpy(), 'previous_integrand': self.previous.cpu().numpy(),
                'intA': self.intA.copy(), 'intA2': np.float64(self.intA2), 'time': np.float64(self.time),
                'sample_index': np.int64(self.index)}

    def restore(self, data):
        """Continue from a ``payload`` (this engine's or the reference's bank)."""
        amplitude = np.asarray(data['amplitude'])
        if amplitude.shape != tuple(self.amplitude.shape):
            raise ValueError(f'accumulator shape {amplitude.shape} does not match {tuple(self.amplitude.shape)}')
        self.amplitude.copy_(torch.as_tensor(amplitude, device=self.device))
        self.previous.copy_(torch.as_tensor(np.asarray(data['previous_integrand']), device=self.device))
        self.intA = np.array(data['intA'], float)
        self.intA2 = float(data['intA2'])
        self.time = float(data['time'])
        self.index = int(data['sample_index'])
