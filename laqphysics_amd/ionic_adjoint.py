"""He+ ionic duals propagated on the device.

Mirror of python/ionic_adjoint.py:12-52 (``IonicAdjoint``: the multi-column
adjoint ionic operator with its field-free resolvent preconditioner) and of
the device branch of python/streaming_surface.py:125-146 (``AdjointStepper``
with ``ionic_device`` cuda:0) and python/streaming_surface.py:171-243
(endpoints of one backward sweep, optionally kept in the reference's
preparation files, then block replay).

The operator G(f) = H_ion(f)^dagger acts on nrhs dual columns at once
(``hex_ionic_apply``); the CF4-Pade steps use this package's GMRES
(``implicit.cf4_step``) at the reference's linear tolerance 1e-12.  The
preconditioner is the exact resolvent of the field-free block in per-l
eigenbases, as the reference's, but in complex128 with the non-flexible
solve (the reference's complex64 + FGMRES reaches the same tolerance).
"""
import ctypes as ct
import json
from pathlib import Path
from types import SimpleNamespace

import numpy as np
import torch
from scipy.linalg import eig, inv

from . import implicit, native
from .hip_backend import _stream, _vp
from .storage import DEFAULT_FILE_BYTES, ShardedArray, atomic_json


class IonicAdjoint:
    """Device adjoint ionic operator on ``nrhs`` dual columns (python/ionic_adjoint.py:12-52).

    Coefficients: the velocity gauge's (A_x, A_y, A_z, |A|^2 / 2) or the mixed
    gauge's 12 (``coefficients``); quadratic coefficients are independent
    affine coefficients, so CF4 combines them as the sparse-LU step does and
    adds no separate A^2 phase (``a2_coefficient = 0``)."""
    a2_coefficient = 0.

    def __init__(self, ion, nrhs, device='cuda:0'):
        self.device = torch.device(device)
        if self.device.type != 'cuda':
            raise ValueError('the device ionic adjoint needs a cuda device')
        if not torch.cuda.is_available():
            raise RuntimeError('the device ionic adjoint needs a GPU')
        self.ion = ion
        self.n, self.na, self.nrhs = ion.n, ion.na, int(nrhs)
        self.ndof = self.n * self.na
        self.size = self.ndof * self.nrhs
        self.mixed = hasattr(ion, 'Pq')
        matrices = [ion.h0, *ion.Pq, *ion.Df, *ion.Q] if self.mixed else [ion.h0, *ion.P, ion.identity]
        matrices = [x.conj().T.tocsr() for x in matrices]
        pattern = sum(abs(x) for x in matrices).tocsr()
        pattern.sort_indices()
        rows = np.repeat(np.arange(self.ndof), np.diff(pattern.indptr))
        cols = pattern.indices
        self.nk = len(matrices) - 1
        self._ptr = np.ascontiguousarray(pattern.indptr, np.int32)
        self._col = np.ascontiguousarray(cols, np.int32)
        data = np.ascontiguousarray([np.asarray(x[rows, cols]).ravel() for x in matrices], complex)
        handle = ct.c_void_p()
        with torch.cuda.device(self.device):
            native.check(native.lib().hex_ionic_create(
                self.device.index or 0, self.ndof, self.nrhs, len(cols), native.ptr(self._ptr), native.ptr(self._col),
                self.nk, native.ptr(data), ct.byref(handle)), 'ionic adjoint')
        self._handle = handle
        # field-free blocks: one eigenbasis per l, replicated per state
        values, vectors, inverses, conditions = {}, {}, {}, {}
        for l in sorted({l for l, m in ion.states}):
            i = next(i for i, (ll, m) in enumerate(ion.states) if ll == l)
            block = matrices[0][i * self.n:(i + 1) * self.n, i * self.n:(i + 1) * self.n].toarray()
            e, S = eig(block)
            values[l], vectors[l], inverses[l] = e, S, inv(S)
            conditions[l] = float(np.linalg.cond(S))
        if max(conditions.values()) > 1e12:
            raise ValueError('ill-conditioned ionic adjoint preconditioner')
        self.condition_numbers = conditions
        S = np.ascontiguousarray([vectors[l] for l, m in ion.states], complex)
        Si = np.ascontiguousarray([inverses[l] for l, m in ion.states], complex)
        E = np.ascontiguousarray([values[l] for l, m in ion.states], complex)
        native.check(native.lib().hex_ionic_set_eigen(self._handle, self.n, self.na, native.ptr(S), native.ptr(Si),
                                                      native.ptr(E)), 'ionic adjoint eigenbasis')
        self.separable = SimpleNamespace(apply=self.precondition, flexible=False)

    def __del__(self):
        handle = getattr(self, '_handle', None)
        if handle is not None and handle.value:
            native.lib().hex_ionic_destroy(handle)
            self._handle = None

    @property
    def device_bytes(self):
        return int(native.lib().hex_ionic_bytes(self._handle))

    def coefficients(self, field):
        """The operator's coefficients for a field (python/ionic_adjoint.py:36-37)."""
        return np.asarray(field, float) if self.mixed else np.r_[field, .5 * np.dot(field, field)]

    def tensor(self, x):
        return torch.as_tensor(np.array(x, copy=True), dtype=torch.complex128, device=self.device)

    def state(self, chi):
        """Device state (flat, row-major (ndof, nrhs)) of host or device duals."""
        if isinstance(chi, torch.Tensor):
            x = chi.to(self.device, torch.complex128)
        else:
            x = self.tensor(np.ascontiguousarray(chi))
        if x.numel() != self.size:
            raise ValueError(f'dual state has {x.numel()} entries, expected {self.size}')
        return x.reshape(-1).contiguous()

    def host(self, x):
        return x.detach().cpu().numpy().reshape(self.ndof, self.nrhs)

    def _check(self, x):
        if x.dtype != torch.complex128 or x.device != self.device or x.numel() != self.size or not x.is_contiguous():
            raise ValueError('dual state must be a contiguous complex128 device tensor of ndof * nrhs entries')

    @torch.no_grad()
    def apply_affine(self, x, field, velocity, a, b, out=None):
        """out = a G(field) x + b x for the coefficients ``field``."""
        x = x.reshape(-1)
        self._check(x)
        coef = np.ascontiguousarray(field, float)
        if coef.shape != (self.nk,):
            raise ValueError(f'the ionic operator takes {self.nk} field coefficients')
        if out is None:
            out = torch.empty_like(x)
        self._check(out)
        ca = (ct.c_double * 2)(complex(a).real, complex(a).imag)
        cb = (ct.c_double * 2)(complex(b).real, complex(b).imag)
        native.check(native.lib().hex_ionic_apply(self._handle, native.ptr(coef), ca, cb, _vp(x), _vp(out),
                                                  _stream(self.device)), 'ionic apply')
        return out

    def apply(self, x, field, velocity=True, out=None):
        return self.apply_affine(x, field, velocity, 1., 0., out)

    @torch.no_grad()
    def precondition(self, x, alpha, shift, out=None):
        """Field-free resolvent (python/ionic_adjoint.py:39-42)."""
        x = x.reshape(-1)
        self._check(x)
        if out is None:
            out = torch.empty_like(x)
        self._check(out)
        alpha = complex(alpha)
        native.check(native.lib().hex_ionic_precondition(self._handle, alpha.real, alpha.imag, float(shift), _vp(x),
                                                         _vp(out), _stream(self.device)), 'ionic precondition')
        return out

    def step(self, x, field, t, dt, tol=1e-12):
        """One CF4-Pade step of the duals (python/streaming_surface.py:141-143)."""
        return implicit.cf4_step(self, x, lambda s: self.coefficients(field(s)), t, dt, 0., tol=tol)


class DeviceDuals:
    """chi(t_k) for every frame time t_k, on the device
    (python/streaming_surface.py:190-243 with ionic_device cuda:0).

    ``prepare`` runs one backward sweep from the dual final states and keeps
    the endpoints every ``block`` frames; a frame inside a block is replayed
    from the block's upper endpoint when first asked for, and the block's
    frames stay on the device until another block is asked for.  With
    ``block = 1`` every frame is an endpoint and nothing is replayed.
    ``substeps`` CF4 steps of dt = sample_dt / substeps separate two frames
    (the propagation's ``surface_stride``).  ``progress(done, total)`` is
    called every 512 sweep frames.

    ``folder``: the endpoints are also written, as they are produced, to
    ``folder/ionic_endpoints`` (a ``storage.ShardedArray`` of shape
    (endpoints, ndof, nrhs)) with ``folder/preparation.json`` recording the
    recipe (``configuration`` signature, frame count, block, device) and the
    lowest endpoint ready (python/streaming_surface.py:171-200).  A folder
    that already holds a preparation of the same recipe is resumed from its
    lowest ready endpoint, a complete one without any sweep; another recipe
    is refused."""

    def __init__(self, ion, unique, times, field, substeps=1, block=128, device='cuda:0', tol=1e-12, progress=None,
                 folder=None, configuration='', max_file_bytes=DEFAULT_FILE_BYTES):
        self.times = np.asarray(times, float)
        if len(self.times) < 2:
            raise ValueError('dual propagation needs at least two frames')
        self.sample_dt = float(self.times[1] - self.times[0])
        if not np.allclose(np.diff(self.times), self.sample_dt, rtol=1e-10, atol=1e-12):
            raise ValueError('online propagation requires uniform surface sampling')
        self.block = int(block)
        if self.block < 1:
            raise ValueError('ionic block must be positive')
        self.field = field
        self.substeps = int(substeps)
        self.tol = tol
        self.engine = IonicAdjoint(ion, len(unique), device)
        self.final = ion.dual_final_states(unique)
        frames = len(self.times)
        self.ends = np.unique(np.r_[np.arange(0, frames, self.block), frames - 1])
        self.end_lookup = {int(v): i for i, v in enumerate(self.ends)}
        self.endpoints = torch.empty((len(self.ends), self.engine.size), dtype=torch.complex128,
                                     device=self.engine.device)
        self.maximum_residual = 0.
        self.linear_iterations = 0
        self.cached_block = None
        self.cached = None
        self.progress = progress
        last = len(self.times) - 1
        self.lowest = last
        self.store = None
        self.endpoints[self.end_lookup[last]].copy_(self.engine.state(self.final))
        if folder is not None:
            self.folder = Path(folder)
            self.folder.mkdir(parents=True, exist_ok=True)
            self.record = self.folder / 'preparation.json'
            self.recipe = {'configuration': configuration, 'kernel': 'laqphysics_amd hex_ionic cf4', 'frames': frames,
                           'block': self.block, 'ionic_device': str(self.engine.device)}
            endpoint_dir = self.folder / 'ionic_endpoints'
            if self.record.exists():
                old = json.loads(self.record.read_text())
                if old['recipe'] != self.recipe:
                    raise ValueError('online projection recipe changed; use a fresh output directory')
                self.lowest = int(old['lowest_ready_frame'])
                self.maximum_residual = float(old.get('maximum_ionic_residual', 0.))
                self.store = ShardedArray(endpoint_dir, mode='r+')
                for k, end in enumerate(self.ends):
                    if end >= self.lowest:
                        self.endpoints[k].copy_(self.engine.state(np.asarray(self.store[k])))
            else:
                self.store = ShardedArray(endpoint_dir, (len(self.ends), self.engine.ndof, self.engine.nrhs),
                                          mode='w+', max_file_bytes=max_file_bytes)
                self.store[len(self.ends) - 1] = self.final
                self.store.flush()
                self._write_record()
        self.prepare()

    def _write_record(self):
        atomic_json(self.record, {'recipe': self.recipe, 'lowest_ready_frame': self.lowest,
                                  'complete': self.lowest == 0, 'maximum_ionic_residual': self.maximum_residual})

    def __len__(self):
        return len(self.times)

    def _backstep(self, x, i):
        d = self.sample_dt / self.substeps
        for sub in range(self.substeps):
            x, info = self.engine.step(x, self.field, self.times[i] - sub * d, -d, self.tol)
            self.maximum_residual = max(self.maximum_residual, info['linear_residual'])
            self.linear_iterations += info['linear_iterations']
        return x

    def prepare(self):
        """Backward sweep from the lowest ready endpoint down to frame 0."""
        last = len(self.times) - 1
        x = self.endpoints[self.end_lookup[self.lowest]].clone()
        for i in range(self.lowest, 0, -1):
            x = self._backstep(x, i)
            if i - 1 in self.end_lookup:
                k = self.end_lookup[i - 1]
                self.endpoints[k].copy_(x)
                if self.store is not None:
                    self.store[k] = x.cpu().numpy().reshape(self.engine.ndof, self.engine.nrhs)
                    self.store.flush()
                    self.lowest = i - 1
                    self._write_record()
            done = last - i + 1
            if self.progress is not None and done % 512 == 0:
                self.progress(done, last)
        self.lowest = 0

    def __getitem__(self, index):
        index = int(index)
        if index < 0:
            index += len(self.times)
        if not 0 <= index < len(self.times):
            raise IndexError('frame index out of range')
        shape = (self.engine.ndof, self.engine.nrhs)
        if index in self.end_lookup:
            return self.endpoints[self.end_lookup[index]].view(shape)
        block = int(np.searchsorted(self.ends, index, side='right')) - 1
        start, end = int(self.ends[block]), int(self.ends[block + 1])
        if self.cached_block != block:
            self.cached = torch.empty((end - start, self.engine.size), dtype=torch.complex128,
                                      device=self.engine.device)
            x = self.endpoints[block + 1].clone()
            for i in range(end, start + 1, -1):
                x = self._backstep(x, i)
                self.cached[i - 1 - start].copy_(x)
            self.cached_block = block
        return self.cached[index - start].view(shape)

    def host(self, index):
        return self[index].cpu().numpy()
