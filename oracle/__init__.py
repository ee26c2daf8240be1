"""ORACLE -- CPU restatement of the reference hot path, used only as the checker.

Test infrastructure: only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this package.  The shipped
library (``laqphysics_amd``) never imports it.

Contents
  tensor_apply      C/OpenMP restatement of src/kernels.f90:119-176 (tensor_apply.c)
  ref_tensor_apply  the reference's own Fortran kernel, compiled by oracle/Makefile
                    into oracle/_ref/ when its source and amdflang are available
  kron              independent Kronecker assembly (python/helium3d.py:105-123)
  arnoldi           NumPy/BLAS Arnoldi exponential (python/propagate.py:6-50)
  implicit          CF4 / [2/2] Pade / GMRES (python/implicit.py:13-109), NumPy
  adjoint_yield     the adjoint objective's He+ yield (duals, projection, Volkov)
                    in PyTorch on the CPU, gradient by autograd
  surface           tSURFF flux and Volkov accumulation (python/helium3d.py:69-77,
                    python/streaming_surface.py:39-64, surface_projection.py:126-144)
"""
import ctypes as ct
import os
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
_LIBS = {}


def build():
    """Compile the C restatement (and oracle/_ref when possible) with make."""
    import subprocess
    subprocess.run(['make', '-s', '-C', str(HERE)], check=True)


def _load(kind):
    if kind not in _LIBS:
        path = HERE / ('_build/liboracle.so' if kind == 'oracle' else '_ref/libhelium_ref.so')
        if not path.exists():
            if kind == 'oracle':
                build()
            if not path.exists():
                raise FileNotFoundError(f'{path} is not built (make -C oracle)')
        _LIBS[kind] = ct.CDLL(str(path))
    return _LIBS[kind]


def have_reference_kernel():
    return (HERE / '_ref/libhelium_ref.so').exists()


def _p(a):
    return a.ctypes.data_as(ct.c_void_p)


def _call(lib, symbol, h, psi, field, velocity):
    z = np.asfortranarray(np.asarray(psi).reshape(h.shape, order='F'), dtype=np.complex128)
    out = np.empty_like(z, order='F')
    f = np.ascontiguousarray(field, dtype=np.float64)
    if f.shape != (3,):
        raise ValueError('field must have three Cartesian components')
    tables = {name: np.ascontiguousarray(getattr(h, name)) for name in
              ('ptr', 'col', 'tv', 'vptr', 'vc', 'vl', 'vcoef', 'dptr', 'dc', 'de', 'ldiff', 'r')}
    diag = np.asfortranarray(h.diag, dtype=np.complex128)
    rad = np.asfortranarray(h.rad, dtype=np.complex128)
    dcoef = np.asfortranarray(h.dcoef, dtype=np.complex128)
    args = [ct.c_int(h.n), ct.c_int(h.nc), ct.c_int(len(tables['tv']))]
    args += [_p(tables['ptr']), _p(tables['col']), _p(tables['tv']), _p(diag)]
    args += [ct.c_int(len(tables['vc']))]
    args += [_p(tables[a]) for a in ('vptr', 'vc', 'vl', 'vcoef')] + [_p(rad)]
    args += [ct.c_int(len(tables['dc']))]
    args += [_p(tables[a]) for a in ('dptr', 'dc', 'de')] + [_p(dcoef)]
    args += [_p(tables['ldiff']), _p(tables['r'])]
    args += [_p(f), ct.c_int(int(velocity)), _p(z), _p(out)]
    getattr(lib, symbol)(*args)
    return out.ravel(order='F')


def tensor_apply(h, psi, field=(0., 0., 0.), velocity=False):
    """H.psi by the C restatement; same call as python/native.py:24-38."""
    return _call(_load('oracle'), 'oracle_tensor_apply', h, psi, field, velocity)


def ref_tensor_apply(h, psi, field=(0., 0., 0.), velocity=False):
    """H.psi by the reference's own Fortran kernel built into oracle/_ref."""
    return _call(_load('ref'), 'tensor_apply', h, psi, field, velocity)


def set_threads(count):
    """OpenMP thread count for the C restatement (the reference uses OMP_NUM_THREADS)."""
    os.environ['OMP_NUM_THREADS'] = str(int(count))
