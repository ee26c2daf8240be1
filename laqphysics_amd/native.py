"""ctypes binding of the HIP engine library (laqphysics_amd/build/libhexhip.so).

Mirrors the reference's loader (python/native.py:7-38): one shared library,
loaded once, F-ordered complex128 buffers, checked arguments.  The library is
the HIP engine; there is no CPU fallback.  A missing library or a failing HIP
call raises, naming the library and the HIP error.

Environment: ``HELIUM_HIP_LIBRARY`` overrides the library path.
"""
import ctypes as ct
import os
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent
LIBRARY = Path(os.environ.get('HELIUM_HIP_LIBRARY', PKG / 'build' / 'libhexhip.so'))

HEX_OK, HEX_EINVAL, HEX_EDEVICE, HEX_ENOTCONV = 0, -1, -2, -3

_lib = None


class HexTables(ct.Structure):
    _fields_ = [('n', ct.c_int), ('nc', ct.c_int), ('lam_count', ct.c_int),
                ('nt', ct.c_int), ('ptr', ct.c_void_p), ('col', ct.c_void_p), ('tv', ct.c_void_p),
                ('r', ct.c_void_p), ('cut', ct.c_void_p), ('d1', ct.c_void_p), ('d2', ct.c_void_p),
                ('nv', ct.c_int), ('vptr', ct.c_void_p), ('vc', ct.c_void_p), ('vl', ct.c_void_p),
                ('vcoef', ct.c_void_p),
                ('nd', ct.c_int), ('dptr', ct.c_void_p), ('dc', ct.c_void_p), ('de', ct.c_void_p),
                ('dcoef', ct.c_void_p), ('ldiff', ct.c_void_p)]


class HexYieldDesc(ct.Structure):
    _fields_ = [('M', ct.c_int), ('nu', ct.c_int), ('na', ct.c_int), ('nr', ct.c_int), ('N', ct.c_int),
                ('dt', ct.c_double), ('h0', ct.c_void_p), ('pz', ct.c_void_p), ('chiN', ct.c_void_p),
                ('nent', ct.c_int), ('nlab', ct.c_int), ('nout', ct.c_int), ('ns', ct.c_int), ('nc', ct.c_int),
                ('ej', ct.c_void_p), ('ec', ct.c_void_p), ('ea', ct.c_void_p), ('eu', ct.c_void_p),
                ('ew', ct.c_void_p), ('nE', ct.c_int), ('nD', ct.c_int), ('tests', ct.c_void_p),
                ('harm', ct.c_void_p), ('energy', ct.c_void_p), ('kz', ct.c_void_p), ('wy', ct.c_void_p),
                ('times', ct.c_void_p)]


def lib():
    """The loaded engine library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not LIBRARY.exists():
            raise RuntimeError(f'HIP engine library {LIBRARY} is missing; build it with '
                               '`python -m laqphysics_amd.build` (there is no CPU fallback)')
        L = ct.CDLL(str(LIBRARY))
        L.hex_last_error.restype = ct.c_char_p
        L.hex_engine_bytes.restype = ct.c_int64
        L.hex_engine_create.argtypes = [ct.POINTER(HexTables), ct.c_int, ct.POINTER(ct.c_void_p)]
        L.hex_engine_create_mixed.argtypes = [ct.POINTER(HexTables), ct.c_void_p, ct.c_int, ct.POINTER(ct.c_void_p)]
        L.hex_engine_set_mixed.argtypes = [ct.c_void_p] * 5 + [ct.c_int] + [ct.c_void_p] * 4
        L.hex_apply_mixed.argtypes = [ct.c_void_p] * 7
        L.hex_engine_destroy.argtypes = [ct.c_void_p]
        L.hex_workspace_create.argtypes = [ct.c_int, ct.POINTER(ct.c_void_p)]
        L.hex_engine_bytes.argtypes = [ct.c_void_p]
        L.hex_apply.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_int, ct.c_double, ct.c_void_p, ct.c_void_p,
                                ct.c_void_p]
        L.hex_apply_adjoint.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_int, ct.c_void_p, ct.c_void_p, ct.c_void_p]
        L.hex_apply_terms.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_int, ct.c_int, ct.c_int, ct.c_void_p,
                                      ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_void_p]
        L.hex_engine_set_launch.argtypes = [ct.c_void_p, ct.c_int]
        L.hex_apply_batched.argtypes = [ct.c_void_p, ct.c_int, ct.c_void_p, ct.c_int, ct.c_void_p, ct.c_void_p,
                                        ct.c_void_p]
        L.hex_apply_affine.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_int, ct.c_void_p, ct.c_void_p, ct.c_void_p,
                                       ct.c_void_p, ct.c_void_p]
        L.hex_axpby.argtypes = [ct.c_int64, ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_void_p,
                                ct.c_void_p]
        L.hex_separable_create.argtypes = [ct.c_int, ct.c_int, ct.c_int, ct.c_int, ct.c_void_p, ct.c_void_p,
                                           ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.POINTER(ct.c_void_p)]
        L.hex_separable_create_c64.argtypes = L.hex_separable_create.argtypes
        L.hex_separable_single.argtypes = [ct.c_void_p]
        L.hex_separable_apply.argtypes = [ct.c_void_p, ct.c_double, ct.c_double, ct.c_double, ct.c_void_p,
                                          ct.c_void_p, ct.c_void_p]
        L.hex_separable_bytes.restype = ct.c_int64
        L.hex_separable_bytes.argtypes = [ct.c_void_p]
        L.hex_separable_destroy.argtypes = [ct.c_void_p]
        L.hex_separable_lanes.argtypes = [ct.c_void_p]
        L.hex_separable_set_lanes.argtypes = [ct.c_void_p, ct.c_int]
        L.hex_surface_contract.argtypes = [ct.c_int] * 5 + [ct.c_void_p] * 9
        L.hex_ionic_create.argtypes = [ct.c_int, ct.c_int64, ct.c_int, ct.c_int, ct.c_void_p, ct.c_void_p, ct.c_int,
                                       ct.c_void_p, ct.POINTER(ct.c_void_p)]
        L.hex_ionic_set_eigen.argtypes = [ct.c_void_p, ct.c_int, ct.c_int, ct.c_void_p, ct.c_void_p, ct.c_void_p]
        L.hex_ionic_apply.argtypes = [ct.c_void_p] * 7
        L.hex_ionic_precondition.argtypes = [ct.c_void_p, ct.c_double, ct.c_double, ct.c_double, ct.c_void_p,
                                             ct.c_void_p, ct.c_void_p]
        L.hex_ionic_bytes.restype = ct.c_int64
        L.hex_ionic_bytes.argtypes = [ct.c_void_p]
        L.hex_ionic_destroy.argtypes = [ct.c_void_p]
        L.hex_volkov_add.argtypes = ([ct.c_int] * 5 + [ct.c_void_p] * 6 + [ct.c_double, ct.c_void_p, ct.c_double,
                                     ct.c_double, ct.c_int, ct.c_void_p, ct.c_void_p, ct.c_void_p])
        L.hex_expm_host.argtypes = [ct.c_int, ct.c_void_p, ct.c_void_p]
        L.hex_expm_batched.argtypes = [ct.c_int, ct.c_int, ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_void_p]
        L.hex_batch_create.argtypes = [ct.c_void_p, ct.c_int, ct.c_int, ct.c_double, ct.c_double, ct.c_int,
                                       ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.POINTER(ct.c_void_p)]
        L.hex_batch_set_launch.argtypes = [ct.c_void_p, ct.c_int]
        L.hex_batch_run.argtypes = [ct.c_void_p, ct.c_int, ct.c_void_p]
        L.hex_batch_status.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_void_p]
        L.hex_batch_bytes.restype = ct.c_int64
        L.hex_batch_bytes.argtypes = [ct.c_void_p]
        L.hex_batch_destroy.argtypes = [ct.c_void_p]
        L.hex_engine_set_surface.argtypes = [ct.c_void_p, ct.c_int, ct.c_int, ct.c_int, ct.c_void_p, ct.c_void_p]
        L.hex_flux.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_void_p]
        L.hex_flux_terms.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_int, ct.c_void_p, ct.c_void_p, ct.c_void_p]
        L.hex_flux_adjoint.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_void_p]
        L.hex_cgs2.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_int64, ct.c_int, ct.c_void_p, ct.c_int64,
                               ct.c_void_p, ct.c_void_p]
        L.hex_cgs2_commit.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_int64, ct.c_int, ct.c_void_p, ct.c_int64,
                                      ct.c_void_p, ct.c_void_p, ct.c_void_p]
        L.hex_arnoldi_check.argtypes = [ct.c_void_p, ct.c_int, ct.c_int, ct.c_int, ct.c_double, ct.c_double,
                                        ct.c_double, ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_void_p]
        L.hex_arnoldi_check_fixed.argtypes = [ct.c_void_p, ct.c_int, ct.c_int, ct.c_int, ct.c_int, ct.c_double,
                                              ct.c_double, ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_void_p]
        L.hex_gmres_check.argtypes = [ct.c_void_p, ct.c_int, ct.c_int, ct.c_int, ct.c_int, ct.c_double, ct.c_double,
                                      ct.c_void_p, ct.c_void_p, ct.c_void_p]
        L.hex_cgs2_commit_noalloc.argtypes = [ct.c_void_p, ct.c_int64, ct.c_int, ct.c_void_p, ct.c_int64,
                                              ct.c_void_p, ct.c_void_p, ct.c_void_p]
        L.hex_norm.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_int64, ct.POINTER(ct.c_double), ct.c_void_p]
        L.hex_scale.argtypes = [ct.c_void_p, ct.c_double, ct.c_void_p, ct.c_int64, ct.c_void_p]
        L.hex_combine.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_int64, ct.c_int, ct.c_void_p, ct.c_void_p,
                                  ct.c_int64, ct.c_void_p]
        L.hex_dots.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_int64, ct.c_int, ct.c_void_p, ct.c_int64,
                               ct.c_void_p, ct.c_void_p]
        L.hex_yield_create.argtypes = [ct.c_int, ct.POINTER(HexYieldDesc), ct.POINTER(ct.c_void_p)]
        L.hex_yield_destroy.argtypes = [ct.c_void_p]
        L.hex_yield_bytes.argtypes = [ct.c_void_p]
        L.hex_yield_bytes.restype = ct.c_int64
        L.hex_yield_value.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_void_p,
                                      ct.c_void_p]
        L.hex_yield_gradient.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_void_p,
                                         ct.c_void_p, ct.c_void_p, ct.c_void_p]
        L.hex_dots_device.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_int64, ct.c_int, ct.c_void_p, ct.c_int64,
                                      ct.c_void_p, ct.c_void_p]
        _lib = L
    return _lib


def check(code, what=''):
    """Map an engine status code onto the reference's exception types."""
    if code == HEX_OK:
        return
    message = lib().hex_last_error().decode(errors='replace')
    prefix = f'{what}: ' if what else ''
    if code == HEX_EINVAL:
        raise ValueError(prefix + message)
    raise RuntimeError(prefix + message)


def ptr(a):
    return a.ctypes.data_as(ct.c_void_p)


def tensor_apply(h, psi, field=(0., 0., 0.), velocity=False):
    """H.psi with host buffers through the HIP drop-in (python/native.py:24-38)."""
    z = np.asfortranarray(np.asarray(psi).reshape(h.shape, order='F'), dtype=np.complex128)
    out = np.empty_like(z, order='F')
    f = np.ascontiguousarray(field, dtype=np.float64)
    if f.shape != (3,):
        raise ValueError('field must have three Cartesian components')
    keep = {name: np.ascontiguousarray(getattr(h, name)) for name in
            ('ptr', 'col', 'tv', 'vptr', 'vc', 'vl', 'vcoef', 'dptr', 'dc', 'de', 'ldiff', 'r')}
    keep['tv'] = keep['tv'].astype(complex)
    keep['r'] = keep['r'].astype(complex)
    diag = np.asfortranarray(h.diag, dtype=np.complex128)
    rad = np.asfortranarray(h.rad, dtype=np.complex128)
    dcoef = np.asfortranarray(h.dcoef, dtype=np.complex128)
    args = [ct.c_int(h.n), ct.c_int(h.nc), ct.c_int(len(keep['tv']))]
    args += [ptr(keep['ptr']), ptr(keep['col']), ptr(keep['tv']), ptr(diag)]
    args += [ct.c_int(len(keep['vc']))]
    args += [ptr(keep[a]) for a in ('vptr', 'vc', 'vl', 'vcoef')] + [ptr(rad)]
    args += [ct.c_int(len(keep['dc']))]
    args += [ptr(keep[a]) for a in ('dptr', 'dc', 'de')] + [ptr(dcoef)]
    args += [ptr(keep['ldiff']), ptr(keep['r'])]
    args += [ptr(f), ct.c_int(int(velocity)), ptr(z), ptr(out)]
    check(lib().tensor_apply(*args), 'tensor_apply')
    return out.ravel(order='F')


def surface_flux(h, psi, field):
    """tSURFF source with host buffers (python/helium3d.py:69-77) on the device."""
    from .hip_backend import HipHamiltonian
    engine = getattr(h, '_flux_engine', None)
    if engine is None:
        engine = HipHamiltonian(h)
        h._flux_engine = engine
    return engine.flux(engine.state(psi), field)
