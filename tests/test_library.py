"""The C-ABI library loads and exports every entry point its header declares.
No compute calls (CPU only)."""
import ctypes as ct
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def declared_functions():
    names = []
    for header in sorted((ROOT / 'include').glob('*.h')):
        text = re.sub(r'/\*.*?\*/', '', header.read_text(), flags=re.S)
        for m in re.finditer(r'^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b([A-Za-z_]\w*)\s*\(', text, flags=re.M):
            name = m.group(1)
            if name not in ('if', 'for', 'while', 'return', 'sizeof'):
                names.append(name)
    return sorted(set(names))


def test_header_declares_entry_points():
    names = declared_functions()
    for required in ('tensor_apply', 'hex_engine_create', 'hex_apply', 'hex_cgs2', 'hex_last_error'):
        assert required in names


def test_library_exports_every_declared_symbol():
    from laqphysics_amd import native
    if not native.LIBRARY.exists():
        from laqphysics_amd import build
        build.build()
    lib = ct.CDLL(str(native.LIBRARY))
    missing = [name for name in declared_functions() if not hasattr(lib, name)]
    assert not missing, missing


def test_missing_library_fails_loudly(tmp_path, monkeypatch):
    from laqphysics_amd import native
    monkeypatch.setattr(native, 'LIBRARY', tmp_path / 'absent.so')
    monkeypatch.setattr(native, '_lib', None)
    with pytest.raises(RuntimeError, match='no CPU fallback'):
        native.lib()


def test_host_expm_matches_scipy_and_closed_forms():
    """The small-matrix exponential of the device Krylov control, run on the host."""
    import numpy as np
    from scipy.linalg import expm
    from laqphysics_amd import native
    L = native.lib()

    def run(A):
        A = np.ascontiguousarray(A, dtype=complex)
        E = np.empty_like(A)
        native.check(L.hex_expm_host(A.shape[0], A.ctypes.data_as(ct.c_void_p), E.ctypes.data_as(ct.c_void_p)))
        return E
    rng = np.random.default_rng(0)
    for m in (1, 3, 8, 12, 30):
        for scale in (1e-3, .3, 2., 20.):
            A = (rng.normal(size=(m, m)) + 1j * rng.normal(size=(m, m))) * scale / np.sqrt(m)
            R = expm(A)
            assert np.linalg.norm(run(A) - R) <= 1e-13 * np.linalg.norm(R)
    # closed forms: diagonal, nilpotent Jordan block, 2x2 rotation generator
    d = np.array([.3, -1.2, 2.5]) - .4j
    np.testing.assert_allclose(run(np.diag(d)), np.diag(np.exp(d)), rtol=1e-14)
    J = np.diag(np.ones(3), 1)
    np.testing.assert_allclose(run(J), np.eye(4) + J + J @ J / 2 + J @ J @ J / 6, atol=1e-15)
    w = 1.7
    np.testing.assert_allclose(run(np.array([[0, -w], [w, 0]])),
                               np.array([[np.cos(w), -np.sin(w)], [np.sin(w), np.cos(w)]]), atol=1e-15)


def test_bad_arguments_return_einval_with_a_message():
    """Argument checks happen before any device work, so they run without a
    GPU: each entry point rejects null or inconsistent arguments with
    HEX_EINVAL and a message, and native.check maps that to ValueError."""
    from laqphysics_amd import native
    L = native.lib()
    null = ct.c_void_p()
    f = (ct.c_double * 3)(0., 0., 1.)
    two = (ct.c_double * 2)(1., 0.)
    cases = [
        (L.hex_apply, (null, f, 1, ct.c_double(1.), null, null, null)),
        (L.hex_apply_adjoint, (null, f, 1, null, null, null)),
        (L.hex_apply_affine, (null, f, 1, two, two, null, null, null)),
        (L.hex_engine_set_launch, (null, 3)),
        (L.hex_cgs2, (null, null, ct.c_int64(1), 1, null, ct.c_int64(1), null, null)),
        (L.hex_separable_apply, (null, ct.c_double(1.), ct.c_double(0.), ct.c_double(0.), null, null, null)),
        (L.hex_batch_run, (null, 1, null)),
        (L.hex_axpby, (ct.c_int64(0), two, null, two, null, null, null)),
    ]
    for fn, args in cases:
        code = fn(*args)
        assert code == native.HEX_EINVAL, fn.__name__
        assert L.hex_last_error(), fn.__name__
        with pytest.raises(ValueError):
            native.check(code, fn.__name__)
