"""The no-allocation CGS2 entry point against the engine-scratch one (gpu).

hex_cgs2_commit_noalloc runs the kernels of hex_cgs2_commit on the library's
fixed device scratch instead of an engine's: on two calls in a row (the
second smaller, so it runs on what the first left in the scratch) the
orthogonalised vector, the Hessenberg column and the next basis vector agree
bit for bit.  A basis beyond the scratch's 63 vectors is refused before any
device work."""
import ctypes as ct

import pytest
import torch

from laqphysics_amd import native

pytestmark = pytest.mark.gpu


def _p(t):
    return ct.c_void_p(t.data_ptr())


def _lib():
    L = native.lib()
    L.hex_cgs2_commit_noalloc.argtypes = [ct.c_void_p, ct.c_int64, ct.c_int, ct.c_void_p, ct.c_int64, ct.c_void_p,
                                          ct.c_void_p, ct.c_void_p]
    L.hex_cgs2_commit.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_int64, ct.c_int, ct.c_void_p, ct.c_int64,
                                  ct.c_void_p, ct.c_void_p, ct.c_void_p]
    return L


def _stream():
    return ct.c_void_p(torch.cuda.current_stream().cuda_stream)


def _case(L, ws, N, k, seed):
    g = torch.Generator(device='cpu').manual_seed(seed)
    A = torch.randn(N, k + 1, dtype=torch.complex128, generator=g)
    basis, _ = torch.linalg.qr(A[:, :k])          # orthonormal columns
    Q = torch.cat([basis.T, A[:, k:].T]).contiguous().to('cuda:0')
    v = Q[k].clone()
    outs = []
    for which in ('engine', 'noalloc'):
        vv = v.clone()
        h = torch.zeros(k + 1, dtype=torch.complex128, device='cuda:0')
        q = torch.empty(N, dtype=torch.complex128, device='cuda:0')
        if which == 'engine':
            rc = L.hex_cgs2_commit(ws, _p(Q), N, k, _p(vv), N, _p(h), _p(q), _stream())
        else:
            rc = L.hex_cgs2_commit_noalloc(_p(Q), N, k, _p(vv), N, _p(h), _p(q), _stream())
        assert rc == native.HEX_OK, native.last_error() if hasattr(native, 'last_error') else rc
        outs.append((vv, h, q))
    torch.cuda.synchronize()
    (v1, h1, q1), (v2, h2, q2) = outs
    assert torch.equal(v1, v2) and torch.equal(h1, h2) and torch.equal(q1, q2)
    # and it is an orthogonalisation: q is a unit vector orthogonal to Q[:k]
    assert abs(float(torch.linalg.norm(q2)) - 1.) < 1e-13
    if k:
        assert float(torch.max(torch.abs(Q[:k].conj() @ q2))) < 1e-12


def test_noalloc_cgs2_matches_engine_scratch_twice(device):
    print('CASE cgs2_commit_noalloc vs cgs2_commit: N 200000 k 9, then N 3001 k 3', flush=True)
    L = _lib()
    ws = ct.c_void_p()
    assert L.hex_workspace_create(0, ct.byref(ws)) == native.HEX_OK
    try:
        _case(L, ws, 200000, 9, 1)
        _case(L, ws, 3001, 3, 2)
    finally:
        L.hex_engine_destroy(ws)


def test_noalloc_cgs2_refuses_a_basis_beyond_its_scratch(device):
    print('CASE cgs2_commit_noalloc refuses k = 64', flush=True)
    L = _lib()
    v = torch.zeros(16, dtype=torch.complex128, device='cuda:0')
    h = torch.zeros(65, dtype=torch.complex128, device='cuda:0')
    Q = torch.zeros(64, 16, dtype=torch.complex128, device='cuda:0')
    rc = L.hex_cgs2_commit_noalloc(_p(Q), 16, 64, _p(v), 16, _p(h), _p(v.clone()), _stream())
    assert rc == native.HEX_EINVAL
