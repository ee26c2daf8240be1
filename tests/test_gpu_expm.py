"""The device small-matrix exponential of the Krylov control (gpu).

`hex_expm_batched` (one wave per matrix, work in global memory) and the
Arnoldi check kernel (work in LDS up to order 22, in global memory above)
run the wave-level scaling-and-squaring Pade algorithm: the LU of the Pade
denominator is spread over the wave (pivot by a wave argmax, entry updates by
lane).  Checked against scipy's expm and the host build of the serial
algorithm, over orders 1..64 and norms that select every Pade degree (3, 5,
7, 9, 13) and up to eight squarings, including matrices whose pivot columns
hold ties (entries of equal magnitude).
"""
import ctypes as ct

import numpy as np
import pytest
import torch
from scipy.linalg import expm

from laqphysics_amd import native
from laqphysics_amd.hip_backend import _stream, _vp

pytestmark = pytest.mark.gpu

REL = 1e-13


def _say(*what):
    print('CASE', *what, flush=True)


def _host(A):
    A = np.ascontiguousarray(A, dtype=complex)
    E = np.empty_like(A)
    native.check(native.lib().hex_expm_host(A.shape[0], A.ctypes.data_as(ct.c_void_p), E.ctypes.data_as(ct.c_void_p)))
    return E


def _device_batched(mats, device):
    count, m, _ = mats.shape
    A = torch.tensor(np.ascontiguousarray(mats), dtype=torch.complex128, device=device)
    E = torch.empty_like(A)
    work = torch.empty(count * 6 * m * m, dtype=torch.complex128, device=device)
    native.check(native.lib().hex_expm_batched(count, m, _vp(A), _vp(E), _vp(work), _stream(device)), 'expm')
    torch.cuda.synchronize(device)
    return E.cpu().numpy()


def _cases(m, rng):
    out = []
    # 1-norms from 1e-3 to ~200: degrees 3, 5, 7, 9 and 13 with 0..8 squarings
    for scale in (1e-3, .1, .5, 1.5, 4., 30., 200.):
        A = (rng.normal(size=(m, m)) + 1j * rng.normal(size=(m, m))) * scale / m
        out.append(A)
    # ties in every pivot column: entries of equal magnitude
    out.append(rng.choice([-1., 1.], size=(m, m)) * (.7 + .7j) / m)
    out.append(np.ones((m, m)) * .05j)
    return out


@pytest.mark.parametrize('m', [1, 2, 4, 8, 12, 13, 22, 30, 64])
def test_expm_batched_against_scipy_and_host(device, m):
    _say('expm batched order', m)
    rng = np.random.default_rng(100 + m)
    mats = np.stack(_cases(m, rng))
    got = _device_batched(mats, device)
    for A, E in zip(mats, got):
        R = expm(A)
        # the squarings carry the rounding of the scaled approximant up with the norm
        nrm = np.abs(A).sum(axis=0).max()
        assert np.linalg.norm(E - R) <= REL * max(1., nrm) * np.linalg.norm(R), (m, nrm)
        H = _host(A)
        assert np.linalg.norm(E - H) <= REL * max(1., nrm) * np.linalg.norm(H), m


@pytest.mark.parametrize('maxdim', [12, 22, 30])
def test_arnoldi_check_coefficients(device, maxdim):
    """The check kernel's accepted coefficients, beta expm(-i dt H_m)[:, 0], at a
    forced dimension m, for the LDS path (m <= 22) and the global-work path."""
    _say('arnoldi check coefficients maxdim', maxdim)
    L = native.lib()
    rng = np.random.default_rng(maxdim)
    hc = np.zeros((64, 64), complex)
    for j in range(maxdim):
        hc[j, :j + 1] = rng.normal(size=j + 1) + 1j * rng.normal(size=j + 1)
        hc[j, j + 1] = abs(rng.normal()) + .5
    hcols = torch.tensor(hc, device=device)
    H = torch.zeros((maxdim + 1) * maxdim, dtype=torch.complex128, device=device)
    res = torch.zeros(64, dtype=torch.complex128, device=device)
    work = torch.zeros(8 * maxdim * maxdim, dtype=torch.complex128, device=device)
    beta = 1.7
    for dt in (.02, .3, 2.):
        for m in sorted({1, 4, maxdim // 2, maxdim}):
            native.check(L.hex_arnoldi_check_fixed(_vp(hcols), 0, m - 1, maxdim, m, ct.c_double(dt),
                                                   ct.c_double(beta), _vp(H), _vp(work), _vp(res),
                                                   _stream(device)), 'check')
            r = res.cpu().numpy()
            assert r[0].real == 1.0 and int(r[0].imag) == m, (dt, m, r[0])
            Hm = np.zeros((m, m), complex)
            for c in range(m):
                Hm[:c + 1, c] = hc[c, :c + 1]
                if c + 1 < m:
                    Hm[c + 1, c] = hc[c, c + 1].real
            A = -1j * dt * Hm
            ref = expm(A)[:, 0] * beta
            nrm = np.abs(A).sum(axis=0).max()
            assert np.linalg.norm(r[2:2 + m] - ref) <= REL * max(1., nrm) * np.linalg.norm(ref), (dt, m)
