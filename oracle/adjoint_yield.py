"""ORACLE -- the adjoint objective's He+ yield term in PyTorch on the CPU, with
its gradient by autograd.  Test infrastructure only (see the package header):
the device yield (laqphysics_amd/csrc/hex_yield.hip) is checked against it.

The yield has no reference counterpart as a function of the control (the
reference has no gradients); its value restates the reference's online tSURFF
chain for recorded frames: the ionic duals by the backward CF4-Pade steps of
python/implicit.py:111-123 (dense solves here), the channel projection of
python/streaming_surface.py:198-214 and the Volkov accumulation of
python/streaming_surface.py:39-64.  It is pinned in tests/test_adjoint_yield.py
against the sparse-LU dual propagation, the oracle's own tSURFF pipeline
(oracle/surface.py) and finite differences.
"""
import numpy as np
import torch

ROOTS = (3 + 1j * np.sqrt(3.), 3 - 1j * np.sqrt(3.))


class TorchYield:
    """Y(theta, F0, Fd) from the host tables of a ``laqphysics_amd.adjoint.SurfaceYield``."""

    def __init__(self, sy):
        c = lambda x: torch.as_tensor(np.ascontiguousarray(x), dtype=torch.complex128)
        self.sy = sy
        self.dt = float(sy.problem.dt)
        self.h0, self.pz, self.chiN = c(sy.h0), c(sy.pz), c(sy.chiN)
        self.eye = torch.eye(self.h0.shape[0], dtype=torch.complex128)
        self.Gs = torch.as_tensor(sy.Gs)
        self.Gc = torch.as_tensor(sy.Gc)
        self.times = torch.as_tensor(sy.times)
        self.energy = torch.as_tensor(sy.energy)
        self.kz = torch.as_tensor(sy.kz)
        self.wy = torch.as_tensor(sy.wy)
        self.tests, self.harm, self.ew = c(sy.tests), c(sy.harm), c(sy.ew)
        self.ej, self.ec, self.ea, self.eu = (torch.as_tensor(x, dtype=torch.long) for x in (sy.ej, sy.ec, sy.ea, sy.eu))

    def duals(self, theta, chunk=64):
        """chi(t_n), n = 0..N: each step's propagator prod_f (2 M_f^-1 - 1) by
        dense solves with identity right-hand sides, ``chunk`` steps at a time."""
        from laqphysics_amd.ionic import A1, A2
        dt = -self.dt
        eye = self.eye
        a = (self.Gc @ theta).to(torch.complex128)   # (N+1, 2)
        chi = self.chiN
        out = [None] * len(self.times)
        out[-1] = chi
        steps = list(range(len(self.times) - 1, 0, -1))
        for c0 in range(0, len(steps), chunk):
            idx = torch.as_tensor(steps[c0:c0 + chunk])
            a1, a2 = a[idx, 0][:, None, None], a[idx, 1][:, None, None]
            h1 = self.h0 + eye * (a1 * a1 / 2) + a1 * self.pz
            h2 = self.h0 + eye * (a2 * a2 / 2) + a2 * self.pz
            prop = None
            for H in (A2 * h1 + A1 * h2, A1 * h1 + A2 * h2):
                for root in ROOTS:
                    factor = 2 * torch.linalg.solve(eye + (1j * dt / root) * H, eye.expand_as(H)) - eye
                    prop = factor if prop is None else factor @ prop
            for q, i in enumerate(steps[c0:c0 + chunk]):
                chi = prop[q] @ chi
                out[i - 1] = chi
        return out

    def value(self, theta, F0, Fd):
        """Y from theta (K,) and the frames F0, Fd, (N+1, nc, ns, n) CPU tensors."""
        sy = self.sy
        A = self.Gs @ theta
        F = F0 + A[:, None, None, None].to(F0.dtype) * Fd
        chis = torch.stack(self.duals(theta)).reshape(-1, sy.na, sy.nr, sy.nu)   # (S, na, n, nu)
        nlab, nout, ns = sy.shape
        dt = self.dt
        G = chis.permute(0, 1, 3, 2)[:, self.ea, self.eu].conj()                  # (S, E, n)
        contrib = self.ew[None, :, None] * torch.einsum('ier,iesr->ies', G, F[:, self.ec])
        S = F.shape[0]
        q = torch.zeros(S, nlab * nout, ns, dtype=torch.complex128).index_add(1, self.ej, contrib)
        q = q.reshape(S, nlab, nout, ns)
        zero = torch.zeros(1, dtype=A.dtype)
        intA = torch.cat([zero, torch.cumsum(dt * (A[1:] + A[:-1]) / 2, 0)])
        intA2 = torch.cat([zero, torch.cumsum(dt * (A[1:] ** 2 + A[:-1] ** 2) / 2, 0)])
        phase = (self.energy[None, :, None] * self.times[:, None, None] + self.kz[None] * intA[:, None, None]
                 + intA2[:, None, None] / 2)                                      # (S, nE, nD)
        radial = torch.einsum('ijls,les->ijel', q, self.tests)
        cur = 1j * np.sqrt(2.) * torch.einsum('ijel,ld->ijed', radial, self.harm) * torch.exp(1j * phase)[:, None]
        amp = dt * (cur[1:] + cur[:-1]).sum(0) / 2
        return torch.sum(self.wy[None] * (amp.real ** 2 + amp.imag ** 2))

    def gradient(self, theta, F0, Fd):
        """(Y, dY/dtheta, dY/dconj(F0), dY/dconj(Fd)) by autograd (torch's
        complex gradient is twice the conjugate Wirtinger derivative)."""
        theta = torch.as_tensor(theta, dtype=torch.float64).clone().requires_grad_(True)
        F0 = F0.clone().requires_grad_(True)
        Fd = Fd.clone().requires_grad_(True)
        Y = self.value(theta, F0, Fd)
        Y.backward()
        return float(Y.detach()), theta.grad.numpy(), F0.grad / 2, Fd.grad / 2
