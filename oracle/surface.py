"""ORACLE -- tSURFF surface source and Volkov accumulation (test infrastructure only).

``flux`` restates python/helium3d.py:69-77: out[i, s, c] = sum_k J[s, k] psi(i, k, c)
plus, for electron-2 dipole edges, coef * sum_k JP[s, k] psi(i, k, dc) with the
surface rows J = [T, Theta] and JP = i [T, Theta] o (r_k - r_s) prepared by
``prepare_surface`` (python/helium3d.py:52-67).
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
