"""ORACLE -- backward CF4-Pade propagation of the He+ duals on the CPU (test infrastructure only).

Restates python/implicit.py:111-123 (``SparseCF4.step`` with a negative step)
as it is driven by python/streaming_surface.py:190-200 (``_backstep``) and
python/surface_projection.py:43-57: for every frame interval, ``substeps``
steps of dt = -sample_dt / substeps from t = t_i - sub * |dt|, each

    chi <- prod_f (2 (I + i alpha_f H_f)^{-1} - I) chi,
    H_f in (A2 G(t + C1 dt) + A1 G(t + C2 dt), A1 G(t + C1 dt) + A2 G(t + C2 dt)),
    alpha_f = dt / (3 + i sqrt3), dt / (3 - i sqrt3) for each H_f,

G(t) = H_ion(f(t))^dagger, with dense LU factors (scipy.linalg.lu_factor) for
dense G, or SuperLU factors with the minimum-degree ordering on A^T + A for
sparse G, as the reference's step factors them (scipy.sparse.linalg.splu).

Pins (tests/test_oracle_ionic.py): field-free, one step is S R(-i dt E / 2)^2
S^{-1} with R the [2/2] Pade approximant of exp (closed form in the
eigenbasis); with a field the propagation converges at fourth order to the
time-ordered exponential of d chi / dt = -i G(t) chi integrated by a
high-order Runge-Kutta method; and it equals the package's sparse-LU host
path (whose spectra reproduce the reference's recorded runs).
"""
import numpy as np
from scipy.linalg import lu_factor, lu_solve
from scipy.sparse import eye as sparse_eye, issparse
from scipy.sparse.linalg import splu

SQRT3 = np.sqrt(3.)
C1 = .5 - SQRT3 / 6
C2 = .5 + SQRT3 / 6
A1 = (3 - 2 * SQRT3) / 12
A2 = (3 + 2 * SQRT3) / 12


def cf4_step(G1, G2, chi, dt):
    """One step from G at the two stages (python/implicit.py:115-123)."""
    sparse = issparse(G1)
    I = sparse_eye(G1.shape[0], format='csc') if sparse else np.eye(G1.shape[0])
    for H in (A2 * G1 + A1 * G2, A1 * G1 + A2 * G2):
        for root in (3 + 1j * SQRT3, 3 - 1j * SQRT3):
            M = I + 1j * (dt / root) * H
            if sparse:
                solved = splu(M.tocsc(), permc_spec='MMD_AT_PLUS_A').solve(chi)
            else:
                solved = lu_solve(lu_factor(M), chi)
            chi = 2 * solved - chi
    return chi


def backward(matrix_at, chi, times, substeps=1):
    """[chi(t_0), ..., chi(t_last)] from chi(t_last) = ``chi``; ``matrix_at(t)``
    returns G(t), dense or sparse."""
    times = np.asarray(times, float)
    out = [None] * len(times)
    out[-1] = np.array(chi, complex)
    d = (times[1] - times[0]) / substeps
    x = out[-1]
    for i in range(len(times) - 1, 0, -1):
        for sub in range(substeps):
            t = times[i] - sub * d
            x = cf4_step(matrix_at(t - C1 * d), matrix_at(t - C2 * d), x, -d)
        out[i - 1] = x
    return out
