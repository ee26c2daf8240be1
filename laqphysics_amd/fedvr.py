"""Normalised finite-element DVR radial grid with an optional irECS Laguerre tail.

Mirror of the reference grid builder (python/fedvr.py:12-77).  The produced
``Grid`` carries the same fields under the same names, so operator tables
built on it are the ones the reference builds:

* Gauss-Lobatto elements on the real segment, kinetic blocks ``D^T W D / h``
  accumulated element by element *before* the DVR normalisation
  (python/fedvr.py:49-52);
* an optional exterior-complex-scaled Laguerre-Radau tail that starts at the
  last real edge, ``r = R + e^{i theta}/alpha * y`` (python/fedvr.py:54-69);
* Dirichlet at the origin, and at the outer edge only without a tail
  (python/fedvr.py:72);
* normalisation by the complex ``sqrt(w_i w_j)`` (no conjugation, no abs) and
  removal of entries below 1e-13 (python/fedvr.py:74-75).
"""
from dataclasses import dataclass
from functools import lru_cache

import numpy as np
from scipy.sparse import csr_matrix
from scipy.special import eval_laguerre, eval_legendre, roots_genlaguerre, roots_jacobi


def derivative(nodes):
    """Lagrange differentiation matrix on ``nodes`` (barycentric form).

    ``D[i, k] = l_k'(x_i)``; the diagonal is minus the off-diagonal row sum so
    constants are differentiated exactly (python/fedvr.py:12-18).
    """
    x = np.asarray(nodes, dtype=float)
    gap = x[:, None] - x[None, :]
    np.fill_diagonal(gap, 1.0)
    bary = 1.0 / np.prod(gap, axis=1)
    out = (bary[None, :] / bary[:, None]) / gap
    np.fill_diagonal(out, 0.0)
    np.fill_diagonal(out, -out.sum(axis=1))
    return out


@dataclass
class Grid:
    """Radial grid: points, complex DVR weights, normalised kinetic CSR."""
    r: np.ndarray
    weights: np.ndarray
    kinetic: csr_matrix
    real_radius: float
    order: object
    ecs_angle: float
    interior_weights: np.ndarray = None
    arguments: dict = None       # make_grid arguments, for the grid's real version

    def real_version(self):
        """The same grid without complex scaling (python/run3d.py:24-25)."""
        args = dict(self.arguments)
        args['ecs_angle'] = 0.
        return make_grid(**args)


@lru_cache(None)
def lobatto(order):
    """Gauss-Lobatto-Legendre nodes, weights and differentiation matrix.

    Interior nodes are the zeros of P'_order, i.e. of the Jacobi polynomial
    P^{(1,1)}_{order-1}; weights are 2/(p(p+1) P_p(x)^2) (python/fedvr.py:30-34).
    """
    inner = roots_jacobi(order - 1, 1, 1)[0]
    x = np.concatenate(([-1.0], inner, [1.0]))
    w = 2.0 / (order * (order + 1) * eval_legendre(order, x) ** 2)
    return x, w, derivative(x)


def _laguerre_tail(tail):
    """Laguerre-Radau tail nodes/weights and the scaled derivative matrix.

    Nodes are 0 plus the zeros of L^{(1)}_tail; the basis functions carry the
    exp(-y/2) decay.  The derivative diagonal is set analytically (1/2 away from
    the origin, -tail/2 at it) instead of by row sums, which lose precision at
    far Laguerre nodes (python/fedvr.py:55-67).
    """
    y = np.concatenate(([0.0], roots_genlaguerre(tail, 1)[0]))
    wy = 1.0 / ((tail + 1) * eval_laguerre(tail, y) ** 2)
    dy = derivative(y)
    np.fill_diagonal(dy, 0.5)
    dy[0, 0] = -tail / 2.0
    scaled = dy * np.exp((y[None, :] - y[:, None]) / 2.0) - 0.5 * np.eye(tail + 1)
    return y, wy * np.exp(y), scaled


def make_grid(edges, order=8, ecs_angle=0.0, tail=0, alpha=1.0):
    """Build the FE-DVR grid (python/fedvr.py:36-77); same arguments and errors."""
    edges = np.asarray(edges, dtype=float)
    bad_edges = (edges.ndim != 1 or len(edges) < 2 or not np.all(np.isfinite(edges))
                 or edges[0] != 0 or np.any(np.diff(edges) <= 0))
    if bad_edges or tail < 0 or alpha <= 0:
        raise ValueError('increasing edges starting at 0; order>=2; tail>=0; alpha>0')
    nel = len(edges) - 1
    orders = np.full(nel, order) if np.ndim(order) == 0 else np.asarray(order)
    if orders.shape != (nel,) or np.any(orders < 2) or np.any(orders != np.round(orders)):
        raise ValueError('one integer order >=2 is required per finite element')
    orders = orders.astype(int)
    if not 0 <= ecs_angle < np.pi / 2:
        raise ValueError('absorbing ECS angle must be in [0,pi/2)')
    if ecs_angle and not tail:
        raise ValueError('ECS requires a tail element')

    total = int(orders.sum()) + 1 + tail
    r = np.zeros(total, complex)
    w = np.zeros(total, complex)
    K = np.zeros((total, total), complex)
    start = 0
    for e in range(nel):
        p = int(orders[e])
        a, b = edges[e], edges[e + 1]
        x, wx, D = lobatto(p)
        sl = slice(start, start + p + 1)
        h = b - a
        r[sl] = a + h * (x + 1.0) / 2.0
        w[sl] += wx * h / 2.0
        K[sl, sl] += D.T @ (wx[:, None] * D) / h
        start += p
    interior = w.real.copy()
    if tail:
        y, wp, d = _laguerre_tail(tail)
        scale = np.exp(1j * ecs_angle) / alpha
        sl = slice(total - tail - 1, total)
        r[sl] = edges[-1] + scale * y
        w[sl] += scale * wp
        K[sl, sl] += d.T @ (wp[:, None] * d) / (2.0 * scale)
    keep = np.arange(1, total if tail else total - 1)
    r = r[keep]
    w = w[keep]
    K = K[np.ix_(keep, keep)]
    K /= np.sqrt(w[:, None] * w[None, :])
    K[np.abs(K) < 1e-13] = 0
    label = int(orders[0]) if np.ndim(order) == 0 else tuple(int(o) for o in orders)
    arguments = {'edges': edges.copy(), 'order': order, 'ecs_angle': ecs_angle, 'tail': tail, 'alpha': alpha}
    return Grid(r, w, csr_matrix(K), float(edges[-1]), label, ecs_angle, interior[keep], arguments)
