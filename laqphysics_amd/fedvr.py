"""Normalised finite-element DVR radial grid with an optional irECS Laguerre tail.

Produces the grid of the reference's ``make_grid`` (python/fedvr.py:36-77):
same arguments, same ``Grid`` fields, the same operator up to rounding.  The
construction here is its own:

* every element is a ``_Element``: quadrature nodes ``x``, weights ``w`` and
  the derivative matrix ``E[q, k] = chi_k'(x_q)`` of its basis functions in
  the element's own coordinate.  The kinetic block is ``E^T diag(w) E`` over
  the element's Jacobian;
* Gauss-Lobatto-Legendre elements use the classical closed form of the GLL
  differentiation matrix, ``P_p(x_q) / (P_p(x_k) (x_q - x_k))`` off the
  diagonal and ``-+ p (p + 1) / 4`` at the two end nodes (zero elsewhere);
* the exterior-complex-scaled tail is one Laguerre-Radau element whose basis
  functions carry ``exp(-(y - y_k)/2)``.  Its nodes are the zeros of
  ``q(y) = y L_N^(1)(y)``, so the Lagrange derivatives follow from ``q'`` at the
  nodes, and the derivative of the decaying basis has an exactly known
  diagonal: ``0`` at ``y > 0`` and ``-(N + 1)/2`` at the bridge ``y = 0``;
* blocks are assembled as COO triplets (continuity through the shared bridge
  nodes), the Dirichlet nodes are dropped (the origin, and the outer edge when
  there is no tail), and the matrix is scaled by ``1/sqrt(w)`` on both sides
  with the complex square root, then entries below 1e-13 are removed.
"""
from dataclasses import dataclass
from functools import lru_cache

import numpy as np
from scipy.sparse import coo_matrix, csr_matrix
from scipy.special import eval_genlaguerre, eval_laguerre, eval_legendre, roots_genlaguerre, roots_jacobi

DROP = 1e-13   # kinetic entries below this magnitude are removed (python/fedvr.py:75)


def derivative(nodes):
    """Lagrange differentiation matrix ``D[i, k] = l_k'(x_i)`` on arbitrary nodes.

    Uses the node polynomial ``pi(x) = prod (x - x_m)``:
    ``l_k'(x_i) = pi'(x_i) / (pi'(x_k) (x_i - x_k))`` for ``i != k`` and
    ``l_i'(x_i) = sum_{m != i} 1 / (x_i - x_m)``.
    """
    x = np.asarray(nodes, dtype=float)
    diff = x[:, None] - x[None, :]
    off = ~np.eye(len(x), dtype=bool)
    safe = np.where(off, diff, 1.0)
    dpi = np.prod(safe, axis=1)                  # pi'(x_i)
    out = np.where(off, dpi[:, None] / (dpi[None, :] * safe), 0.0)
    out[np.diag_indices(len(x))] = np.sum(np.where(off, 1.0 / safe, 0.0), axis=1)
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


@dataclass(frozen=True)
class _Element:
    x: np.ndarray      # nodes in the element's own coordinate
    w: np.ndarray      # quadrature weights there
    E: np.ndarray      # E[q, k] = derivative of basis function k at node q


@lru_cache(None)
def _gll(order):
    """Gauss-Lobatto-Legendre element of polynomial order ``order`` on [-1, 1]."""
    x = np.concatenate(([-1.0], roots_jacobi(order - 1, 1, 1)[0], [1.0]))
    P = eval_legendre(order, x)
    w = 2.0 / (order * (order + 1) * P * P)
    gap = x[:, None] - x[None, :]
    np.fill_diagonal(gap, 1.0)
    E = (P[:, None] / P[None, :]) / gap
    np.fill_diagonal(E, 0.0)
    corner = order * (order + 1) / 4.0
    E[0, 0] = -corner
    E[-1, -1] = corner
    return _Element(x, w, E)


def lobatto(order):
    """(nodes, weights, derivative matrix) of the GLL element (python/fedvr.py:30-34)."""
    el = _gll(order)
    return el.x, el.w, el.E


@lru_cache(None)
def _radau_laguerre(points):
    """Laguerre-Radau element with ``points`` nodes after the bridge node y = 0.

    Basis ``chi_k(y) = l_k(y) exp(-(y - y_k)/2)``; the weights carry the
    matching ``exp(y)``.  Returns nodes, weights and ``E[q, k] = chi_k'(y_q)``.
    """
    N = points
    y = np.concatenate(([0.0], roots_genlaguerre(N, 1)[0]))
    w = np.exp(y) / ((N + 1) * eval_laguerre(N, y) ** 2)
    # q(y) = y L_N^(1)(y): q'(0) = L_N^(1)(0) = N + 1, and at a zero y_k of L_N^(1),
    # q'(y_k) = y_k d/dy L_N^(1)(y_k) = -y_k L_{N-1}^(2)(y_k)
    dq = np.empty(N + 1)
    dq[0] = N + 1.0
    dq[1:] = -y[1:] * eval_genlaguerre(N - 1, 2, y[1:])
    gap = y[:, None] - y[None, :]
    np.fill_diagonal(gap, 1.0)
    E = (dq[:, None] / dq[None, :]) / gap * np.exp((y[None, :] - y[:, None]) / 2.0)
    np.fill_diagonal(E, 0.0)
    E[0, 0] = -(N + 1) / 2.0
    return _Element(y, w, E)


def _check(edges, order, ecs_angle, tail, alpha):
    edges = np.asarray(edges, dtype=float)
    if (edges.ndim != 1 or edges.size < 2 or not np.isfinite(edges).all() or edges[0] != 0
            or (np.diff(edges) <= 0).any() or tail < 0 or alpha <= 0):
        raise ValueError('increasing edges starting at 0; order>=2; tail>=0; alpha>0')
    count = edges.size - 1
    orders = np.asarray(order) if np.ndim(order) else np.full(count, order)
    if orders.shape != (count,) or (orders < 2).any() or (orders != np.round(orders)).any():
        raise ValueError('one integer order >=2 is required per finite element')
    if not 0 <= ecs_angle < np.pi / 2:
        raise ValueError('absorbing ECS angle must be in [0,pi/2)')
    if ecs_angle and not tail:
        raise ValueError('ECS requires a tail element')
    return edges, orders.astype(int)


def make_grid(edges, order=8, ecs_angle=0.0, tail=0, alpha=1.0):
    """FE-DVR grid on ``edges`` with per-element ``order`` and an optional
    Laguerre-Radau tail of ``tail`` points scaled by ``exp(i ecs_angle)/alpha``
    (the reference's ``make_grid``, python/fedvr.py:36-77)."""
    edges, orders = _check(edges, order, ecs_angle, tail, alpha)
    # node k of element e sits at global index start[e] + k; elements share bridges
    starts = np.concatenate(([0], np.cumsum(orders)))
    size = int(starts[-1]) + 1 + tail
    r = np.zeros(size, complex)
    mass = np.zeros(size, complex)
    rows, cols, vals = [], [], []
    for e, p in enumerate(orders):
        el = _gll(int(p))
        a, b = edges[e], edges[e + 1]
        half = (b - a) / 2.0
        at = starts[e] + np.arange(p + 1)
        r[at] = a + half * (el.x + 1.0)
        mass[at] += half * el.w
        # (1/2) int chi_k' chi_l' dr with dr = half dx: sum_q w_q E_qk E_ql / (2 half)
        block = el.E.T @ (el.w[:, None] * el.E) / (b - a)
        rows.append(np.repeat(at, p + 1))
        cols.append(np.tile(at, p + 1))
        vals.append(block.ravel())
    real_mass = mass.real.copy()
    if tail:
        el = _radau_laguerre(tail)
        scale = np.exp(1j * ecs_angle) / alpha
        at = size - tail - 1 + np.arange(tail + 1)
        r[at] = edges[-1] + scale * el.x
        mass[at] += scale * el.w
        # dr = scale dy: (1/2) sum_q (scale w_q) E_qk E_ql / scale^2
        block = el.E.T @ (el.w[:, None] * el.E) / (2.0 * scale)
        rows.append(np.repeat(at, tail + 1))
        cols.append(np.tile(at, tail + 1))
        vals.append(block.ravel())
    K = coo_matrix((np.concatenate(vals).astype(complex), (np.concatenate(rows), np.concatenate(cols))),
                   shape=(size, size)).toarray()
    # Dirichlet at the origin; at the outer edge only when no tail carries the
    # functions to infinity
    keep = slice(1, size if tail else size - 1)
    r, mass, real_mass = r[keep], mass[keep], real_mass[keep]
    s = 1.0 / np.sqrt(mass)
    K = s[:, None] * K[keep, keep] * s[None, :]
    K[np.abs(K) < DROP] = 0.0
    label = int(orders[0]) if not np.ndim(order) else tuple(int(o) for o in orders)
    arguments = {'edges': edges.copy(), 'order': order, 'ecs_angle': ecs_angle, 'tail': tail, 'alpha': alpha}
    return Grid(r, mass, csr_matrix(K), float(edges[-1]), label, ecs_angle, real_mass, arguments)
