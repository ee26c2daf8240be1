"""Condon-Shortley Gaunt algebra for the product spherical-harmonic channels.

Mirror of python/angular.py:8-35.  Coefficients are evaluated exactly with
sympy's Wigner 3j symbols and rounded to float once, as the reference does,
so the Coulomb and dipole edge coefficients agree bit for bit.
"""
from functools import lru_cache

import numpy as np
from scipy.special import eval_genlaguerre, factorial
from sympy.physics.wigner import wigner_3j


@lru_cache(None)
def C(bra, k, q, ket):
    """<l m| C^k_q |j n> with C^k_q = sqrt(4 pi/(2k+1)) Y_kq (python/angular.py:8-12)."""
    l, m = bra
    j, n = ket
    if m != n + q or abs(q) > k or not abs(l - j) <= k <= l + j or (l + k + j) % 2:
        return 0.0
    parity_3j = wigner_3j(l, k, j, 0, 0, 0)
    magnetic_3j = wigner_3j(l, k, j, -m, q, n)
    return float((-1) ** m * np.sqrt((2 * l + 1) * (2 * j + 1)) * parity_3j * magnetic_3j)


def cartesian(bra, ket):
    """Cartesian (x, y, z) components of the rank-1 tensor C^1 (python/angular.py:14-16)."""
    minus = C(bra, 1, -1, ket)
    plus = C(bra, 1, 1, ket)
    root2 = np.sqrt(2)
    return np.array([(minus - plus) / root2, 1j * (minus + plus) / root2, C(bra, 1, 0, ket)], complex)


def coulomb(bra, ket, lam):
    """Angular factor of the order-``lam`` multipole of 1/r12 (python/angular.py:18-24).

    Only q = m1 - m1' survives magnetic conservation.
    """
    (l1, m1), (l2, m2) = bra
    (k1, n1), (k2, n2) = ket
    q = m1 - n1
    if abs(q) > lam or m2 - n2 != -q:
        return 0.0
    first = C((l1, m1), lam, q, (k1, n1))
    if not first:
        return 0.0
    return (-1.0) ** q * first * C((l2, m2), lam, -q, (k2, n2))


def channels(lmax, M=0):
    """Product channels ((l1,m1),(l2,m2)) with m1+m2 == M, or all when M is None."""
    if lmax < 0:
        raise ValueError('lmax>=0')
    single = [(l, m) for l in range(lmax + 1) for m in range(-l, l + 1)]
    return [(a, b) for a in single for b in single if M is None or a[1] + b[1] == M]


def radial_hydrogen(n, l, r, Z=2.0):
    """Hydrogenic radial function R_nl(r) for charge Z (python/angular.py:31-35)."""
    if n < 1 or l < 0 or l >= n:
        raise ValueError('requires 0<=l<n')
    x = 2 * Z * np.asarray(r) / n
    norm = (2 * Z / n) ** 1.5 * np.sqrt(factorial(n - l - 1) / (2 * n * factorial(n + l)))
    return norm * np.exp(-x / 2) * x ** l * eval_genlaguerre(n - l - 1, 2 * l + 1, x)


# ---------------------------------------------------------------------------
# Racah algebra for the coupled bases.  Every coefficient is formed exactly in
# sympy and rounded to float once.

@lru_cache(None)
def _reduced_C_exact(l, k, lp):
    """<l||C^k||l'> = (-1)^l sqrt((2l+1)(2l'+1)) (l k l'; 0 0 0) (Edmonds 5.4.5)."""
    from sympy import sqrt
    return (-1) ** l * sqrt((2 * l + 1) * (2 * lp + 1)) * wigner_3j(l, k, lp, 0, 0, 0)


@lru_cache(None)
def clebsch(l1, m1, l2, m2, L, M):
    """<l1 m1 l2 m2 | L M> as a float."""
    from sympy.physics.wigner import clebsch_gordan
    return float(clebsch_gordan(l1, l2, L, m1, m2, M))


@lru_cache(None)
def scalar_coupled(l1, l2, k1, k2, L, k):
    """<(l1 l2) L M | C^k(1) . C^k(2) | (k1 k2) L M>, independent of M (Edmonds 7.1.6).

    The multipole factor of 1/r12 between two coupled channels of equal (L, M)."""
    from sympy.physics.wigner import wigner_6j
    if (l1 + k1 + k) % 2 or (l2 + k2 + k) % 2:
        return 0.0
    value = ((-1) ** (k1 + l2 + L) * wigner_6j(l1, l2, L, k2, k1, k)
             * _reduced_C_exact(l1, k, k1) * _reduced_C_exact(l2, k, k2))
    return float(value)


@lru_cache(None)
def tensor_coupled(bra, ket, k, q, electron):
    """<(l1 l2) L M | C^k_q(electron) | (k1 k2) J N> in the coupled basis.

    Wigner-Eckart in the two-electron space, with the reduced element of a
    tensor acting on one part of a coupled pair (Edmonds 7.1.7 / 7.1.8)."""
    from sympy import sqrt
    from sympy.physics.wigner import wigner_6j
    l1, l2, L, M = bra
    k1, k2, J, N = ket
    if M != N + q:
        return 0.0
    if electron == 1:
        if l2 != k2 or (l1 + k1 + k) % 2:
            return 0.0
        part = (-1) ** (l1 + l2 + J + k) * wigner_6j(l1, L, l2, J, k1, k) * _reduced_C_exact(l1, k, k1)
    else:
        if l1 != k1 or (l2 + k2 + k) % 2:
            return 0.0
        part = (-1) ** (l1 + k2 + L + k) * wigner_6j(l2, L, l1, J, k2, k) * _reduced_C_exact(l2, k, k2)
    value = (-1) ** (L - M) * wigner_3j(L, k, J, -M, q, N) * sqrt((2 * L + 1) * (2 * J + 1)) * part
    return float(value)


def spherical_to_cartesian(minus, zero, plus):
    """(x, y, z) components of a rank-1 tensor from its q = -1, 0, +1 parts."""
    root2 = np.sqrt(2)
    return ((minus - plus) / root2, 1j * (minus + plus) / root2, zero)
