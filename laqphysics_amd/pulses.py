"""sin^2-envelope pulses: E(t) and its exact vector potential A(t) = -int E.

Mirror of python/pulses.py:5-48 (same fields, validation and closed forms) and
of the polarisation rules of python/run3d.py:41-70.
"""
from dataclasses import dataclass, replace

import numpy as np


@dataclass(frozen=True)
class Pulse:
    omega: float = 1.5
    cycles: float = 60
    field: float = .0534
    start: float = 0.
    cep: float = 0.

    def __post_init__(self):
        values = [self.omega, self.cycles, self.field, self.start, self.cep]
        if (not np.all(np.isfinite(values)) or self.omega <= 0 or self.cycles <= 1
                or self.field < 0 or self.start < 0):
            raise ValueError('finite omega>0, cycles>1, field>=0, start>=0 and finite CEP are required')

    @property
    def duration(self):
        return 2 * np.pi * self.cycles / self.omega

    def envelope(self, t):
        s = np.asarray(t) - self.start
        inside = (s >= 0) & (s <= self.duration)
        return np.where(inside, np.sin(np.pi * s / self.duration) ** 2, 0.)

    def electric(self, t):
        return self.field * self.envelope(t) * np.cos(self.omega * (np.asarray(t) - self.start) + self.cep)

    def vector(self, t):
        """Exact A(t) = -int_start^t E (python/pulses.py:28-40).

        With sin^2(pi u/T) = 1/2 - cos(b u)/2 (b = 2 pi/T) the integrand is a sum
        of three cosines; int_0^u cos(cep + w s) ds = u cos(cep + w u/2) sinc(w u/2 pi)
        stays accurate when w -> 0.  Integer cycle counts >= 2 give exactly zero
        area, so endpoint round-off is removed after the pulse.
        """
        u = np.clip(np.asarray(t) - self.start, 0, self.duration)
        b = 2 * np.pi / self.duration

        def cosine_integral(w):
            return u * np.cos(self.cep + w * u / 2) * np.sinc(w * u / (2 * np.pi))

        value = -self.field * (.5 * cosine_integral(self.omega) - .25 * cosine_integral(self.omega + b)
                               - .25 * cosine_integral(self.omega - b))
        if round(self.cycles) >= 2 and abs(self.cycles - round(self.cycles)) < 1e-12:
            value = np.where(np.asarray(t) >= self.start + self.duration, 0., value)
        return value

    def area_envelope(self, t):
        u = np.clip(np.asarray(t) - self.start, 0, self.duration)
        return u / 2 - self.duration * np.sin(2 * np.pi * u / self.duration) / (4 * np.pi)

    def area_squared(self, t):
        u = np.clip(np.asarray(t) - self.start, 0, self.duration)
        T = self.duration
        return 3 * u / 8 - T * np.sin(2 * np.pi * u / T) / (4 * np.pi) + T * np.sin(4 * np.pi * u / T) / (32 * np.pi)


def vector_function(pulses):
    """A(t) as a 3-vector for ``[(Pulse, polarization), ...]`` (python/run3d.py:41-56).

    'z' is linear; 'sigma+' / 'sigma-' add A_x = A/sqrt2 and
    A_y = +-A(cep - pi/2)/sqrt2 (sigma+ is Delta m = +1 absorption).
    """
    pulses = list(pulses)
    for p, pol in pulses:
        if pol not in ('z', 'sigma+', 'sigma-'):
            raise ValueError('polarization must be z, sigma+, or sigma-')

    def avec(t):
        v = np.zeros(3)
        for p, pol in pulses:
            if pol == 'z':
                v[2] += p.vector(t)
            else:
                v[0] += p.vector(t) / np.sqrt(2)
                v[1] += replace(p, cep=p.cep - np.pi / 2).vector(t) / np.sqrt(2) * (1 if pol == 'sigma+' else -1)
        return v
    return avec


def electric_function(pulses):
    """E(t) as a 3-vector (python/run3d.py:58-70)."""
    pulses = list(pulses)

    def electric(t):
        f = np.zeros(3)
        for p, pol in pulses:
            if pol == 'z':
                f[2] += p.electric(t)
            elif pol in ('sigma+', 'sigma-'):
                f[0] += p.electric(t) / np.sqrt(2)
                f[1] += replace(p, cep=p.cep - np.pi / 2).electric(t) / np.sqrt(2) * (1 if pol == 'sigma+' else -1)
            else:
                raise ValueError('unsupported polarization')
        return f
    return electric
