"""Frequency-dependent wall impedance: passive admittance models, fitting them to octave-band
absorption, and their discrete (IIR) form for the time loop.

The reference plans "frequency-dependent impedance via an IIR filter fitted to octave-band
absorption data" (``PLAN.md:21-24``) and explains why a real admittance is only the first step
(``src/ac_fdtd/boundaries.py:1-30``, ``:76-87``); it does not implement it. Here a wall's
normalised admittance ``xi = rho c Y`` is

    xi(s) = g + sum_b 1 / (m_b s + r_b + k_b / s),     g, m_b, r_b, k_b >= 0,

a conductance in parallel with mass-spring-damper branches. Every term is positive real, so
the wall can only absorb energy, at any frequency. Each branch is discretised by the bilinear
transform at the time step, which maps positive-real to positive-real, so the discrete wall is
passive too, and gives a biquad ``Y_b(z) = beta (1 - z^-2) / (A0 + A1 z^-1 + A2 z^-2)`` relating
the face's mean pressure ``(p^{n+1} + p^n) / 2`` to its outward velocity at ``n + 1/2``. The
instantaneous part ``g + sum_b beta / A0`` enters the reference's time-centred wall closed form
exactly as a real admittance does (``a = c dt xi / (2 dx)``); the branches' memory (their filter
states) enters as a known velocity on the wall face, which the pressure sweep's divergence
already reads. :mod:`ac_fdtd_amd.iir_walls` runs it.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field
from functools import cached_property

import numpy as np

__all__ = [
    "FrequencyDependentWall",
    "FrequencyDependentWalls",
    "ImpedanceBranch",
    "absorption_from_admittance",
    "fit_wall",
]

#: The octave bands absorption tables are quoted in, Hz.
OCTAVE_CENTRES = (125.0, 250.0, 500.0, 1000.0, 2000.0, 4000.0)


@dataclass(frozen=True)
class ImpedanceBranch:
    """One branch ``1 / (mass s + resistance + stiffness / s)`` of a normalised admittance:
    ``mass`` in s, ``resistance`` dimensionless, ``stiffness`` in 1/s; all non-negative, mass
    and stiffness positive (a branch resonating at ``sqrt(stiffness / mass)`` rad/s)."""

    mass: float
    resistance: float
    stiffness: float

    def __post_init__(self) -> None:
        if not (self.mass > 0.0 and self.stiffness > 0.0 and self.resistance >= 0.0):
            raise ValueError("an impedance branch needs mass > 0, stiffness > 0 and "
                             f"resistance >= 0; got {self}")

    def admittance(self, omega: np.ndarray) -> np.ndarray:
        s = 1j * np.asarray(omega, dtype=float)
        return 1.0 / (self.mass * s + self.resistance + self.stiffness / s)

    def biquad(self, dt: float) -> tuple[float, float, float, float]:
        """``(b0, b2, a1, a2)`` of the bilinear-transformed branch:
        ``y[n] = b0 x[n] + b2 x[n-2] - a1 y[n-1] - a2 y[n-2]`` (``b1 = 0``)."""
        g = 2.0 / dt
        mg2, rg, k = self.mass * g * g, self.resistance * g, self.stiffness
        a0 = mg2 + rg + k
        return g / a0, -g / a0, (2.0 * k - 2.0 * mg2) / a0, (mg2 - rg + k) / a0


def absorption_from_admittance(xi: np.ndarray) -> np.ndarray:
    """Normal-incidence absorption ``1 - |(1 - xi) / (1 + xi)|^2`` of a normalised admittance."""
    xi = np.asarray(xi)
    return 1.0 - np.abs((1.0 - xi) / (1.0 + xi)) ** 2


@dataclass(frozen=True)
class FrequencyDependentWall:
    """A passive locally reacting wall: conductance ``conductance`` (the frequency-independent
    part of ``xi``) in parallel with resonant ``branches``."""

    conductance: float = 0.0
    branches: tuple[ImpedanceBranch, ...] = field(default_factory=tuple)

    def __post_init__(self) -> None:
        if self.conductance < 0.0:
            raise ValueError(f"a passive wall needs conductance >= 0, got {self.conductance}")

    def admittance(self, frequencies) -> np.ndarray:
        """Normalised admittance ``xi(j 2 pi f)`` (complex)."""
        omega = 2.0 * math.pi * np.asarray(frequencies, dtype=float)
        total = np.full(omega.shape, self.conductance, dtype=complex)
        for branch in self.branches:
            total = total + branch.admittance(omega)
        return total

    def discrete_admittance(self, frequencies, dt: float) -> np.ndarray:
        """The admittance the time loop realises at ``frequencies``: each branch's biquad on the
        unit circle ``z = exp(j 2 pi f dt)``."""
        z1 = np.exp(-2j * math.pi * np.asarray(frequencies, dtype=float) * dt)
        total = np.full(z1.shape, self.conductance, dtype=complex)
        for branch in self.branches:
            b0, b2, a1, a2 = branch.biquad(dt)
            total = total + (b0 + b2 * z1**2) / (1.0 + a1 * z1 + a2 * z1**2)
        return total

    def absorption(self, frequencies) -> np.ndarray:
        return absorption_from_admittance(self.admittance(frequencies))

    def instantaneous(self, dt: float) -> float:
        """The part of the discrete admittance acting on the present mean pressure:
        ``conductance + sum_b b0``."""
        return self.conductance + sum(branch.biquad(dt)[0] for branch in self.branches)

    @classmethod
    def fit(cls, octave_absorption: dict, branches: int | None = None) -> "FrequencyDependentWall":
        return fit_wall(octave_absorption, branches)


def _xi_for_absorption(alpha: float) -> float:
    """The real admittance below 1 with absorption ``alpha``: ``4 xi / (1 + xi)^2 = alpha``."""
    alpha = min(max(alpha, 1e-6), 1.0 - 1e-9)
    root = math.sqrt(1.0 - alpha)
    return (1.0 - root) / (1.0 + root)


def fit_wall(octave_absorption: dict, branches: int | None = None) -> FrequencyDependentWall:
    """Fit a passive wall to normal-incidence absorption per octave band
    (``{centre_hz: alpha}``): one resonant branch per band (or ``branches`` spread over the
    bands) and a conductance, by least squares on the absorption at each band's centre and
    edges (targets interpolated in log frequency), with every parameter kept positive by
    fitting its logarithm."""
    from scipy.optimize import least_squares

    centres = np.array(sorted(octave_absorption), dtype=float)
    targets = np.array([octave_absorption[c] for c in centres], dtype=float)
    if np.any(targets < 0.0) or np.any(targets >= 1.0):
        raise ValueError("absorption coefficients must lie in [0, 1)")
    count = len(centres) if branches is None else int(branches)
    if count < 1:
        raise ValueError("a fitted wall needs at least one branch")
    samples = np.exp(np.linspace(np.log(centres[0] / math.sqrt(2.0)),
                                 np.log(centres[-1] * math.sqrt(2.0)), 8 * len(centres)))
    wanted = np.interp(np.log(samples), np.log(centres), targets)
    places = np.exp(np.linspace(np.log(centres[0]), np.log(centres[-1]), count))
    start = []
    for f in places:
        xi = _xi_for_absorption(float(np.interp(np.log(f), np.log(centres), targets)))
        start += [math.log(0.5 / max(xi, 1e-3)), math.log(2.0 * math.pi * f), math.log(0.7)]
    start.append(math.log(0.05))

    def build(theta):
        parts = []
        for b in range(count):
            resistance, omega, q = np.exp(theta[3 * b: 3 * b + 3])
            parts.append(ImpedanceBranch(mass=resistance * q / omega, resistance=resistance,
                                         stiffness=resistance * q * omega))
        return FrequencyDependentWall(conductance=float(np.exp(theta[-1])), branches=tuple(parts))

    def residual(theta):
        return build(theta).absorption(samples) - wanted

    # Resonances kept within two octaves of the table, quality factors and resistances within
    # physical ranges: the fit then stays well inside the band the grid resolves.
    lower, upper = [], []
    for _ in range(count):
        lower += [math.log(1e-3), math.log(2.0 * math.pi * centres[0] / 4.0), math.log(0.05)]
        upper += [math.log(1e4), math.log(2.0 * math.pi * centres[-1] * 4.0), math.log(20.0)]
    lower.append(-30.0)
    upper.append(math.log(10.0))
    result = least_squares(residual, np.clip(np.array(start), lower, upper), method="trf",
                           bounds=(lower, upper), max_nfev=4000)
    return build(result.x)


@dataclass(frozen=True)
class FrequencyDependentWalls:
    """The six walls (``FACES`` order: x_min, x_max, y_min, y_max, z_min, z_max), each a
    :class:`FrequencyDependentWall`, a real normalised admittance (a float; 0 is rigid), or a
    wall fitted from ``{centre_hz: absorption}``."""

    x_min: object = 0.0
    x_max: object = 0.0
    y_min: object = 0.0
    y_max: object = 0.0
    z_min: object = 0.0
    z_max: object = 0.0

    @classmethod
    def uniform(cls, wall) -> "FrequencyDependentWalls":
        return cls(wall, wall, wall, wall, wall, wall)

    @cached_property
    def walls(self) -> tuple:
        """The six walls resolved (tables fitted once, on first use)."""
        out = []
        for value in (self.x_min, self.x_max, self.y_min, self.y_max, self.z_min, self.z_max):
            if isinstance(value, dict):
                value = fit_wall(value)
            elif not isinstance(value, FrequencyDependentWall):
                value = float(value)
                if value < 0.0:
                    raise ValueError(f"a real wall admittance must be >= 0, got {value}")
            out.append(value)
        return tuple(out)

    def instantaneous(self, dt: float) -> tuple[float, ...]:
        """Each wall's admittance on the present mean pressure (real admittances as given)."""
        return tuple(w.instantaneous(dt) if isinstance(w, FrequencyDependentWall) else w
                     for w in self.walls)

    def branch_count(self) -> int:
        return max([len(w.branches) for w in self.walls
                    if isinstance(w, FrequencyDependentWall)], default=0)

    def branch_table(self, dt: float) -> np.ndarray:
        """``(6, branches, 4)`` fp64 table of ``(b0, b2, a1, a2)``; unused entries 0 (a zero
        branch keeps zero state)."""
        table = np.zeros((6, self.branch_count(), 4))
        for face, wall in enumerate(self.walls):
            if isinstance(wall, FrequencyDependentWall):
                for b, branch in enumerate(wall.branches):
                    table[face, b] = branch.biquad(dt)
        return table

    def face_mask(self) -> int:
        return sum(1 << f for f, w in enumerate(self.walls)
                   if isinstance(w, FrequencyDependentWall) and w.branches)
