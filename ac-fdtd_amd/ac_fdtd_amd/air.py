"""ISO 9613-1 air absorption and the relaxation processes that carry it in the time domain.

Follows the reference's ``src/ac_fdtd/air.py:67-237``. Each relaxation process ``nu`` adds one
cell-sized state ``psi_nu`` driven by the velocity divergence,

    tau_nu dpsi/dt + psi = rho c_0^2 Delta_nu div v,

whose attenuation has exactly the shape of an ISO 9613-1 relaxation term, so the strengths are
read off the standard: ``Delta_nu = C_nu c / pi``, ``tau_nu = 1 / (2 pi F_nu)``. The classical
``f^2`` term is the same shape with its relaxation frequency placed at the Nyquist frequency of
the time step. With both switches on there are three processes: O2, N2, classical.

The ISO constants and the order of every arithmetic step are kept as in the reference so that
the coefficients handed to the kernels agree with it to the last bit.
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

__all__ = ["AirAbsorption", "RelaxationProcess"]

_P_REF = 101325.0  # Pa, ISO 9613-1 reference static pressure
_T_REF = 293.15  # K, reference temperature
_T_TRIPLE = 273.16  # K, triple point of water
_KELVIN = 273.15
_DB_PER_NEPER = 8.685889638065035  # 20 log10(e)


@dataclass(frozen=True)
class RelaxationProcess:
    """``strength`` = Delta (fractional c^2 increase), ``relaxation_time`` = tau in s."""

    strength: float
    relaxation_time: float

    @property
    def frequency(self) -> float:
        return 1.0 / (2.0 * math.pi * self.relaxation_time)

    def attenuation(self, frequencies: np.ndarray, sound_speed: float) -> np.ndarray:
        """Np/m of this process alone: ``(pi Delta / c) f^2 F / (F^2 + f^2)`` (air.py:52-64)."""
        relax = self.frequency
        return (
            self.strength
            * math.pi
            / sound_speed
            * frequencies**2
            * relax
            / (relax**2 + frequencies**2)
        )


@dataclass(frozen=True)
class AirAbsorption:
    """Atmospheric conditions (deg C, % RH, Pa) and which loss terms to model (air.py:67-112)."""

    temperature: float = 20.0
    relative_humidity: float = 50.0
    static_pressure: float = _P_REF
    classical: bool = True
    relaxation: bool = True

    def __post_init__(self) -> None:
        if not 0.0 <= self.relative_humidity <= 100.0:
            raise ValueError(f"relative_humidity is a percentage, got {self.relative_humidity}")
        if self.temperature <= -_KELVIN:
            raise ValueError(f"temperature must be above absolute zero, got {self.temperature}")

    @property
    def kelvin(self) -> float:
        return self.temperature + _KELVIN

    @property
    def _pa(self) -> float:
        return self.static_pressure / _P_REF

    @property
    def _tr(self) -> float:
        return self.kelvin / _T_REF

    @property
    def molar_water_concentration(self) -> float:
        """``h`` in molar %, from the ISO saturation-pressure fit (air.py:137-146)."""
        exponent = -6.8346 * (_T_TRIPLE / self.kelvin) ** 1.261 + 4.6151
        return self.relative_humidity * 10.0**exponent / self._pa

    @property
    def oxygen_relaxation_frequency(self) -> float:
        h = self.molar_water_concentration
        return self._pa * (24.0 + 4.04e4 * h * (0.02 + h) / (0.391 + h))

    @property
    def nitrogen_relaxation_frequency(self) -> float:
        h = self.molar_water_concentration
        return (
            self._pa
            * self._tr**-0.5
            * (9.0 + 280.0 * h * math.exp(-4.170 * (self._tr ** (-1 / 3) - 1.0)))
        )

    @property
    def _classical(self) -> float:
        """``B`` of the classical term ``B f^2`` in Np/(m Hz^2)."""
        return 1.84e-11 / self._pa * self._tr**0.5

    @property
    def _oxygen(self) -> float:
        return self._tr**-2.5 * 0.01275 * math.exp(-2239.1 / self.kelvin)

    @property
    def _nitrogen(self) -> float:
        return self._tr**-2.5 * 0.1068 * math.exp(-3352.0 / self.kelvin)

    def _relaxation_terms(self):
        return (
            (self._oxygen, self.oxygen_relaxation_frequency),
            (self._nitrogen, self.nitrogen_relaxation_frequency),
        )

    def attenuation(self, frequencies: np.ndarray) -> np.ndarray:
        """ISO 9613-1 attenuation in dB/m — the target, not the model (air.py:177-190)."""
        f = np.asarray(frequencies, dtype=float)
        total = np.zeros_like(f)
        if self.classical:
            total += self._classical * f**2
        if self.relaxation:
            for coefficient, relax in self._relaxation_terms():
                total += coefficient * f**2 * relax / (relax**2 + f**2)
        return _DB_PER_NEPER * total

    def processes(self, sound_speed: float, dt: float) -> tuple[RelaxationProcess, ...]:
        """O2, N2 (if ``relaxation``), then classical (if ``classical``) (air.py:192-218)."""
        out = []
        if self.relaxation:
            for coefficient, relax in self._relaxation_terms():
                out.append(
                    RelaxationProcess(
                        strength=coefficient * sound_speed / math.pi,
                        relaxation_time=1.0 / (2.0 * math.pi * relax),
                    )
                )
        if self.classical:
            nyquist = 0.5 / dt
            out.append(
                RelaxationProcess(
                    strength=self._classical * nyquist * sound_speed / math.pi,
                    relaxation_time=1.0 / (2.0 * math.pi * nyquist),
                )
            )
        return tuple(out)

    def model_attenuation(self, frequencies: np.ndarray, sound_speed: float, dt: float):
        """dB/m of the implemented relaxing medium, before discretisation (air.py:220-233)."""
        f = np.asarray(frequencies, dtype=float)
        total = np.zeros_like(f)
        for process in self.processes(sound_speed, dt):
            total += process.attenuation(f, sound_speed)
        return _DB_PER_NEPER * total

    def total_strength(self, sound_speed: float, dt: float) -> float:
        return sum(process.strength for process in self.processes(sound_speed, dt))
