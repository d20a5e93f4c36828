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
class _Atmosphere:
    """The ISO 9613-1 quantities of one atmosphere, worked out once: absolute temperature,
    pressure and temperature ratios to the reference state, the molar water concentration
    ``h`` (%), the O2 and N2 relaxation frequencies (Hz), and the coefficients of the classical
    term (Np/(m Hz^2)) and of the two relaxation terms (Np/m at the relaxation maximum per
    unit frequency ratio)."""

    kelvin: float
    pressure_ratio: float
    temperature_ratio: float
    water: float
    oxygen_frequency: float
    nitrogen_frequency: float
    classical: float
    oxygen: float
    nitrogen: float

    @classmethod
    def of(cls, celsius: float, humidity: float, pressure: float) -> "_Atmosphere":
        # Each expression keeps the reference's order of operations (air.py:137-175), so the
        # strengths handed to the kernels match it bit for bit.
        kelvin = celsius + _KELVIN
        pr = pressure / _P_REF
        tr = kelvin / _T_REF
        saturation_exponent = -6.8346 * (_T_TRIPLE / kelvin) ** 1.261 + 4.6151
        h = humidity * 10.0**saturation_exponent / pr
        return cls(
            kelvin=kelvin,
            pressure_ratio=pr,
            temperature_ratio=tr,
            water=h,
            oxygen_frequency=pr * (24.0 + 4.04e4 * h * (0.02 + h) / (0.391 + h)),
            nitrogen_frequency=pr * tr**-0.5
            * (9.0 + 280.0 * h * math.exp(-4.170 * (tr ** (-1 / 3) - 1.0))),
            classical=1.84e-11 / pr * tr**0.5,
            oxygen=tr**-2.5 * 0.01275 * math.exp(-2239.1 / kelvin),
            nitrogen=tr**-2.5 * 0.1068 * math.exp(-3352.0 / kelvin),
        )

    def relaxations(self) -> tuple[tuple[float, float], ...]:
        """(coefficient, relaxation frequency) of O2, then N2."""
        return ((self.oxygen, self.oxygen_frequency), (self.nitrogen, self.nitrogen_frequency))


@dataclass(frozen=True)
class AirAbsorption:
    """Atmospheric conditions (deg C, % RH, Pa) and which loss terms to model (air.py:67-112):
    ``classical`` the viscous, thermal and rotational ``f^2`` term, ``relaxation`` the O2 and
    N2 relaxations that dominate the audio band."""

    temperature: float = 20.0
    relative_humidity: float = 50.0
    static_pressure: float = _P_REF
    classical: bool = True
    relaxation: bool = True

    def __post_init__(self) -> None:
        if not 0.0 <= self.relative_humidity <= 100.0:
            raise ValueError(f"relative humidity is given in percent (0 to 100); "
                             f"{self.relative_humidity} is outside that")
        if self.temperature <= -_KELVIN:
            raise ValueError(f"{self.temperature} C lies at or below absolute zero")

    @property
    def _atmosphere(self) -> _Atmosphere:
        return _Atmosphere.of(self.temperature, self.relative_humidity, self.static_pressure)

    @property
    def kelvin(self) -> float:
        return self._atmosphere.kelvin

    @property
    def molar_water_concentration(self) -> float:
        """``h`` in molar %, from the ISO saturation-pressure fit."""
        return self._atmosphere.water

    @property
    def oxygen_relaxation_frequency(self) -> float:
        return self._atmosphere.oxygen_frequency

    @property
    def nitrogen_relaxation_frequency(self) -> float:
        return self._atmosphere.nitrogen_frequency

    def attenuation(self, frequencies: np.ndarray) -> np.ndarray:
        """ISO 9613-1 attenuation in dB/m: the curve the model is matched to (air.py:177-190)."""
        f = np.asarray(frequencies, dtype=float)
        atmosphere = self._atmosphere
        terms = []
        if self.classical:
            terms.append(atmosphere.classical * f**2)
        if self.relaxation:
            terms += [a * f**2 * fr / (fr**2 + f**2) for a, fr in atmosphere.relaxations()]
        nepers = np.zeros_like(f)
        for term in terms:
            nepers += term
        return _DB_PER_NEPER * nepers

    def processes(self, sound_speed: float, dt: float) -> tuple[RelaxationProcess, ...]:
        """The relaxation states the time step carries: O2 and N2 (``relaxation``), then the
        classical term as a process relaxing at the Nyquist frequency of ``dt``
        (``classical``): ``Delta = a c / pi``, ``tau = 1 / (2 pi F)`` (air.py:192-218)."""
        atmosphere = self._atmosphere
        pairs = list(atmosphere.relaxations()) if self.relaxation else []
        if self.classical:
            nyquist = 0.5 / dt
            pairs.append((atmosphere.classical * nyquist, nyquist))
        return tuple(
            RelaxationProcess(strength=a * sound_speed / math.pi,
                              relaxation_time=1.0 / (2.0 * math.pi * fr))
            for a, fr in pairs
        )

    def model_attenuation(self, frequencies: np.ndarray, sound_speed: float, dt: float):
        """dB/m of the relaxing medium the solver implements, before discretisation: the ISO
        curve except where the classical term's surrogate relaxation rolls off."""
        f = np.asarray(frequencies, dtype=float)
        nepers = np.zeros_like(f)
        for process in self.processes(sound_speed, dt):
            nepers += process.attenuation(f, sound_speed)
        return _DB_PER_NEPER * nepers

    def total_strength(self, sound_speed: float, dt: float) -> float:
        return sum(process.strength for process in self.processes(sound_speed, dt))
