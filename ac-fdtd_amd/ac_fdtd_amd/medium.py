"""Homogeneous fluid: sound speed and density (reference ``src/ac_fdtd/medium.py:20-61``)."""

from __future__ import annotations

from dataclasses import dataclass

__all__ = ["Medium", "AIR"]

_GAS_CONSTANT_DRY_AIR = 287.058  # J/(kg K)
_KELVIN_OFFSET = 273.15


@dataclass(frozen=True)
class Medium:
    sound_speed: float
    density: float

    @property
    def impedance(self) -> float:
        """Characteristic impedance ``rho * c``; wall admittances are normalised by it."""
        return self.density * self.sound_speed

    @classmethod
    def air(cls, temperature: float = 20.0, static_pressure: float = 101325.0) -> Medium:
        """Dry air: ``c = 331.3 sqrt(T / 273.15)``, ideal-gas density (medium.py:41-57).

        The expression keeps the reference's ``** 0.5`` (a ``pow`` call, not ``sqrt``), so
        that the time step and every coefficient derived from ``c`` agree to the last bit.
        """
        kelvin = temperature + _KELVIN_OFFSET
        return cls(
            sound_speed=331.3 * (kelvin / _KELVIN_OFFSET) ** 0.5,
            density=static_pressure / (_GAS_CONSTANT_DRY_AIR * kelvin),
        )


AIR = Medium.air()
