"""Excitation signals and deconvolution (reference ``src/ac_fdtd/sources.py:40-108``).

A source signal is a volume velocity per unit volume (1/s), one sample per step; sample ``n``
acts on the step from ``p^n`` to ``p^{n+1}``.
"""

from __future__ import annotations

import numpy as np

__all__ = ["band_limited_impulse", "deconvolve", "gaussian_derivative_pulse", "gaussian_pulse"]

_GAUSSIAN_BANDWIDTH_DB = 40.0


def gaussian_pulse(n_steps: int, dt: float, max_frequency: float) -> np.ndarray:
    """``exp(-((t - 4w)/w)^2)`` at ``t = (n + 1/2) dt`` with ``w = sqrt(ln 100) / (pi f)``.

    The spectrum is 40 dB down at ``max_frequency`` (sources.py:40-51).
    """
    width = np.sqrt(np.log(10.0 ** (_GAUSSIAN_BANDWIDTH_DB / 20.0))) / (np.pi * max_frequency)
    instants = (np.arange(n_steps) + 0.5) * dt
    return np.exp(-(((instants - 4.0 * width) / width) ** 2))


def gaussian_derivative_pulse(n_steps: int, dt: float, width: float) -> np.ndarray:
    """``-(t - t0)/w^2 * exp(-((t - t0)/w)^2)`` with ``t0 = 4 w``, at ``t = (n + 1/2) dt``.

    Zero-mean, so it injects no net volume: the throughput workloads use it with ``w`` set
    from a width in cells (``w = w_cells dx / c``). Not in the reference, which has only the
    Gaussian itself.
    """
    instants = (np.arange(n_steps) + 0.5) * dt
    shifted = instants - 4.0 * width
    return -shifted / width**2 * np.exp(-((shifted / width) ** 2))


def band_limited_impulse(n_steps: int, dt: float, max_frequency: float) -> np.ndarray:
    """Hann-windowed sinc flat to ``max_frequency``, ``2h + 1`` taps, ``h = ceil(4 / (f dt))``.

    Zero after the kernel (sources.py:54-77).
    """
    if max_frequency >= 0.5 / dt:
        raise ValueError(
            f"a band limit of {max_frequency} Hz does not lie below this time step's Nyquist "
            f"frequency, {0.5 / dt} Hz"
        )
    half = int(np.ceil(4.0 / (max_frequency * dt)))
    taps = 2 * half + 1
    if taps > n_steps:
        raise ValueError(
            f"{n_steps} steps cannot hold the {taps}-tap impulse band-limited to "
            f"{max_frequency} Hz"
        )
    band = 2.0 * max_frequency * dt
    kernel = band * np.sinc(band * np.arange(-half, half + 1))
    kernel *= np.hanning(taps)
    signal = np.zeros(n_steps)
    signal[:taps] = kernel
    return signal


def deconvolve(
    recorded: np.ndarray,
    excitation: np.ndarray,
    dt: float,
    max_frequency: float,
    regularisation: float = 1e-6,
) -> np.ndarray:
    """Regularised spectral division of the excitation out of a recording (sources.py:80-108)."""
    recorded = np.asarray(recorded, dtype=float)
    n = recorded.shape[-1]
    spectrum = np.fft.rfft(recorded, axis=-1)
    source = np.fft.rfft(np.asarray(excitation, dtype=float))
    power = np.abs(source) ** 2
    response = spectrum * np.conjugate(source) / (power + regularisation * power.max())
    response[..., np.fft.rfftfreq(n, dt) > max_frequency] = 0.0
    return np.fft.irfft(response, n=n, axis=-1)
