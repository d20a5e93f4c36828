"""Room-acoustics parameters of impulse responses (reference ``src/ac_fdtd/metrics.py:41-151``).

Host-side post-processing of the recordings the GPU time loop produces: octave-band filtering,
Schroeder decay, T20/T30, EDT, clarity, and the Sabine/Eyring estimates they are compared
with. A simulated response has no noise floor, so Schroeder integration runs to the end.
"""

from __future__ import annotations

import numpy as np
from scipy import signal as scipy_signal

__all__ = [
    "air_absorption_coefficient",
    "clarity",
    "early_decay_time",
    "eyring_time",
    "octave_band",
    "octave_band_reverberation",
    "reverberation_time",
    "sabine_time",
    "schroeder_decay",
    "schroeder_frequency",
]

#: 24 ln(10) / c at 343 m/s, the constant of Sabine's formula (metrics.py:38).
_SABINE = 24.0 * np.log(10.0) / 343.0


def octave_band(response: np.ndarray, sample_rate: float, centre: float) -> np.ndarray:
    """Zero-phase 4th-order Butterworth band-pass over ``[centre/sqrt2, centre*sqrt2]``."""
    nyquist = 0.5 * sample_rate
    low = centre / np.sqrt(2.0) / nyquist
    high = min(centre * np.sqrt(2.0) / nyquist, 0.999)
    if low >= high:
        raise ValueError(f"the {centre} Hz octave band does not fit below {nyquist} Hz")
    sections = scipy_signal.butter(4, (low, high), btype="bandpass", output="sos")
    return scipy_signal.sosfiltfilt(sections, response, axis=-1)


def schroeder_decay(response: np.ndarray, floor_db: float = -120.0) -> np.ndarray:
    """Backward-integrated energy decay in dB, 0 at the start, floored at ``floor_db``."""
    energy = np.asarray(response, dtype=float) ** 2
    integrated = np.cumsum(energy[..., ::-1], axis=-1)[..., ::-1]
    with np.errstate(divide="ignore"):
        decay = 10.0 * np.log10(np.maximum(integrated / integrated[..., :1], 1e-300))
    return np.maximum(decay, floor_db)


def _decay_slope(decay: np.ndarray, sample_rate: float, upper: float, lower: float) -> float:
    """Seconds per 60 dB from a least-squares line between two levels of the decay."""
    start = int(np.argmax(decay <= upper))
    end = int(np.argmax(decay <= lower))
    if end <= start:
        raise ValueError(
            f"the decay never reaches {lower} dB: the response is too short, or the room is "
            "too live for the run length"
        )
    instants = np.arange(start, end) / sample_rate
    slope, _ = np.polyfit(instants, decay[start:end], 1)
    return -60.0 / slope


def reverberation_time(response, sample_rate: float, evaluation_range: float = 20.0) -> float:
    """T20 (default) or T30: fit from -5 dB down ``evaluation_range`` dB, extrapolated to 60."""
    return _decay_slope(schroeder_decay(response), sample_rate, -5.0, -5.0 - evaluation_range)


def early_decay_time(response, sample_rate: float) -> float:
    """EDT: the first 10 dB of decay, extrapolated to 60."""
    return _decay_slope(schroeder_decay(response), sample_rate, 0.0, -10.0)


def clarity(response, sample_rate: float, split: float = 0.05) -> float:
    """C50 (or C80 with ``split=0.08``): early-to-late energy in dB."""
    boundary = int(round(split * sample_rate))
    energy = np.asarray(response, dtype=float) ** 2
    return 10.0 * np.log10(float(np.sum(energy[..., :boundary])) / float(np.sum(energy[..., boundary:])))


def octave_band_reverberation(response, sample_rate: float, centres=(125.0, 250.0, 500.0, 1000.0),
                              evaluation_range: float = 20.0) -> dict[float, float]:
    """T20 per octave band; bands whose decay is too short are reported as NaN."""
    out = {}
    for centre in centres:
        try:
            out[centre] = reverberation_time(octave_band(response, sample_rate, centre),
                                             sample_rate, evaluation_range)
        except ValueError:
            out[centre] = float("nan")
    return out


def air_absorption_coefficient(attenuation_db_per_metre: float) -> float:
    """dB/m of pressure to the energy coefficient ``m`` (1/m) of Sabine's ``4 m V`` term."""
    return attenuation_db_per_metre / (10.0 * np.log10(np.e))


def sabine_time(volume, surface, absorption, air_coefficient=0.0) -> float:
    return _SABINE * volume / (surface * absorption + 4.0 * air_coefficient * volume)


def eyring_time(volume, surface, absorption, air_coefficient=0.0) -> float:
    return _SABINE * volume / (-surface * np.log(1.0 - absorption) + 4.0 * air_coefficient * volume)


def schroeder_frequency(volume: float, reverberation: float) -> float:
    """``2000 sqrt(T60 / V)``: above it modes overlap and the statistical formulae apply."""
    return 2000.0 * np.sqrt(reverberation / volume)
