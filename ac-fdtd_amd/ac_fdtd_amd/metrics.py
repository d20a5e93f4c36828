"""Room-acoustics parameters of the recordings the GPU time loop produces.

The public functions keep the names, arguments and definitions of the reference's metrics
(``src/ac_fdtd/metrics.py:41-151``): a zero-phase 4th-order Butterworth octave band over
``[f/sqrt2, f sqrt2]``, Schroeder's backward energy integral run to the end of the response
(a simulation has no noise floor to truncate), reverberation times from a least-squares line
fitted between two levels of that decay (T20/T30 from -5 dB, EDT from 0 dB) extrapolated to
60 dB, clarity as an early-to-late energy ratio, and the Sabine and Eyring estimates with the
constant ``24 ln 10 / c`` at c = 343 m/s.

The work is organised around :class:`DecayCurve`, which integrates a response once and then
answers every decay-time question about it; the line fit is the closed-form least-squares
slope over centred sample times. Everything here is host post-processing in fp64.
"""

from __future__ import annotations

import math

import numpy as np
from scipy import signal as scipy_signal

__all__ = [
    "DecayCurve",
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

#: Sound speed the statistical formulae are quoted at, m/s.
_FORMULA_SOUND_SPEED = 343.0
#: 60 dB of energy decay is ln(10^6) nepers; Sabine's constant is 4 x that / c = 24 ln 10 / c.
_SABINE_CONSTANT = 4.0 * math.log(1e6) / _FORMULA_SOUND_SPEED
_BAND_ORDER = 4
#: The upper band edge is held just below Nyquist, where the filter design still converges.
_NYQUIST_MARGIN = 0.999


def _band_edges(centre: float, sample_rate: float) -> tuple[float, float]:
    """Octave-band edges in Hz, the upper one kept below Nyquist."""
    half_octave = math.sqrt(2.0)
    nyquist = sample_rate / 2.0
    return centre / half_octave, min(centre * half_octave, _NYQUIST_MARGIN * nyquist)


def octave_band(response: np.ndarray, sample_rate: float, centre: float) -> np.ndarray:
    """One octave band of ``response`` (time along the last axis), filtered forward and back so
    the band's energy envelope keeps its timing."""
    low, high = _band_edges(centre, sample_rate)
    if not low < high:
        raise ValueError(
            f"an octave band centred at {centre} Hz needs a sample rate above "
            f"{2.0 * low / _NYQUIST_MARGIN:.1f} Hz; this response is sampled at {sample_rate} Hz"
        )
    sos = scipy_signal.butter(_BAND_ORDER, (low, high), btype="bandpass", fs=sample_rate,
                              output="sos")
    return scipy_signal.sosfiltfilt(sos, response, axis=-1)


class DecayCurve:
    """Schroeder's energy decay of a response: the energy still to arrive after each sample,
    in dB relative to the total, clipped below at ``floor_db``.

    ``level`` holds the curve; :meth:`decay_time` fits a line to any stretch of it.
    """

    def __init__(self, response, sample_rate: float = 1.0, floor_db: float = -120.0) -> None:
        power = np.square(np.asarray(response, dtype=np.float64))
        remaining = np.flip(np.cumsum(np.flip(power, axis=-1), axis=-1), axis=-1)
        total = remaining[..., :1]
        ratio = np.divide(remaining, total, out=np.zeros_like(remaining), where=total > 0.0)
        with np.errstate(divide="ignore"):
            level = 10.0 * np.log10(ratio)
        self.level = np.maximum(level, floor_db)
        self.sample_rate = float(sample_rate)

    def _first_at_or_below(self, threshold_db: float) -> int | None:
        hits = np.flatnonzero(self.level <= threshold_db)
        return int(hits[0]) if hits.size else None

    def decay_time(self, start_db: float, stop_db: float) -> float:
        """Seconds for 60 dB of decay, from the least-squares line through the curve between
        the first samples at or below ``start_db`` and ``stop_db``."""
        first = self._first_at_or_below(start_db)
        last = self._first_at_or_below(stop_db)
        if first is None or last is None or last <= first:
            raise ValueError(
                f"the energy decay does not fall from {start_db} dB to {stop_db} dB within the "
                "response; run the simulation longer or add absorption"
            )
        times = np.arange(first, last, dtype=np.float64) / self.sample_rate
        levels = self.level[first:last]
        centred = times - times.mean()
        slope = float(np.dot(centred, levels - levels.mean()) / np.dot(centred, centred))
        return -60.0 / slope

    def t(self, evaluation_range: float = 20.0) -> float:
        """T20 (or T30 with ``evaluation_range=30``): the fit starts 5 dB down."""
        return self.decay_time(-5.0, -5.0 - evaluation_range)

    def edt(self) -> float:
        """Early decay time: the fit over the first 10 dB."""
        return self.decay_time(0.0, -10.0)


def schroeder_decay(response: np.ndarray, floor_db: float = -120.0) -> np.ndarray:
    """The decay curve in dB (0 at the start, non-increasing, clipped at ``floor_db``)."""
    return DecayCurve(response, floor_db=floor_db).level


def reverberation_time(response, sample_rate: float, evaluation_range: float = 20.0) -> float:
    """T20 by default, T30 with ``evaluation_range=30``; seconds."""
    return DecayCurve(response, sample_rate).t(evaluation_range)


def early_decay_time(response, sample_rate: float) -> float:
    """EDT in seconds."""
    return DecayCurve(response, sample_rate).edt()


def clarity(response, sample_rate: float, split: float = 0.05) -> float:
    """C50 (C80 with ``split=0.08``): energy before ``split`` seconds over energy after, dB."""
    power = np.square(np.asarray(response, dtype=np.float64))
    boundary = int(round(split * sample_rate))
    early, late = float(power[..., :boundary].sum()), float(power[..., boundary:].sum())
    return 10.0 * math.log10(early / late)


def octave_band_reverberation(response, sample_rate: float, centres=(125.0, 250.0, 500.0, 1000.0),
                              evaluation_range: float = 20.0) -> dict[float, float]:
    """T20 (or T30) of each octave band; NaN for a band whose decay is too short to fit."""
    times = {}
    for centre in centres:
        try:
            band = octave_band(response, sample_rate, centre)
            times[centre] = DecayCurve(band, sample_rate).t(evaluation_range)
        except ValueError:
            times[centre] = float("nan")
    return times


def air_absorption_coefficient(attenuation_db_per_metre: float) -> float:
    """ISO 9613-1 attenuation (dB of pressure per metre) as the energy coefficient ``m`` (1/m)
    of the ``4 m V`` term: a pressure level falls by 10 log10(e) dB per neper of energy."""
    return attenuation_db_per_metre * math.log(10.0) / 10.0


def _statistical_time(volume, loss_area) -> float:
    return _SABINE_CONSTANT * volume / loss_area


def sabine_time(volume, surface, absorption, air_coefficient=0.0) -> float:
    """Sabine's T60 in seconds: equivalent absorption area ``S alpha`` plus ``4 m V``."""
    return _statistical_time(volume, surface * absorption + 4.0 * air_coefficient * volume)


def eyring_time(volume, surface, absorption, air_coefficient=0.0) -> float:
    """Eyring's T60 in seconds: ``-S ln(1 - alpha)`` in place of ``S alpha``."""
    return _statistical_time(volume, -surface * np.log1p(-absorption)
                             + 4.0 * air_coefficient * volume)


def schroeder_frequency(volume: float, reverberation: float) -> float:
    """Hz above which modes overlap enough for the statistical formulae to hold:
    ``2000 sqrt(T60 / V)``."""
    return 2000.0 * math.sqrt(reverberation / volume)
