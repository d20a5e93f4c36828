"""Impulse responses to audio files: resampling to an audio rate and 32-bit float WAV output.

The grid fixes the sample rate (``1 / dt``, e.g. 17 320.5 Hz at the 1 kHz room grid), so a
response is resampled before it is written. Same public functions and meaning as the
reference's ``src/ac_fdtd/io.py:27-69``: a polyphase filter at a rational ratio (denominator at
most 1000, exact to a part in 10^4) along the last axis, and a float WAV with one channel per
response, peak-normalised by default. The WAV container is written here directly: a RIFF
header with an IEEE-float ``fmt`` chunk (format tag 3) and little-endian float32 frames.
"""

from __future__ import annotations

import struct
from fractions import Fraction
from pathlib import Path

import numpy as np
from scipy import signal as scipy_signal

__all__ = ["resample", "write_wav"]

#: Largest denominator of the rational resampling ratio.
MAX_DENOMINATOR = 1000
_IEEE_FLOAT = 3


def resampling_ratio(source_rate: float, target_rate: float) -> Fraction:
    """``target_rate / source_rate`` as the nearest fraction with a denominator of at most
    :data:`MAX_DENOMINATOR`."""
    if not (source_rate > 0.0 and target_rate > 0.0):
        raise ValueError(f"sample rates must both be positive; {source_rate} Hz and "
                         f"{target_rate} Hz were given")
    return Fraction(target_rate / source_rate).limit_denominator(MAX_DENOMINATOR)


def resample(signal: np.ndarray, source_rate: float, target_rate: float) -> np.ndarray:
    """``signal`` (time along the last axis) from ``source_rate`` to ``target_rate``, through
    scipy's polyphase filter (Kaiser-windowed FIR) at the rational ratio."""
    ratio = resampling_ratio(source_rate, target_rate)
    return scipy_signal.resample_poly(np.asarray(signal, dtype=np.float64), ratio.numerator,
                                      ratio.denominator, axis=-1)


def _wav_bytes(frames: np.ndarray, sample_rate: int) -> bytes:
    """RIFF/WAVE bytes of ``frames`` (samples, channels) as 32-bit IEEE floats."""
    channels = frames.shape[1]
    data = np.ascontiguousarray(frames, dtype="<f4").tobytes()
    block = 4 * channels
    fmt = struct.pack("<HHIIHH", _IEEE_FLOAT, channels, sample_rate, sample_rate * block,
                      block, 32)
    chunks = b"WAVE" + b"fmt " + struct.pack("<I", len(fmt)) + fmt
    chunks += b"data" + struct.pack("<I", len(data)) + data
    return b"RIFF" + struct.pack("<I", len(chunks)) + chunks


def write_wav(path: str | Path, signal: np.ndarray, sample_rate: float,
              target_rate: float | None = 48000.0, normalise: bool = True) -> Path:
    """Write one response, or a stack of them (time along the last axis, one channel each),
    to a 32-bit float WAV at ``target_rate`` (``None`` keeps ``sample_rate``), scaled to a
    peak of 1 unless ``normalise`` is False. Returns the path; missing folders are made."""
    stack = np.atleast_2d(np.asarray(signal, dtype=np.float64))
    rate = float(sample_rate)
    if target_rate is not None:
        stack = resample(stack, rate, target_rate)
        rate = float(target_rate)
    if normalise:
        peak = float(np.max(np.abs(stack))) if stack.size else 0.0
        if peak > 0.0:
            stack = stack / peak
    path = Path(path)
    path.parent.mkdir(parents=True, exist_ok=True)
    path.write_bytes(_wav_bytes(stack.T, int(round(rate))))
    return path
