"""Seeded synthetic channel data (point scatterers and speckle).

Echo model after the reference's point-scatterer simulator (@fresnel/fresnel.m:87-137):
the echo of scatterer s on element n for transmit w arrives at
tau = tx_delay(s, w) + |x_s - x_n| / c, with the transmit delay of das.m's
wavefront conventions (+midprocess/das.m:110-206). The pulse is uff.pulse's
Gaussian-windowed cosine cos(2 pi f0 t) exp(-1.7886 (t bw f0)^2)
(+uff/pulse.m:76-79).

- RF: sum_s a_s * cos(2 pi f0 (t - tau)) * env(t - tau)
- IQ (demodulated at f0): sum_s a_s * env(t - tau) * exp(-1j 2 pi f0 tau),
  the down-mix sign of +preprocess/fast_demodulation.m:134.

Only samples within +-3.5 envelope widths of tau are written (the envelope
is below 1e-9 beyond); the result is deterministic (each trace is
accumulated by one thread in scatterer order).
Output layout: (F, W, N, L) complex64 or float32 (the MEX memory layout).
"""
from __future__ import annotations

import math
import os

import numpy as np

from .enums import Wavefront
from .uff import Probe, Wave


def pulse_bandwidth_for_cycles(cycles: float) -> float:
    """Fractional bandwidth whose envelope FWHM is `cycles` periods:
    exp(-1.7886 (t bw f0)^2) = 1/2 at t = cycles/(2 f0)."""
    return 2.0 * math.sqrt(math.log(2.0) / 1.7886) / cycles


def tx_delay_points(wave: Wave, x, y, z, model_hybrid_margin=1e-3):
    """Transmit delay [s] of scatterers (das.m:110-206 conventions, spherical
    model for focused sources)."""
    src = wave.source
    if wave.wavefront == Wavefront.plane or math.isinf(src.distance):
        ca, sa = math.cos(src.azimuth), math.sin(src.azimuth)
        ce, se = math.cos(src.elevation), math.sin(src.elevation)
        d = z * ca * ce + x * sa * ce + y * se
    elif wave.wavefront == Wavefront.photoacoustic:
        d = np.zeros_like(x)
    else:
        sx, sy, sz = src.xyz
        r = np.sqrt((sx - x) ** 2 + (sy - y) ** 2 + (sz - z) ** 2)
        sgn = np.where(z < sz, -1.0, 1.0)
        d = sgn * r - abs(src.distance) if sz < 0 else sgn * r + src.distance
    return d / wave.sound_speed - wave.delay


def echoes(probe: Probe, sequence: list, scatterers: np.ndarray, amplitudes: np.ndarray, *, fs: float,
           n_samples: int, f0: float, bandwidth: float, iq: bool, t0: float = 0.0, c: float = 1540.0,
           n_frames: int = 1, frame_offsets: np.ndarray | None = None, threads: int | None = None,
           support: float = 3.5):
    """scatterers: (S, 3) [m]; amplitudes: (S,) or (F, S); frame_offsets:
    (F, 3) displacement of all scatterers per frame. Returns (F, W, N, L)
    complex64 (iq) or float32 (RF), computed by the native generator."""
    import ctypes

    from . import _lib as L

    S = scatterers.shape[0]
    F, W, N, Ls = int(n_frames), len(sequence), probe.N_elements, int(n_samples)
    pos = np.empty((F, S, 3), dtype=np.float64)
    txd = np.empty((F, W, S), dtype=np.float64)
    for f in range(F):
        pos[f] = scatterers + (frame_offsets[f] if frame_offsets is not None else 0.0)
        for w, wave in enumerate(sequence):
            txd[f, w] = tx_delay_points(wave, pos[f, :, 0], pos[f, :, 1], pos[f, :, 2])
    amps = np.ascontiguousarray(np.broadcast_to(np.asarray(amplitudes, dtype=np.float64), (F, S)))
    out = np.zeros((F, W, N, Ls), dtype=np.complex64 if iq else np.float32)
    d = L.SynthDesc(L=Ls, N=N, W=W, F=F, S=S, iq=1 if iq else 0,
                    half=int(math.ceil(support / (bandwidth * f0) * fs)) + 1,
                    threads=int(threads or min(16, os.cpu_count() or 1)), fs=fs, t0=t0, f0=f0,
                    bandwidth=bandwidth, c=c)
    ex, ey, ez = (np.ascontiguousarray(a, dtype=np.float64) for a in (probe.x, probe.y, probe.z))
    rc = L.lib().ustbx_synth_echoes(ctypes.byref(d), ex.ctypes.data, ey.ctypes.data, ez.ctypes.data,
                                    pos.ctypes.data, amps.ctypes.data, txd.ctypes.data, out.ctypes.data)
    if rc != 0:
        raise ValueError("ustbx_synth_echoes: bad arguments")
    return out


def analytic(rf: np.ndarray) -> np.ndarray:
    """Analytic signal along time (the RF branch of das.m:211-214): FFT,
    zero negative frequencies, double positive ones (MATLAB hilbert)."""
    L = rf.shape[-1]
    X = np.fft.fft(rf.astype(np.float64), axis=-1)
    h = np.zeros(L)
    h[0] = 1.0
    if L % 2 == 0:
        h[L // 2] = 1.0
        h[1:L // 2] = 2.0
    else:
        h[1:(L + 1) // 2] = 2.0
    return np.fft.ifft(X * h, axis=-1).astype(np.complex64)
