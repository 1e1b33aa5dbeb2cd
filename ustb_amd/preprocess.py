"""preprocess.fast_demodulation on the device.

``FastDemodulation`` mirrors the MATLAB class
(+preprocess/fast_demodulation.m): properties modulation_frequency,
downsample_frequency and lowpass_frequency_vector, ``go()`` returning channel
data whose IQ samples stay on the device. ``flavour = "python"`` follows the
Python reference's FastDemodulation instead
(python/src/ustb/preprocess/fast_demodulation.py). The FIR taps are designed
on the host (scipy.signal, the published Kaiser and windowed-sinc designs);
mixing, filtering and decimation run in ustbx_demodulate. A missing
modulation frequency is estimated from the RF power spectrum
(ustbx_estimate_modulation_frequency, fast_demodulation.m:57-73).
"""
from __future__ import annotations

import math
import warnings

import numpy as np
import torch

from . import _lib as L
from . import ops
from .uff import ChannelData


def matlab_lowpass(f_pass: float, f_stop: float, fs: float, dev: float = 1e-3) -> np.ndarray:
    """kaiserord([f_pass f_stop], [1 0], [dev dev], fs) + fir1(N, Wn, 'low',
    kaiser(N+1, beta), 'noscale') (fast_demodulation.m:267-275)."""
    from scipy.signal import firwin, kaiserord

    numtaps, beta = kaiserord(-20.0 * math.log10(dev), (f_stop - f_pass) / (fs / 2.0))
    return firwin(numtaps, (f_pass + f_stop) / 2.0, window=("kaiser", beta), scale=False, fs=fs)


def python_lowpass(n_samples: int, cutoff: float, fs: float) -> np.ndarray:
    """firwin(numtaps, cutoff, fs=fs), numtaps = min(65, L // 3 * 2 - 1), at
    least 3 and odd (fast_demodulation.py:60-66)."""
    from scipy.signal import firwin

    numtaps = min(65, n_samples // 3 * 2 - 1)
    if numtaps < 3:
        numtaps = 3
    if numtaps % 2 == 0:
        numtaps += 1
    return firwin(numtaps, cutoff, fs=fs)


class FastDemodulation:
    """RF -> IQ: down-mix, low-pass, decimate (fast_demodulation.m:58-183)."""

    def __init__(self):
        self.input: ChannelData | None = None
        self.modulation_frequency: float | None = None
        self.downsample_frequency: float | None = None
        self.lowpass_frequency_vector = (0.5, 1.0)
        self.flavour = "matlab"
        self.output: ChannelData | None = None

    def _rf(self) -> torch.Tensor:
        cd = self.input
        if cd.device_data is not None:
            x = cd.device_data
        else:
            x = torch.from_numpy(np.ascontiguousarray(np.asarray(cd.data).reshape(
                cd.data.shape + (1,) * (4 - cd.data.ndim)).transpose(3, 2, 1, 0))).to("cuda")
        if x.is_complex():
            raise ValueError("fast demodulation takes RF (real) channel data")
        return x.float().contiguous()

    def go(self) -> ChannelData:
        cd = self.input
        fs, t0 = float(cd.sampling_frequency), float(cd.initial_time)
        x = self._rf()
        Ls = x.shape[-1]
        if self.flavour == "python":
            fc = self.modulation_frequency if self.modulation_frequency is not None else cd.modulation_frequency
            if not fc:
                raise ValueError("modulation_frequency (or the channel data's) must be set")
            fc = float(fc)
            b = python_lowpass(Ls, fc * self.lowpass_frequency_vector[0], fs)
            dec = (max(1, int(math.floor(fs / self.downsample_frequency))) if self.downsample_frequency is not None
                   else max(1, int(math.floor(fs / (4 * fc)))))
            iq = ops.demodulate(x, fs, t0, fc, b, dec, L.FIR_ZERO_PHASE)
            fs_out, t0_out, fmod = fs / dec, t0, fc
        else:
            if self.modulation_frequency is None:
                warnings.warn("The modulation frequency is not specified. The estimated central frequency will be "
                              "used.")
                self.modulation_frequency = ops.estimate_modulation_frequency(x, fs)
            fmod = float(self.modulation_frequency)
            fdown = self.downsample_frequency
            if fdown is None:
                warnings.warn("The downsampling frequency is not specified. Using 2 * modulation_frequency.")
                fdown = 2.0 * fmod
            if math.fmod(fs, fdown) != 0:
                warnings.warn("The input sample frequency is not a multiple of the specified downsample frequency. "
                              "Rounding up the downsample frequency to the next higher downsample frequency.")
            Ndown = int(math.floor(fs / fdown))
            self.downsample_frequency = fs / Ndown
            lp = self.lowpass_frequency_vector
            b = matlab_lowpass(lp[0] * fmod, lp[1] * fmod, fs)
            iq = ops.demodulate(x, fs, t0, fmod, b, Ndown, L.FIR_CAUSAL)
            fs_out, t0_out = fs / Ndown, t0 - (len(b) - 1) / 2.0 / fs
        self.b = b
        self.output = ChannelData(probe=cd.probe, sequence=cd.sequence, sampling_frequency=fs_out,
                                  initial_time=t0_out, sound_speed=cd.sound_speed, modulation_frequency=fmod,
                                  device_data=iq)
        return self.output
