"""Pins of the oracle's RF -> IQ restatement (CPU): preprocess.fast_demodulation
(fast_demodulation.m:58-183, 267-275) and the Python reference's
FastDemodulation (python/src/ustb/preprocess/fast_demodulation.py:24-85).

- the Kaiser design against scipy.signal.kaiserord / firwin (the same
  published formulas in a library), and its pass-band / stop-band bounds;
- the causal FIR against scipy.signal.lfilter and the zero-phase FIR against
  scipy.signal.filtfilt (its default padding and initial conditions);
- closed forms: a tone at the modulation frequency comes out as the constant
  0.5 exp(i phi) after the filter's transient; output rate and time origin.
"""
import math

import numpy as np
import pytest
from scipy import signal

from oracle import oracle as O


@pytest.mark.parametrize("fs,fmod", [(41.6e6, 5.2e6), (20.8e6, 5.2e6), (50e6, 3e6)])
def test_kaiser_design_matches_scipy(fs, fmod):
    b = O.kaiser_lowpass(0.5 * fmod, fmod, fs)
    numtaps, beta = signal.kaiserord(60.0, 0.5 * fmod / (fs / 2))
    ref = signal.firwin(numtaps, 0.75 * fmod, window=("kaiser", beta), scale=False, fs=fs)
    assert len(b) == numtaps
    np.testing.assert_allclose(b, ref, rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(b, b[::-1], rtol=0, atol=1e-15)   # linear phase
    w, h = signal.freqz(b, 1, worN=[0.0, 0.5 * fmod, fmod, 1.5 * fmod], fs=fs)
    assert abs(abs(h[0]) - 1.0) <= 2e-3 and abs(abs(h[1]) - 1.0) <= 2e-3
    assert abs(h[2]) <= 2e-3 and abs(h[3]) <= 2e-3             # 60 dB ripple spec, up to rounding


def test_fir_restatements_match_scipy():
    rng = np.random.default_rng(3)
    x = rng.standard_normal((2, 3, 700)) + 1j * rng.standard_normal((2, 3, 700))
    b = signal.firwin(31, 0.2)
    np.testing.assert_allclose(O.fir_causal(b, x), signal.lfilter(b, 1, x, axis=-1), rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(O.fir_zero_phase(b, x), signal.filtfilt(b, 1, x, axis=-1), rtol=1e-11, atol=1e-11)
    with pytest.raises(ValueError, match="padlen"):
        O.fir_zero_phase(b, x[..., :93])
    # the impulse response of the causal filter is b
    imp = np.zeros(50)
    imp[0] = 1.0
    np.testing.assert_array_equal(O.fir_causal(b, imp)[:31], b)


def test_demodulated_tone_is_constant():
    fs, fmod, L, phi = 41.6e6, 5.2e6, 800, 0.7
    t0 = 2e-6
    t = t0 + np.arange(L) / fs
    rf = np.cos(2 * np.pi * fmod * t + phi)[None, None, None, :]
    iq, fs_out, t0_out, b = O.fast_demodulation(rf, fs, t0, fmod)
    Ndown = int(fs // (2 * fmod))
    assert fs_out == fs / Ndown and iq.shape[-1] == math.ceil(L / Ndown)
    assert t0_out == pytest.approx(t0 - (len(b) - 1) / 2 / fs, rel=1e-15)
    steady = iq[0, 0, 0, math.ceil(len(b) / Ndown):]
    assert np.max(np.abs(steady - 0.5 * np.exp(1j * phi))) <= 2e-3
    # the Python reference's zero-phase form: unit DC gain, no transient
    iqp, fsp, t0p, bp = O.python_fast_demodulation(rf, fs, t0, fmod)
    assert t0p == t0 and fsp == fs / math.floor(fs / (4 * fmod))
    mid = iqp[0, 0, 0, 20:-20]
    assert np.max(np.abs(mid - 0.5 * np.exp(1j * phi))) <= 2e-3


# ---------------------------------------------------------------------------
# modulation-frequency estimate (fast_demodulation.m:57-73, power_spectrum.m)
# ---------------------------------------------------------------------------
def _pulses(f0, fs=40e6, L=1024, seed=0, sigma=0.4e-6, shape=(2, 3, 8)):
    rng = np.random.default_rng(seed)
    t = np.arange(L) / fs
    tau = rng.uniform(5e-6, 15e-6, shape)[..., None]
    return np.exp(-((t - tau) / sigma) ** 2) * np.cos(2 * np.pi * f0 * (t - tau))


def test_power_spectrum_parseval_and_symmetry():
    """Unnormalised, the spectrum sums to L times the mean trace energy
    (Parseval); normalised its maximum is 1; a real signal's spectrum is even
    about the centre bin after fftshift; fx runs from -fs/2 in steps fs/L."""
    fs, L = 40e6, 1000
    x = _pulses(4.1e6, fs=fs, L=L)
    fx, pw = O.power_spectrum(x, fs, normalised=False)
    np.testing.assert_allclose(pw.sum(), L * np.mean(np.sum(x ** 2, axis=-1)), rtol=1e-12)
    np.testing.assert_allclose(pw[1:L // 2], pw[L // 2 + 1:][::-1], rtol=1e-9, atol=1e-12 * pw.max())
    np.testing.assert_allclose(fx, -fs / 2 + np.arange(L) * fs / L, rtol=0, atol=1e-6)
    _, pn = O.power_spectrum(x, fs)
    assert pn.max() == 1.0


def test_power_spectrum_uses_at_most_100_frames():
    x = np.random.default_rng(1).standard_normal((120, 1, 2, 64))
    _, a = O.power_spectrum(x, 1.0, normalised=False)
    _, b = O.power_spectrum(x[:100], 1.0, normalised=False)
    np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("f0", [5.0e6, 4.3e6, 3.77e6])
def test_estimate_of_a_gaussian_pulse_is_its_centre(f0):
    """A Gaussian-windowed tone's spectrum is symmetric about f0, so the
    centre of its -6 dB points is f0 (to the linear interpolation between
    bins, well under a tenth of a bin)."""
    fs, L = 40e6, 1024
    est = O.estimate_modulation_frequency(_pulses(f0, fs=fs, L=L), fs)
    assert abs(est - f0) < 0.1 * fs / L


def test_interp1_sorts_its_sample_points():
    """MATLAB interp1 sorts distinct sample points: the -6 dB search over a
    non-monotonic stretch of the spectrum interpolates in the sorted order."""
    assert O._interp1_unsorted([0.1, 0.9, 0.4], [10.0, 90.0, 40.0], 0.5) == pytest.approx(50.0)
    assert np.isnan(O._interp1_unsorted([0.1, 0.4], [1.0, 4.0], 0.5))
    with pytest.raises(ValueError):
        O._interp1_unsorted([0.1, 0.1], [1.0, 2.0], 0.1)


def test_estimate_pin_rejects_planted_mistakes():
    """The centre-frequency check fails for a sign error or for taking the
    upper -6 dB point from the positive half."""
    import inspect
    import textwrap

    src = textwrap.dedent(inspect.getsource(O.estimate_modulation_frequency))
    for old, new in (("return -(lo + up) / 2", "return (lo + up) / 2"),
                     ("up = _interp1_unsorted(pw[ic:mid], fx[ic:mid], dc / 2)",
                      "up = _interp1_unsorted(pw[mid:], fx[mid:], dc / 2)")):
        ns = dict(vars(O))
        exec(compile(src.replace(old, new), "<mutant>", "exec"), ns)
        fs, L = 40e6, 1024
        est = ns["estimate_modulation_frequency"](_pulses(4.3e6, fs=fs, L=L), fs)
        assert not abs(est - 4.3e6) < 0.1 * fs / L
