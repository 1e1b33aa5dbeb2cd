"""GPU parity of the RF -> IQ front end (ustbx_demodulate) against the
oracle's restatement of preprocess.fast_demodulation (MATLAB: Kaiser FIR,
causal, fast_demodulation.m:58-183) and of the Python reference's
FastDemodulation (zero-phase filtfilt, fast_demodulation.py:24-85).

Seeded RF (noise plus tones), several sampling / modulation / downsampling
frequencies including a downsampling frequency that does not divide the
sampling frequency; tolerance 1e-5 of the peak (FP32 against FP64). The
facade's output beamformed through DAS focuses a simulated point target.
"""
import math

import numpy as np
import pytest
import torch

from oracle import oracle as O
from ustb_amd import _lib as L
from ustb_amd import ops
from ustb_amd.preprocess import FastDemodulation, matlab_lowpass, python_lowpass
from ustb_amd.uff import ChannelData

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-30))


def _rf(seed, shape, fs, fmod):
    rng = np.random.default_rng(seed)
    L = shape[-1]
    t = np.arange(L) / fs
    x = 0.3 * rng.standard_normal(shape) + np.cos(2 * np.pi * fmod * 1.1 * t + rng.uniform(0, 6, shape[:-1] + (1,)))
    return x.astype(np.float32)


@pytest.mark.parametrize("fs,fmod,fdown", [(41.6e6, 5.2e6, None), (20.8e6, 5.2e6, 10.4e6), (50e6, 3e6, 7e6)])
def test_matlab_demodulation_matches_oracle(cuda, fs, fmod, fdown):
    print(f"case demod matlab fs={fs} fmod={fmod} fdown={fdown}", flush=True)
    rf = _rf(1, (2, 3, 4, 777), fs, fmod)
    t0 = 1.5e-6
    ref, fs_o, t0_o, b = O.fast_demodulation(rf, fs, t0, fmod, fdown)
    bd = matlab_lowpass(0.5 * fmod, fmod, fs)
    np.testing.assert_allclose(bd, b, rtol=1e-12, atol=1e-15)  # the product's design equals the restatement
    Nd = int(math.floor(fs / (fdown or 2 * fmod)))
    got = ops.demodulate(torch.from_numpy(rf).to(cuda), fs, t0, fmod, bd, Nd, L.FIR_CAUSAL).cpu().numpy()
    assert got.shape == ref.shape
    assert _rel(got, ref) <= 1e-5


@pytest.mark.parametrize("fs,fc,fdown", [(41.6e6, 5.2e6, None), (20.8e6, 5.2e6, 10.4e6)])
def test_python_demodulation_matches_oracle(cuda, fs, fc, fdown):
    print(f"case demod python fs={fs} fc={fc} fdown={fdown}", flush=True)
    rf = _rf(2, (2, 2, 3, 600), fs, fc)
    ref, fs_o, t0_o, b = O.python_fast_demodulation(rf, fs, 0.0, fc, fdown)
    bd = python_lowpass(600, 0.5 * fc, fs)
    np.testing.assert_array_equal(bd, b)
    dec = int(math.floor(fs / fdown)) if fdown else int(math.floor(fs / (4 * fc)))
    got = ops.demodulate(torch.from_numpy(rf).to(cuda), fs, 0.0, fc, bd, dec, L.FIR_ZERO_PHASE).cpu().numpy()
    assert got.shape == ref.shape
    assert _rel(got, ref) <= 1e-5


def test_mixing_tables_past_the_cache_limit(cuda):
    """The mixing-factor tables (one per device, length, fs, t0 and fmod) are
    kept up to 64 per device, the least recently used released first: 70
    initial times in a row, then the first again, each equal to the oracle."""
    fs, fmod = 41.6e6, 5.2e6
    rf = _rf(3, (1, 1, 2, 300), fs, fmod)
    bd = matlab_lowpass(0.5 * fmod, fmod, fs)
    Nd = int(math.floor(fs / (2 * fmod)))
    x = torch.from_numpy(rf).to(cuda)
    for k in list(range(70)) + [0, 1]:
        t0 = 1e-6 + k * 1e-8
        ref = O.fast_demodulation(rf, fs, t0, fmod, None)[0]
        got = ops.demodulate(x, fs, t0, fmod, bd, Nd, L.FIR_CAUSAL).cpu().numpy()
        assert _rel(got, ref) <= 1e-5, k


def test_zero_phase_needs_more_samples_than_padlen(cuda):
    rf = torch.zeros((1, 1, 1, 60), device=cuda)
    with pytest.raises(L.USTBXError, match="padlen"):
        ops.demodulate(rf, 40e6, 0.0, 5e6, np.ones(21) / 21, 1, L.FIR_ZERO_PHASE)


def test_fast_demodulation_facade_then_das_focuses(cuda):
    """RF of a point target -> FastDemodulation -> DAS (IQ with the reference
    phase): the image peaks within 2 mm of the target, as the RF path does."""
    from ustb_amd import synth
    from ustb_amd.das import DAS
    from ustb_amd.enums import Window
    from ustb_amd.uff import Apodization, ChannelData, LinearArray, LinearScan, Wave

    prb = LinearArray(64, 0.3e-3)
    seq = [Wave.plane(a) for a in np.linspace(-0.2, 0.2, 3)]
    fs, f0 = 41.6e6, 5.2e6
    rf = synth.echoes(prb, seq, np.array([(1e-3, 0.0, 20e-3)]), np.ones(1), fs=fs, n_samples=2400, f0=f0,
                      bandwidth=0.6, iq=False)
    cd = ChannelData(probe=prb, sequence=seq, sampling_frequency=fs, modulation_frequency=0.0,
                     data=np.ascontiguousarray(rf.transpose(3, 2, 1, 0)))
    dm = FastDemodulation()
    dm.input = cd
    dm.modulation_frequency = f0
    dm.downsample_frequency = 2 * f0
    iq = dm.go()
    assert iq.sampling_frequency == fs / 4 and iq.modulation_frequency == f0
    assert iq.initial_time == pytest.approx(-(len(dm.b) - 1) / 2 / fs)
    assert tuple(iq.device_data.shape) == (1, 3, 64, 600)
    d = DAS()
    d.channel_data = iq
    d.scan = LinearScan(np.linspace(-5e-3, 5e-3, 41), np.linspace(15e-3, 25e-3, 81))
    d.receive_apodization = Apodization(Window.hanning, f_number=1.7)
    img = d.go().data.cpu().numpy()[:, 0, 0, 0]
    k = int(np.argmax(np.abs(img)))
    assert abs(d.scan.x[k] - 1e-3) <= 2e-3 and abs(d.scan.z[k] - 20e-3) <= 2e-3


def _pulses(f0, fs, L, seed, shape=(3, 4, 8), sigma=0.4e-6):
    rng = np.random.default_rng(seed)
    t = np.arange(L) / fs
    tau = rng.uniform(0.2 * L / fs, 0.6 * L / fs, shape)[..., None]
    x = np.exp(-((t - tau) / sigma) ** 2) * np.cos(2 * np.pi * f0 * (t - tau))
    return (x + 0.01 * rng.standard_normal(x.shape)).astype(np.float32)


@pytest.mark.parametrize("f0,fs,L", [(5.2e6, 41.6e6, 1024), (3.1e6, 25e6, 1000), (7.0e6, 50e6, 777)])
def test_modulation_frequency_estimate_matches_oracle(cuda, f0, fs, L):
    """ustbx_estimate_modulation_frequency against the oracle's restatement
    (fast_demodulation.m:57-73, power_spectrum.m): the normalised power
    spectrum within 1e-5 (FP32 FFT against FP64), the estimate within a
    hundredth of a bin, and within half a bin of the pulses' centre
    frequency (the linear interpolation of the -6 dB points between bins of
    a Gaussian spectrum; odd and even L)."""
    print(f"case f0 estimate f0={f0} fs={fs} L={L}", flush=True)
    rf = _pulses(f0, fs, L, seed=int(f0) % 97)
    est, fx, pw = ops.estimate_modulation_frequency(torch.from_numpy(rf).to(cuda), fs, return_spectrum=True)
    ofx, opw = O.power_spectrum(rf.astype(np.float64), fs)
    np.testing.assert_allclose(pw, opw, atol=1e-5)
    np.testing.assert_allclose(fx, ofx, rtol=0, atol=1e-6 * fs)
    ref = O.estimate_modulation_frequency(rf.astype(np.float64), fs)
    assert abs(est - ref) < 0.01 * fs / L, (est, ref)
    assert abs(est - f0) < 0.5 * fs / L, (est, f0)


def test_fast_demodulation_estimates_a_missing_modulation_frequency(cuda):
    """FastDemodulation without a modulation frequency takes the estimate
    (warning as the MATLAB class does) and then demodulates exactly as with
    that frequency set."""
    fs, f0 = 41.6e6, 5.2e6
    rf = _pulses(f0, fs, 1024, seed=3)
    cd = ChannelData(probe=None, sequence=[None] * rf.shape[1], sampling_frequency=fs,
                     device_data=torch.from_numpy(rf).to(cuda))
    a = FastDemodulation()
    a.input = cd
    a.downsample_frequency = 10.4e6
    with pytest.warns(UserWarning, match="modulation frequency is not specified"):
        out_a = a.go()
    est = a.modulation_frequency
    assert abs(est - f0) < 0.5 * fs / 1024
    b = FastDemodulation()
    b.input = cd
    b.downsample_frequency = 10.4e6
    b.modulation_frequency = est
    out_b = b.go()
    assert torch.equal(out_a.device_data, out_b.device_data)
    assert out_a.modulation_frequency == est
