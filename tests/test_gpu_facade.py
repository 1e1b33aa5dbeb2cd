"""The midprocess.das facade (ustb_amd.das.DAS) and the device Hilbert step.

- hilbert: the analytic signal of das.m:211-214 / das.py:144-145 against
  scipy.signal.hilbert (the oracle's), odd and even lengths;
- DAS.go(): IQ data in the reference layout [time, channel, wave, frame],
  delays and apodization on the device, the reference phase of das.m:306-309,
  output [pixel, channel, wave, frame], against the FP64 oracle on the
  oracle's geometry (das_c's wd and FP32 index, as in test_gpu_geom.py);
  RF data through the Hilbert step; the plan cache;
- DAS._beamform: the Python reference's static contract (das.py:203-261)
  against the oracle's restatement of it.
"""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from tests.cases import channel_noise, geom_cases, matrix_problem, oracle_matrices
from ustb_amd.das import DAS
from ustb_amd.enums import Dimension
from ustb_amd.ops import hilbert
from ustb_amd.uff import ChannelData

pytestmark = pytest.mark.gpu
CASES = geom_cases()
TOL = 1e-5


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-30))


@pytest.mark.parametrize("L", [777, 1001, 512, 1024, 2048, 4096, 600, 1000, 1536, 2100, 3500])
def test_hilbert_matches_scipy(cuda, L):
    """hipFFT path (777 = 3 7 37, 1001 = 7 11 13), the one-block LDS FFT path
    (powers of two 512 .. 4096) and its mixed-radix form (600, 1000, 1536,
    2100, 3500: products of 2, 3, 5, 7); 21 traces, so the last block's
    pair has one trace."""
    print(f"case hilbert L={L}", flush=True)
    rng = np.random.default_rng(L)
    x = rng.standard_normal((1, 3, 7, L)).astype(np.float32)
    got = hilbert(torch.from_numpy(x).to(cuda)).cpu().numpy()
    ref = O.hilbert(x.astype(np.float64), axis=-1)
    # the real part comes back through the FFT round trip, as in MATLAB's and
    # scipy's hilbert (ifft(fft(x) .* h)): equal to x up to FP32 rounding
    assert _rel(got.real, x) <= TOL
    assert _rel(got, ref) <= TOL
    # a pure tone at bin k: the analytic signal is exp(i 2 pi k n / L) exactly
    n = np.arange(L)
    k = 37
    tone = np.cos(2 * np.pi * k * n / L).astype(np.float32)[None]
    z = hilbert(torch.from_numpy(np.repeat(tone, 3, 0)).to(cuda)).cpu().numpy()
    assert np.max(np.abs(z - np.exp(2j * np.pi * k * n / L)[None])) <= 2e-5


def test_hilbert_plan_cache_past_its_limit(cuda):
    """The hipFFT path keeps at most 8 plans per device (one per length,
    trace count and stream) and releases the least recently used one: 11
    trace counts in a row, then the first again on a second stream, all
    match scipy."""
    L = 777
    rng = np.random.default_rng(5)
    s2 = torch.cuda.Stream(device=cuda)
    for k, T in enumerate(list(range(1, 12)) + [1, 2]):
        print(f"case hilbert cache T={T}", flush=True)
        x = rng.standard_normal((T, L)).astype(np.float32)
        xd = torch.from_numpy(x).to(cuda)
        if k >= 11:
            s2.wait_stream(torch.cuda.current_stream(cuda))
            with torch.cuda.stream(s2):
                got = hilbert(xd)
            torch.cuda.current_stream(cuda).wait_stream(s2)
        else:
            got = hilbert(xd)
        assert _rel(got.cpu().numpy(), O.hilbert(x.astype(np.float64), axis=-1)) <= TOL


def _das(case, data, dim=Dimension.both):
    d = DAS()
    d.channel_data = ChannelData(probe=case["probe"], sequence=case["seq"], sampling_frequency=case["fs"],
                                 modulation_frequency=case["fd"], data=data)
    d.scan = case["scan"]
    d.receive_apodization = case["rx"]
    d.transmit_apodization = case["tx"]
    d.dimension = dim
    d.spherical_transmit_delay_model = case["model"]
    return d


@pytest.mark.parametrize("name,dim", [("cpwc_tukey", Dimension.both), ("dw_sector", Dimension.both),
                                      ("cpwc_tukey", Dimension.transmit)])
def test_go_iq_matches_oracle(cuda, name, dim):
    case = CASES[name]
    ch = channel_noise(case, F=2, seed=5)                     # (F, W, N, L)
    d = _das(case, np.ascontiguousarray(ch.transpose(3, 2, 1, 0)), dim)
    got = d.go().data.cpu().numpy()                            # (P, N', W', F)
    ta, ra, td, rd = oracle_matrices(case)
    ref = O.beamform_fp64(ch, case["fs"], 0.0, ta, ra, td, rd, case["fd"], int(dim),
                          w0=O.dasc_wd(case["fd"]), index="fp32")
    ref = ref * O.reference_phase(case["scan"].reference_distance, case["fd"], 1540.0)[None, None, None, :]
    assert got.shape == ref.transpose(3, 2, 1, 0).shape
    assert _rel(got, ref.transpose(3, 2, 1, 0)) <= TOL
    # a second go() reuses the cached plan and gives the same image
    again = d.go().data.cpu().numpy()
    assert len(d._plans) == 1 and np.array_equal(again, got)
    assert d.elapsed_time is not None and d.elapsed_time > 0


def test_go_keeps_the_last_plans_only(cuda):
    """go() over MAX_PLANS + 2 scans keeps the last MAX_PLANS plans (the
    oldest are released) and re-running an evicted configuration builds it
    again with the same image."""
    from ustb_amd.uff import LinearScan

    case = CASES["cpwc_tukey"]
    ch = channel_noise(case, F=1, seed=9)
    d = _das(case, np.ascontiguousarray(ch.transpose(3, 2, 1, 0)))
    imgs = []
    for k in range(DAS.MAX_PLANS + 2):
        d.scan = LinearScan(np.linspace(-3e-3, 3e-3, 16), np.linspace(8e-3 + 1e-4 * k, 14e-3, 24))
        imgs.append(d.go().data.cpu().numpy())
        assert len(d._plans) == min(k + 1, DAS.MAX_PLANS)
    d.scan = LinearScan(np.linspace(-3e-3, 3e-3, 16), np.linspace(8e-3, 14e-3, 24))  # evicted first
    assert np.array_equal(d.go().data.cpu().numpy(), imgs[0])
    assert len(d._plans) == DAS.MAX_PLANS


def test_go_rf_goes_through_hilbert(cuda):
    case = CASES["pw_rf_hanning"]
    rng = np.random.default_rng(17)
    F, W, N, L = 1, len(case["seq"]), case["probe"].N_elements, case["L"]
    rf = rng.standard_normal((F, W, N, L)).astype(np.float32)
    d = _das(case, np.ascontiguousarray(rf.transpose(3, 2, 1, 0)))
    got = d.go().data.cpu().numpy()
    analytic = O.hilbert(rf.astype(np.float64), axis=-1).astype(np.complex64)
    ta, ra, td, rd = oracle_matrices(case)
    ref = O.beamform_fp64(analytic, case["fs"], 0.0, ta, ra, td, rd, 0.0, 3, index="fp32")
    assert _rel(got, ref.transpose(3, 2, 1, 0)) <= TOL


def test_go_complex_rf_device_data_goes_through_hilbert(cuda):
    """Complex RF with modulation frequency 0, handed over as a device
    tensor, is made analytic from its real part as host data is (das.m:212:
    hilbert whenever |w0| < eps)."""
    case = CASES["pw_rf_hanning"]
    rng = np.random.default_rng(18)
    F, W, N, L = 1, len(case["seq"]), case["probe"].N_elements, case["L"]
    rf = rng.standard_normal((F, W, N, L)).astype(np.float32)
    host = _das(case, np.ascontiguousarray(rf.transpose(3, 2, 1, 0))).go().data.cpu().numpy()
    d = _das(case, None)
    d.channel_data.device_data = torch.from_numpy(rf.astype(np.complex64)).to(cuda)
    got = d.go().data.cpu().numpy()
    np.testing.assert_array_equal(got, host)


@pytest.mark.parametrize("dim", [0, 1, 2, 3])
def test_beamform_static_matches_python_reference(cuda, dim):
    pb = matrix_problem(seed=21 + dim, P=300, F=2)
    ch = np.ascontiguousarray(pb["ch"].transpose(3, 2, 1, 0))  # (L, N, W, F)
    t = np.arange(ch.shape[0], dtype=np.float32)
    w0 = np.float32(2.0 * np.pi * np.float32(0.05))
    args = (ch, t, pb["tx_apo"].T, pb["rx_apo"].T, pb["tx_delay"].T, pb["rx_delay"].T, w0, dim)
    got = DAS._beamform(*args, N_pixels=300, N_channels=ch.shape[1], N_waves=ch.shape[2], N_frames=ch.shape[3])
    ref = O.python_beamform(*args)
    assert got.shape == ref.shape
    assert _rel(got, ref) <= TOL


@pytest.mark.parametrize("s,dim", [(s, d) for s in (0, 1, 2, 3) for d in (0, 1, 2, 3)])
def test_beamform_static_matches_reference_vectors(cuda, s, dim):
    """DAS._beamform on the Python reference's own output (golden vectors in
    physical units; USTBX_FLAG_PYTHON_ARITHMETIC): RF bit for bit, IQ within
    2e-6 of the peak (its cos/sin and complex product are libm's and NumPy's
    on the host), NaN positions identical."""
    from tests.test_golden_python import golden_cases

    for s_, dim_, args, w0, ref in golden_cases():
        if (s_, dim_) != (s, dim):
            continue
        print(f"case python golden set={s} dim={dim}", flush=True)
        ch, t, ta, ra, td, rd = args
        L_, N, W, F = ch.shape
        got = DAS._beamform(ch, t, ta, ra, td, rd, w0, dim, N_pixels=ta.shape[0], N_channels=N, N_waves=W,
                            N_frames=F)
        assert got.shape == ref.shape and got.dtype == ref.dtype
        assert np.array_equal(np.isnan(got), np.isnan(ref))
        if s in (0, 3):  # RF
            assert np.array_equal(got, ref, equal_nan=True)
        else:
            ok = ~np.isnan(ref)
            assert np.max(np.abs(got[ok] - ref[ok])) <= 2e-6 * np.max(np.abs(ref[ok]))
