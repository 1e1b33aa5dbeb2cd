"""The Python reference's DAS._beamform (python das.py:203-261) on committed
golden vectors (tests/golden/python_das_beamform.npz, written by
tests/golden/make_python_das_vectors.py from the reference itself): physical
units (20.8 MHz, t0 = 1.3 us, float32 time vector), RF and IQ, all four
dimension modes, both trace edges, zero weights and a pixel with NaN receive
weights; and (sets 2 and 3) non-finite delays and samples where a weight is
zero, which reach the sum because the reference skips only all-zero weight
columns. Pins the oracle's restatement bit for bit (CPU); the device path
through the facade is checked against the same vectors in
test_gpu_facade.py."""
import os

import numpy as np
import pytest

from oracle import oracle as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "python_das_beamform.npz")
SETS = (0, 1, 2, 3)


def golden_cases():
    z = np.load(GOLDEN)
    for s in SETS:
        args = tuple(z[f"{s}_{k}"] for k in ("ch", "t", "tx_apo", "rx_apo", "tx_delay", "rx_delay"))
        for dim in (0, 1, 2, 3):
            yield s, dim, args, z[f"{s}_w0"], z[f"{s}_out_{dim}"]


@pytest.mark.parametrize("s,dim", [(s, d) for s in SETS for d in (0, 1, 2, 3)])
def test_oracle_restatement_equals_reference_vectors(s, dim):
    for s_, dim_, args, w0, ref in golden_cases():
        if (s_, dim_) != (s, dim):
            continue
        ch, t, ta, ra, td, rd = args
        got = O.python_beamform(ch, t, ta, ra, td, rd, w0, dim)
        assert got.dtype == ref.dtype and got.shape == ref.shape
        assert np.array_equal(got, ref, equal_nan=True)
        # the NaN receive weights of pixel 20 reach every output of that
        # pixel (sets 2, 3: but channel 4's, whose weights are all zero)
        assert np.isnan(ref[20]).all() if s < 2 else np.isnan(ref[20]).any()
        if s < 2:
            assert not np.isnan(np.delete(ref, 20, axis=0)).any()
        else:  # a NaN sample read with zero weight (pixel 32); in IQ also the
            # zero-weight NaN and infinite delays (pixels 30, 31) via the phase
            nan_pixels = set(np.nonzero(np.isnan(ref).reshape(ref.shape[0], -1).any(axis=1))[0].tolist())
            assert {20, 32} <= nan_pixels
            assert ({30, 31} <= nan_pixels) == (s == 2)


def test_golden_inputs_cover_the_window_edges():
    """Indices (delay - t0) / dt of the first pixels hit -0.5, 0, 0.5, 1, L-2,
    L-1.5, L-1 and L samples, so both edge rules are exercised."""
    z = np.load(GOLDEN)
    t = z["0_t"]
    dt = np.float32(float(t[1] - t[0]))
    idx = ((z["0_tx_delay"][:8, 0] + z["0_rx_delay"][:8, 0]) - np.float32(t[0])) / dt
    L = t.size
    np.testing.assert_allclose(np.sort(idx), [-0.5, 0.0, 0.5, 1.0, L - 2, L - 1.5, L - 1, L], atol=2e-3)
