"""Writes python_das_beamform.npz beside this script: input/output vectors of
the Python reference's DAS._beamform (python/src/ustb/midprocess/das.py:203-261)
on small seeded inputs in physical units (20.8 MHz, t0 = 1.3 us, float32
time vector as DAS.go builds it, das.py:151), for the four dimension modes,
RF (w0 = 0) and IQ (w0 = 2 pi 5.2 MHz as a float32, das.py:54-55), delays
that reach both trace edges, zero apodization, and a pixel whose receive
weights are all NaN; and (sets 2, IQ, and 3, RF) non-finite values where a
weight is zero: a NaN receive delay and an infinite transmit delay at pixels
whose weight for that channel / wave is zero, a NaN sample that a zero-weight
term reads, and a channel whose weights are all zero holding a NaN delay
(the reference skips only all-zero columns, so the others reach the sum).

Reads the reference where it lies (it is not part of this repository) and is
run by hand, in a container that holds it:
    python tests/golden/make_python_das_vectors.py
"""
import math
import os
import sys

import numpy as np

sys.path.insert(0, "/root/reference/python/src")
from ustb.midprocess.das import DAS  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))


def case(seed, P=128, N=6, W=3, F=2, L=80, fs=20.8e6, t0=1.3e-6):
    rng = np.random.default_rng(seed)
    ch = (rng.standard_normal((L, N, W, F)) + 1j * rng.standard_normal((L, N, W, F))).astype(np.complex64)
    # delays in seconds; sum (index) from about -3 to L + 3 samples
    td = (t0 + rng.uniform(-3.0, L * 0.5, (P, W)) / fs).astype(np.float32)
    rd = (rng.uniform(-3.0, L * 0.5 + 3.0, (P, N)) / fs).astype(np.float32)
    td[:8, :] = (t0 + np.array([0.0, 1.0, L - 2.0, L - 1.0, L, 0.5, L - 1.5, -0.5])[:, None] / fs).astype(np.float32)
    rd[:8, :] = 0.0
    ta = rng.uniform(-0.2, 1.0, (P, W)).astype(np.float32)
    ra = rng.uniform(-0.2, 1.0, (P, N)).astype(np.float32)
    ta[rng.random((P, W)) < 0.3] = 0.0
    ra[rng.random((P, N)) < 0.3] = 0.0
    ra[20, :] = np.nan  # a pixel at depth 0: NaN receive weights
    t = t0 + np.arange(L, dtype=np.float32) / fs
    return ch, t, ta, ra, td, rd


def nonfinite_case(seed):
    """case() with non-finite values where a weight is zero (see the module
    docstring); pixels 30, 31, 32 and channel 4."""
    ch, t, ta, ra, td, rd = case(seed)
    L, fs, t0 = ch.shape[0], 20.8e6, 1.3e-6
    ra[30, 1] = 0.0
    rd[30, 1] = np.nan                     # NaN delay, zero receive weight
    ta[31, 0] = 0.0
    td[31, 0] = np.inf                     # infinite delay, zero transmit weight
    ra[32, 2] = 0.0
    ta[32, 0] = 0.7
    td[32, 0] = np.float32(t0 + 10.25 / fs)
    rd[32, 2] = np.float32(0.0)
    ch[10, 2, 0, 0] = np.nan               # read by pixel 32's zero-weight term (channel 2, wave 0)
    ra[:, 4] = 0.0                         # an all-zero channel: skipped, its NaN delay never read
    rd[5, 4] = np.nan
    return ch, t, ta, ra, td, rd


def main():
    """Input set s (s = 0: RF, w0 = 0; s = 1: IQ; s = 2: IQ and s = 3: RF
    with the non-finite values of nonfinite_case) under keys "s_*"; the
    output of each dimension mode d under "s_out_d"."""
    out = {}
    w_iq = np.float32(2.0 * np.pi * np.float32(5.2e6))
    for s, w0 in enumerate((0.0, w_iq, w_iq, 0.0)):
        ch, t, ta, ra, td, rd = case(seed=100 + s) if s < 2 else nonfinite_case(seed=100 + s)
        L, N, W, F = ch.shape
        P = ta.shape[0]
        for name, arr in (("ch", ch), ("t", t), ("tx_apo", ta), ("rx_apo", ra), ("tx_delay", td),
                          ("rx_delay", rd)):
            out[f"{s}_{name}"] = arr
        out[f"{s}_w0"] = np.float32(w0)
        for dim in (0, 1, 2, 3):
            out[f"{s}_out_{dim}"] = DAS._beamform(ch, t, ta, ra, td, rd, w0, dim, P, N, W, F)
    np.savez_compressed(os.path.join(HERE, "python_das_beamform.npz"), **out)
    print("wrote 4 input sets x 4 modes")


if __name__ == "__main__":
    main()
