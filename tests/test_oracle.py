"""Pins of the oracle (CPU). Each check is fixed by something outside the
oracle's own formulas: a closed form, an invariant of the sum, a library
routine (numpy.interp), or brute force in FP64."""
import math

import numpy as np
import pytest

from oracle import oracle as O
from tests.cases import matrix_problem


def _one(L=32, td=0.0, rd=0.0, ta=1.0, ra=1.0, x=None, fd=0.0, dim=3):
    ch = np.zeros((1, 1, 1, L), np.complex64) if x is None else np.asarray(x, np.complex64).reshape(1, 1, 1, L)
    f = lambda v: np.array([[v]], np.float32)
    return O.das_c(ch, 1.0, 0.0, f(ta), f(ra), f(td), f(rd), fd, dim)[0, 0, 0, 0]


def test_impulse_closed_form():
    # d = 3.25 + 7.0 = 10.25 exactly: a = 0.75, b = 0.25
    x = np.zeros(32, np.complex64)
    x[10], x[11] = 2.0, 6.0
    assert _one(td=3.25, rd=7.0, x=x) == pytest.approx(0.75 * 2 + 0.25 * 6, abs=0)
    assert _one(td=3.25, rd=7.0, ta=0.5, ra=4.0, x=x) == 2.0 * 3.0


@pytest.mark.parametrize("d,expect", [(0.5, 0.0), (0.999, 0.0), (1.0, 1.0), (30.0, 30.0), (30.999, 30.999),
                                      (31.0, 0.0), (-0.5, 0.0), (-3.0, 0.0), (1e9, 0.0)])
def test_edge_rule_dasc(d, expect):
    """das_c accepts 0 < trunc(d) < L-1 (das_c.cpp:377,381); a ramp x[l] = l
    interpolates to d inside the window."""
    x = np.arange(32, dtype=np.complex64)
    got = _one(td=np.float32(d), x=x)
    assert got == pytest.approx(expect, rel=1e-6, abs=1e-6)


def test_linear_ramp_is_exact():
    rng = np.random.default_rng(0)
    L = 64
    alpha, beta = 0.7 - 0.2j, 1.5 + 0.3j
    x = (alpha * np.arange(L) + beta).astype(np.complex64)
    for _ in range(50):
        d = np.float32(rng.uniform(1, L - 1.01))
        assert abs(_one(L=L, td=d, x=x) - (alpha * float(d) + beta)) < 1e-4


def test_phase_rotation_closed_form():
    """Constant data: out = apo * exp(i * wd * tau), wd = 2*(float)pi*fd."""
    rng = np.random.default_rng(1)
    x = np.ones(64, np.complex64)
    for _ in range(20):
        td, rd = np.float32(rng.uniform(2, 30)), np.float32(rng.uniform(2, 30))
        fd = 0.05
        got = _one(L=64, td=td, rd=rd, ta=0.5, ra=0.8, x=x, fd=fd)
        ref = 0.4 * np.exp(1j * O.dasc_wd(fd) * (float(td) + float(rd)))
        assert abs(got - ref) < 2e-6


def test_fp64_linear_matches_numpy_interp():
    rng = np.random.default_rng(2)
    L, P = 50, 300
    tr = (rng.standard_normal(L) + 1j * rng.standard_normal(L))
    d = rng.uniform(0, L - 1, P)
    v, valid = O._sample(tr[None, :], d, "linear", "interp1")
    ref = np.interp(d, np.arange(L), tr.real) + 1j * np.interp(d, np.arange(L), tr.imag)
    assert valid.all()
    assert np.allclose(v[0], ref, atol=1e-12)


def test_fp64_cubic_reproduces_quadratics():
    """Keys cubic convolution (a = -1/2) is exact on polynomials of degree <= 2."""
    L = 40
    l = np.arange(L, dtype=np.float64)
    tr = (0.3 * l * l - 2.0 * l + 5.0) + 1j * (-0.1 * l * l + l)
    d = np.linspace(1.0, L - 2.0001, 97)
    v, valid = O._sample(tr[None, :], d, "cubic", "dasc")
    ref = (0.3 * d * d - 2.0 * d + 5.0) + 1j * (-0.1 * d * d + d)
    assert valid.all()
    assert np.allclose(v[0], ref, atol=1e-9)


@pytest.mark.parametrize("fd", [0.0, 0.05])
def test_mode_consistency(fd):
    """BOTH = sum NONE over (n, w); RX = sum over n; TX = sum over w."""
    pb = matrix_problem(seed=3, edges=False)
    args = (pb["ch"], pb["fs"], pb["t0"], pb["tx_apo"], pb["rx_apo"], pb["tx_delay"], pb["rx_delay"], fd)
    none = O.das_c(*args, 0).astype(np.complex128)
    rx, tx, both = (O.das_c(*args, d) for d in (1, 2, 3))
    peak = np.abs(both).max()
    assert np.abs(rx - none.sum(2, keepdims=True)).max() <= 1e-5 * np.abs(rx).max()
    assert np.abs(tx - none.sum(1, keepdims=True)).max() <= 1e-5 * np.abs(tx).max()
    assert np.abs(both - none.sum((1, 2), keepdims=True)).max() <= 1e-5 * peak


def test_frames_independent_and_thread_count_invariant():
    pb = matrix_problem(seed=4, F=3)
    args = (pb["fs"], pb["t0"], pb["tx_apo"], pb["rx_apo"], pb["tx_delay"], pb["rx_delay"], 0.05, 3)
    allf = O.das_c(pb["ch"], *args)
    for f in range(3):
        assert np.array_equal(allf[f:f + 1], O.das_c(pb["ch"][f:f + 1], *args))
    assert np.array_equal(O.das_c(pb["ch"], *args, nthreads=1), O.das_c(pb["ch"], *args, nthreads=7))


def test_linearity_and_zero_apodization():
    pb = matrix_problem(seed=5, F=1, edges=False)
    rest = (pb["fs"], pb["t0"], pb["tx_apo"], pb["rx_apo"], pb["tx_delay"], pb["rx_delay"], 0.05, 3)
    x, y = pb["ch"], np.roll(pb["ch"], 3, axis=-1)
    lhs = O.das_c((2 * x - 0.5 * y).astype(np.complex64), *rest)
    rhs = 2 * O.das_c(x, *rest).astype(np.complex128) - 0.5 * O.das_c(y, *rest)
    assert np.abs(lhs - rhs).max() <= 1e-5 * np.abs(rhs).max()
    z = O.das_c(x, pb["fs"], pb["t0"], np.zeros_like(pb["tx_apo"]), pb["rx_apo"], pb["tx_delay"],
                pb["rx_delay"], 0.05, 3)
    assert not z.any()


def test_nan_apodization_propagates_like_dasc():
    """das_c tests `if (c_rx_apo)`: NaN is nonzero, so the term is computed."""
    x = np.arange(32, dtype=np.complex64)
    assert np.isnan(_one(td=5.0, ra=float("nan"), x=x))
    # tx NaN passes the tx test; a zero rx weight still skips the term
    assert _one(td=5.0, ra=0.0, ta=float("nan"), x=x) == 0
    assert np.isnan(_one(td=5.0, ta=float("nan"), x=x))


@pytest.mark.parametrize("dim", [0, 1, 2, 3])
def test_dasc_vs_fp64_brute_force(dim):
    """das_c (FP32) vs FP64 on random inputs, at das_c's FP32 sample index."""
    pb = matrix_problem(seed=6 + dim, P=300, edges=False)
    args = (pb["ch"], pb["fs"], pb["t0"], pb["tx_apo"], pb["rx_apo"], pb["tx_delay"], pb["rx_delay"], 0.05, dim)
    a = O.das_c(*args)
    b = O.beamform_fp64(*args, w0=O.dasc_wd(0.05), index="fp32")
    assert np.abs(a - b).max() <= 1e-4 * np.abs(b).max()


def test_planted_mistake_is_caught():
    """A restatement with the interpolation weights swapped (a <-> b) fails
    the impulse closed form: the checks above can see such a mistake."""
    def swapped(td, rd, x):
        d = np.float32(np.float32(td) + np.float32(rd))
        k = int(d)
        b = float(d - math.floor(d))
        return b * x[k] + (1 - b) * x[k + 1]

    x = np.zeros(32)
    x[10], x[11] = 2.0, 6.0
    assert swapped(3.25, 7.0, x) != pytest.approx(0.75 * 2 + 0.25 * 6)
    assert _one(td=3.25, rd=7.0, x=x.astype(np.complex64)) == pytest.approx(0.75 * 2 + 0.25 * 6)


def test_python_restatement_matches_dasc_inside_window():
    """The Python reference's _beamform differs from das_c only at the trace
    edge (0 <= floor(d) vs 0 < trunc(d)) and in index rounding (/dt vs *fs)."""
    pb = matrix_problem(seed=9, P=200, F=2, edges=False)
    pb["tx_delay"] = np.abs(pb["tx_delay"]) + 1.0   # keep d >= 2: away from the edge rules
    pb["rx_delay"] = np.abs(pb["rx_delay"]) + 1.0
    ch = np.ascontiguousarray(pb["ch"].transpose(3, 2, 1, 0))  # (L, N, W, F)
    L = ch.shape[0]
    t = np.arange(L, dtype=np.float32)
    w0 = np.float32(2.0 * np.pi * np.float32(0.05))
    for dim in (0, 3):
        py = O.python_beamform(ch, t, pb["tx_apo"].T, pb["rx_apo"].T, pb["tx_delay"].T, pb["rx_delay"].T, w0, dim)
        dc = O.das_c(pb["ch"], 1.0, 0.0, pb["tx_apo"], pb["rx_apo"], pb["tx_delay"], pb["rx_delay"], 0.05, dim)
        dc = dc.transpose(3, 2, 1, 0)
        assert np.abs(py - dc).max() <= 1e-4 * np.abs(dc).max()
