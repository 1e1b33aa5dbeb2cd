"""GPU parity of the matrix (das_c MEX) contract against the das_c oracle.

Bar: bit-identical for RF (fd = 0: no sin/cos) when the das_c summation order
is requested; otherwise |gpu - oracle| <= 1e-5 * max|oracle| (FP32 tolerance
stated by the north-star contract). Cubic interpolation has no reference and
is checked against the FP64 oracle with the same bar.
"""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from tests.cases import matrix_problem
from ustb_amd import ops
from ustb_amd import _lib as L

pytestmark = pytest.mark.gpu
TOL = 1e-5


def _dev(pb, cuda):
    return {k: (torch.from_numpy(v).to(cuda) if isinstance(v, np.ndarray) else v) for k, v in pb.items()}


def _run(pb, cuda, fd, dim, **kw):
    d = _dev(pb, cuda)
    out = ops.das(d["ch"], d["fs"], d["t0"], d["tx_apo"], d["rx_apo"], d["tx_delay"], d["rx_delay"], fd, dim, **kw)
    torch.cuda.synchronize()
    return out.cpu().numpy()


def _rel(a, b):
    return float(np.max(np.abs(a.astype(np.complex128) - b)) / max(np.max(np.abs(b)), 1e-30))


@pytest.mark.parametrize("dim", [0, 1, 2, 3])
def test_rf_bit_exact_dasc_order(cuda, dim):
    pb = matrix_problem(seed=10 + dim)
    ref = O.das_c(pb["ch"], pb["fs"], pb["t0"], pb["tx_apo"], pb["rx_apo"], pb["tx_delay"], pb["rx_delay"],
                  0.0, dim)
    print(f"case rf dasc_order dim={dim}", flush=True)
    got = _run(pb, cuda, 0.0, dim, dasc_order=True)
    assert got.shape == ref.shape
    assert np.array_equal(got, ref), f"max diff {np.max(np.abs(got - ref))}"


@pytest.mark.parametrize("dim", [0, 1, 2, 3])
@pytest.mark.parametrize("split", [1, 2, 4, 8])
@pytest.mark.parametrize("fd", [0.0, 0.05])
def test_split_tolerance(cuda, dim, split, fd):
    pb = matrix_problem(seed=20 + dim)
    ref = O.das_c(pb["ch"], pb["fs"], pb["t0"], pb["tx_apo"], pb["rx_apo"], pb["tx_delay"], pb["rx_delay"], fd,
                  dim)
    print(f"case split={split} fd={fd} dim={dim}", flush=True)
    got = _run(pb, cuda, fd, dim, split=split)
    assert got.shape == ref.shape
    assert _rel(got, ref) <= TOL
    if dim in (0, 2) and fd == 0.0:
        # no cross-channel reduction: the das_c order is kept for any split
        assert np.array_equal(got, ref)


@pytest.mark.parametrize("dim", [0, 3])
def test_iq_dasc_order_tolerance(cuda, dim):
    pb = matrix_problem(seed=30 + dim, F=1)
    ref = O.das_c(pb["ch"], pb["fs"], pb["t0"], pb["tx_apo"], pb["rx_apo"], pb["tx_delay"], pb["rx_delay"],
                  0.05, dim)
    got = _run(pb, cuda, 0.05, dim, dasc_order=True)
    assert _rel(got, ref) <= TOL


@pytest.mark.parametrize("dim", [0, 3])
def test_python_edge_rule(cuda, dim):
    pb = matrix_problem(seed=40 + dim, F=1)
    ref = O.beamform_fp64(pb["ch"], pb["fs"], pb["t0"], pb["tx_apo"], pb["rx_apo"], pb["tx_delay"],
                          pb["rx_delay"], 0.0, dim, edge="python")
    got = _run(pb, cuda, 0.0, dim, edge=L.EDGE_PYTHON)
    assert _rel(got, ref) <= TOL


@pytest.mark.parametrize("dim", [0, 1, 2, 3])
def test_cubic_vs_fp64(cuda, dim):
    pb = matrix_problem(seed=50 + dim, F=2)
    ref = O.beamform_fp64(pb["ch"], pb["fs"], pb["t0"], pb["tx_apo"], pb["rx_apo"], pb["tx_delay"],
                          pb["rx_delay"], 0.05, dim, interp="cubic")
    got = _run(pb, cuda, 0.05, dim, interp=L.INTERP_CUBIC)
    assert _rel(got, ref) <= TOL


def test_empty_and_degenerate(cuda):
    pb = matrix_problem(seed=3, P=5, N=2, W=1, F=1, L=2)  # L = 2: no valid sample (0 < n0 < 1)
    got = _run(pb, cuda, 0.0, 3)
    assert np.all(got == 0)
    pb = matrix_problem(seed=3, P=0, N=2, W=1, F=1, L=16)
    got = _run(pb, cuda, 0.0, 3)
    assert got.shape == (1, 1, 1, 0)


def test_host_entry_matches_device(cuda):
    import ctypes

    pb = matrix_problem(seed=7, F=2)
    ref = O.das_c(pb["ch"], pb["fs"], pb["t0"], pb["tx_apo"], pb["rx_apo"], pb["tx_delay"], pb["rx_delay"], 0.0, 1)
    F, W, N, Ls = pb["ch"].shape
    P = pb["tx_delay"].shape[1]
    out = np.zeros(ref.shape, np.complex64)
    d = L.DasDesc(L=Ls, N=N, W=W, F=F, P=P, fs=1.0, t0=0.0, fd=0.0, dim=1, interp=0, edge=0,
                  flags=L.FLAG_DASC_ORDER, split=0)
    st = L.Status()
    rc = L.lib().ustbx_das_run_host(ctypes.byref(d), pb["ch"].ctypes.data, pb["tx_apo"].ctypes.data,
                                    pb["rx_apo"].ctypes.data, pb["tx_delay"].ctypes.data,
                                    pb["rx_delay"].ctypes.data, out.ctypes.data, 0, ctypes.byref(st))
    L.check(rc, st)
    assert np.array_equal(out, ref)


def test_host_entry_chunked_staging_matches_device(cuda):
    """The host entry stages inputs and outputs through two 8 MiB pinned
    buffers with a threaded CPU copy and keeps its device buffer across calls:
    inputs and outputs of several chunks (a 16.8 MB receive delay matrix, a
    33.6 MB dimension-none output), then a smaller call and a larger one
    (the buffer is reused, then grown), each equal to the device entry's
    image bit for bit."""
    import ctypes

    from ustb_amd import ops

    def problem(seed, P, N, W, F, Ls):
        rng = np.random.default_rng(seed)
        ch = (rng.standard_normal((F, W, N, Ls)) + 1j * rng.standard_normal((F, W, N, Ls))).astype(np.complex64)
        ta = rng.uniform(0, 1, (W, P)).astype(np.float32)
        ra = rng.uniform(0, 1, (N, P)).astype(np.float32)
        td = rng.uniform(0, Ls * 0.4, (W, P)).astype(np.float32)
        rd = rng.uniform(0, Ls * 0.4, (N, P)).astype(np.float32)
        return ch, ta, ra, td, rd

    for seed, (P, N, W, F, Ls, dim) in enumerate(((65536, 64, 2, 1, 256, 0), (4096, 16, 3, 2, 128, 3),
                                                    (131072, 48, 1, 1, 512, 1))):
        print(f"case host entry chunks P={P} N={N} W={W} F={F} L={Ls} dim={dim}", flush=True)
        ch, ta, ra, td, rd = problem(seed, P, N, W, F, Ls)
        shape = ops.out_shape(dim, P, N, W, F)
        out = np.zeros(shape, np.complex64)
        d = L.DasDesc(L=Ls, N=N, W=W, F=F, P=P, fs=1.0, t0=0.0, fd=0.0, dim=dim, interp=0, edge=0, flags=0,
                      split=0)
        st = L.Status()
        rc = L.lib().ustbx_das_run_host(ctypes.byref(d), ch.ctypes.data, ta.ctypes.data, ra.ctypes.data,
                                        td.ctypes.data, rd.ctypes.data, out.ctypes.data, 0, ctypes.byref(st))
        L.check(rc, st)
        T = lambda a: torch.from_numpy(a).to(cuda)
        ref = ops.das(T(ch), 1.0, 0.0, T(ta), T(ra), T(td), T(rd), 0.0, dim).cpu().numpy()
        assert np.array_equal(out, ref)


@pytest.mark.parametrize("dim", [0, 2, 3])
def test_nonfinite_and_huge_delays_and_nan_weights(cuda, dim):
    """das_c converts d = (tau - t0) * fs with (int): NaN, +-Inf and |d| >= 2^31
    fail its 0 < n0 < L-1 test and the term is skipped; a NaN weight is nonzero
    for das_c's `if (apo)` tests, so it propagates into the pixel's sum."""
    pb = matrix_problem(seed=60 + dim, P=300, N=6, W=3, F=2, L=64, edges=False)
    bad = np.array([np.nan, np.inf, -np.inf, 3e9, -3e9, 2.2e9, 1e38], np.float32)
    pb["tx_delay"][:, :len(bad)] = bad
    pb["rx_delay"][:, 10:10 + len(bad)] = bad
    pb["rx_apo"][2, 20] = np.nan
    pb["tx_apo"][1, 21] = np.nan
    pb["rx_apo"][:, 21] = 1.0
    ref = O.das_c(pb["ch"], pb["fs"], pb["t0"], pb["tx_apo"], pb["rx_apo"], pb["tx_delay"], pb["rx_delay"], 0.0, dim)
    assert np.isnan(ref).any() and np.isfinite(ref[..., :len(bad)]).all()
    print(f"case nonfinite delays dim={dim}", flush=True)
    got = _run(pb, cuda, 0.0, dim, dasc_order=True)
    assert np.array_equal(got, ref, equal_nan=True)
    for split in (2, 8):
        got = _run(pb, cuda, 0.05, dim, split=split)
        ref = O.das_c(pb["ch"], pb["fs"], pb["t0"], pb["tx_apo"], pb["rx_apo"], pb["tx_delay"], pb["rx_delay"],
                      0.05, dim)
        assert np.array_equal(np.isnan(got), np.isnan(ref))
        ok = ~np.isnan(ref)
        assert float(np.max(np.abs(got[ok] - ref[ok]))) <= TOL * float(np.max(np.abs(ref[ok])))


@pytest.mark.parametrize("dim", [0, 1, 2, 3])
@pytest.mark.parametrize("fd", [0.0, 0.05])
def test_torch_op_matches_dasc(cuda, dim, fd):
    """torch.ops.ustbx.das (the dispatcher-registered operator over the C-ABI)
    against the das_c oracle, within 1e-5 of the peak (the operator runs the
    default channel split, not das_c's summation order)."""
    from ustb_amd import torch_ops  # noqa: F401  (registers the operator)

    pb = matrix_problem(seed=70 + dim)
    print(f"case torch op dim={dim} fd={fd}", flush=True)
    ref = O.das_c(pb["ch"], pb["fs"], pb["t0"], pb["tx_apo"], pb["rx_apo"], pb["tx_delay"], pb["rx_delay"], fd, dim)
    d = _dev(pb, cuda)
    got = torch.ops.ustbx.das(d["ch"], d["fs"], d["t0"], d["tx_apo"], d["rx_apo"], d["tx_delay"], d["rx_delay"],
                              fd, dim)
    torch.cuda.synchronize()
    got = got.cpu().numpy()
    assert got.shape == ref.shape
    assert _rel(got, ref) <= TOL
