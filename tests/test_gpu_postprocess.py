"""GPU parity of the postprocess steps against the oracle's restatement.

- ustbx_coherent_compounding vs oracle.coherent_compounding (sums and the
  moving sum of coherent_compounding.m) on seeded data, all dimensions;
- ustbx_active_elements / ustbx_plan_active_elements vs
  oracle.active_elements: identical counts;
- ustbx_coherence_factor vs oracle.coherence_factor (MATLAB and Python
  flavours), including zero pixels (NaN -> 0) and pixels with no active
  element (1/M = Inf, as in MATLAB);
- ustbx_plan_run_cf (transmit kernel with the coherence factor in its
  epilogue) against ustbx_plan_run followed by ustbx_coherence_factor, over
  tile shapes, and the oracle's das_c + coherence factor;
- the facade classes (ustb_amd.postprocess) on a simulated point target: the
  reference integration tests' known answers.
Tolerances: FP32 sums against FP64, 1e-6 of the peak for compounding, 1e-5
of the peak for images and 1e-5 absolute for CF values (in [0, 1]).
"""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from tests.cases import channel_noise, geom_cases, oracle_matrices
from ustb_amd import _lib as L
from ustb_amd import ops
from ustb_amd.geometry import make_plan

pytestmark = pytest.mark.gpu
CASES = geom_cases()


def _rel(a, b):
    """max |a - b| / max |b| over the entries that are not NaN; the NaN
    positions must agree."""
    na, nb = np.isnan(a), np.isnan(b)
    assert np.array_equal(na, nb), "NaN positions differ"
    a, b = a[~na], b[~nb]
    return float(np.max(np.abs(a - b), initial=0.0) / max(np.max(np.abs(b), initial=0.0), 1e-30))


def _data(seed, F=3, W=5, N=7, P=301):
    rng = np.random.default_rng(seed)
    x = (rng.standard_normal((F, W, N, P)) + 1j * rng.standard_normal((F, W, N, P))).astype(np.complex64)
    x[:, :, :, :3] = 0  # pixels with no signal
    return x


@pytest.mark.parametrize("dim", [1, 2, 3])
@pytest.mark.parametrize("k", [0, 1, 2, 5])
def test_coherent_compounding_matches_oracle(cuda, dim, k):
    print(f"case compounding dim={dim} window={k}", flush=True)
    x = _data(10 + dim + k)
    got = ops.coherent_compounding(torch.from_numpy(x).to(cuda), dim, k).cpu().numpy()
    ref = O.coherent_compounding(x, dim, k or None)
    assert got.shape == ref.shape
    assert _rel(got, ref) <= 1e-6


def test_coherent_compounding_window_too_large(cuda):
    x = torch.from_numpy(_data(3)).to(cuda)
    with pytest.raises(L.USTBXError, match="window size greater"):
        ops.coherent_compounding(x, L.DIM_TX, 6)


@pytest.mark.parametrize("dim", [1, 2, 3])
def test_active_elements_match_oracle(cuda, dim):
    rng = np.random.default_rng(dim)
    W, N, P = 5, 7, 301
    ta = rng.uniform(-0.1, 1.0, (W, P)).astype(np.float32)
    ra = rng.uniform(-0.1, 1.0, (N, P)).astype(np.float32)
    ra[:, 5] = 0.0  # a pixel with no active receive element
    ta[:, 5] = 0.0
    ra[0, 7] = np.float32(0.16)  # exactly at the criterium in FP32: above 0.16 in FP64
    got = ops.active_elements(torch.from_numpy(ta).to(cuda), torch.from_numpy(ra).to(cuda), dim, P, N, W).cpu()
    ref = O.active_elements(dim, ta, ra)
    np.testing.assert_array_equal(got.numpy(), ref)
    ones = ops.active_elements(None, None, dim, P, N, W).cpu().numpy()
    np.testing.assert_array_equal(ones, {1: N, 2: W, 3: N * W}[dim])


@pytest.mark.parametrize("dim", [1, 2, 3])
@pytest.mark.parametrize("flavour", ["matlab", "python"])
def test_coherence_factor_matches_oracle(cuda, dim, flavour):
    print(f"case coherence factor dim={dim} {flavour}", flush=True)
    x = _data(20 + dim)
    F, W, N, P = x.shape
    count = {1: N, 2: W, 3: N * W}[dim]
    M = np.random.default_rng(dim).integers(1, count + 1, P).astype(np.float32)
    M[4] = 0.0  # no active element: 1/M = Inf
    Md = torch.from_numpy(M).to(cuda) if flavour == "matlab" else None
    fl = L.CF_MATLAB if flavour == "matlab" else L.CF_PYTHON
    img, cf = ops.coherence_factor(torch.from_numpy(x).to(cuda), dim, Md, fl)
    img, cf = img.cpu().numpy(), cf.cpu().numpy()
    rimg, rcf = O.coherence_factor(x, dim, M=M if flavour == "matlab" else None, flavour=flavour)
    assert img.shape == rimg.shape and cf.shape == rcf.shape
    fin = np.isfinite(rcf)
    np.testing.assert_array_equal(np.isfinite(cf), fin)
    if flavour == "matlab":
        assert not fin[..., 4].all()  # the Inf pixel
    assert np.all(cf[..., :3] == 0) and np.all(img[..., :3] == 0)
    assert np.max(np.abs(cf[fin] - rcf[fin]) / np.maximum(1.0, np.abs(rcf[fin]))) <= 1e-5
    fi = np.broadcast_to(fin, img.shape)
    assert _rel(img[fi], rimg[fi]) <= 1e-5


@pytest.mark.parametrize("name", ["dw_sector", "dw_sector_patch", "cpwc_tukey", "pw_rf_hanning", "pw_rf_patch",
                                  "cpwc_depth0", "cpwc_depth0_patch", "focused_hybrid"])
@pytest.mark.parametrize("separable", [True, False])
def test_fused_coherence_factor_equals_unfused(cuda, name, separable):
    """ustbx_plan_run_cf against ustbx_plan_run + ustbx_coherence_factor
    (receive) within FP32 rounding of the grouped sums (2e-6 of the peak
    image, 1e-6 on the factor), and the
    oracle's das_c + coherence factor within 1e-5. cpwc_depth0 and
    focused_hybrid hold NaN weights (depth 0, focal depth): their tiles take
    the SAFE body and the NaN positions must agree."""
    case = CASES[name]
    F = 9
    print(f"case fused cf {name} separable={separable}", flush=True)
    plan = make_plan(case["probe"], case["scan"], case["seq"], n_samples=case["L"], n_frames=F,
                     sampling_frequency=case["fs"], initial_time=0.0, modulation_frequency=case["fd"],
                     sound_speed=1540.0, receive_apodization=case["rx"], transmit_apodization=case["tx"],
                     dimension=2, tx_model=case["model"], separable_phase=separable)
    assert plan.variant[0] == 2
    ch = channel_noise(case, F=F, seed=41)
    chd = torch.from_numpy(ch).to(cuda)
    M = plan.active_elements(L.DIM_RX)
    per_channel = plan.run(chd)
    img_u, cf_u = ops.coherence_factor(per_channel, L.DIM_RX, M, L.CF_MATLAB)
    for fc, cs in ((8, 8), (4, 4), (8, 4)):
        plan.set_tiles(fc, cs)
        img_f, cf_f = plan.run_cf(chd, M, L.CF_MATLAB)
        a, b = img_f.cpu().numpy(), img_u[:, 0, 0].cpu().numpy()
        assert _rel(a, b) <= 2e-6
        assert np.max(np.abs(cf_f.cpu().numpy() - cf_u[:, 0, 0].cpu().numpy())) <= 1e-6  # NaN -> 0
    img_p, _ = plan.run_cf(chd, None, L.CF_PYTHON)
    img_pu, _ = ops.coherence_factor(per_channel, L.DIM_RX, None, L.CF_PYTHON)
    assert _rel(img_p.cpu().numpy(), img_pu[:, 0, 0].cpu().numpy()) <= 2e-6
    # against the oracle: das_c per-channel images, then the coherence factor
    ta, ra, td, rd = oracle_matrices(case)
    ref_bf = O.das_c(ch, case["fs"], 0.0, ta, ra, td, rd, case["fd"], 2)
    Mo = O.active_elements(O.DIM_RX, None, ra)
    np.testing.assert_array_equal(M.cpu().numpy(), Mo)
    rimg, rcf = O.coherence_factor(ref_bf, O.DIM_RX, M=Mo)
    got_img, got_cf = img_f.cpu().numpy(), cf_f.cpu().numpy()
    fin = np.isfinite(rcf[:, 0, 0])
    tol = 1e-5 if (not separable or case["fd"] == 0.0) else 1e-4  # separable: das_c's own phase rounding
    assert np.max(np.abs(got_cf[fin] - rcf[:, 0, 0][fin])) <= tol
    assert _rel(got_img[fin], rimg[:, 0, 0][fin]) <= tol
    plan.close()


def test_fused_coherence_factor_needs_transmit_plan(cuda):
    case = CASES["cpwc_tukey"]
    plan = make_plan(case["probe"], case["scan"], case["seq"], n_samples=case["L"], n_frames=1,
                     sampling_frequency=case["fs"], initial_time=0.0, modulation_frequency=case["fd"],
                     sound_speed=1540.0, receive_apodization=case["rx"], transmit_apodization=case["tx"],
                     dimension=3, tx_model=case["model"])
    ch = torch.from_numpy(channel_noise(case, F=1)).to(cuda)
    with pytest.raises(L.USTBXError, match="dimension-transmit"):
        plan.run_cf(ch)
    plan.close()


def _point_target_das(cuda, dim):
    from ustb_amd import synth
    from ustb_amd.das import DAS
    from ustb_amd.enums import Dimension, Window
    from ustb_amd.uff import Apodization, ChannelData, LinearArray, LinearScan, Wave

    prb = LinearArray(64, 0.3e-3)
    seq = [Wave.plane(a) for a in np.linspace(-0.3, 0.3, 5)]
    rf = synth.echoes(prb, seq, np.array([(0.0, 0.0, 20e-3)]), np.ones(1), fs=41.6e6, n_samples=2400,
                      f0=5.2e6, bandwidth=0.6, iq=False)
    d = DAS()
    d.channel_data = ChannelData(probe=prb, sequence=seq, sampling_frequency=41.6e6, modulation_frequency=0.0,
                                 data=np.ascontiguousarray(rf.transpose(3, 2, 1, 0)))
    d.scan = LinearScan(np.linspace(-3e-3, 3e-3, 40), np.linspace(18e-3, 22e-3, 40))
    d.receive_apodization = Apodization(Window.boxcar, f_number=1.7)
    d.transmit_apodization = Apodization(Window.none)
    d.dimension = dim
    return d, d.go()


def test_coherence_factor_facade_reference_known_answers(cuda):
    """integration_coherence_factor_test.m:4-123 through the device path:
    das with dimension none, then the coherence factor (dimension both) with
    the das apodization; and the same against the oracle on the same data."""
    from ustb_amd.enums import Dimension
    from ustb_amd.postprocess import CoherenceFactor

    d, bf = _point_target_das(cuda, Dimension.none)
    cf = CoherenceFactor()
    cf.input = bf
    cf.receive_apodization = d.receive_apodization
    cf.transmit_apodization = d.transmit_apodization
    out = cf.go()
    assert tuple(out.data.shape) == (1600, 1, 1, 1)
    img = out.data.cpu().numpy()
    cfv = cf.CF.data.cpu().numpy()
    assert np.all(np.isfinite(img)) and np.all(np.isfinite(cfv))
    assert cfv.min() >= 0 and cfv.max() <= 1 + 1e-6
    assert cfv.max() > 0.3
    # the oracle on the device's own beamformed data and apodization
    x = bf.data.permute(3, 2, 1, 0).contiguous().cpu().numpy()
    plan = d._plans[next(iter(d._plans))]
    ta, ra, _, _ = [t.cpu().numpy() for t in plan.materialize()]
    M = O.active_elements(O.DIM_BOTH, ta, ra)
    rimg, rcf = O.coherence_factor(x, O.DIM_BOTH, M=M)
    assert np.max(np.abs(cfv[:, 0, 0, 0] - rcf[0, 0, 0])) <= 1e-5
    assert _rel(img[:, 0, 0, 0], rimg[0, 0, 0]) <= 1e-5


def test_coherent_compounding_facade_reference_known_answers(cuda):
    """integration_coherent_compounding_test.m:3-60: one wave after
    compounding, finite, the peak within 2 mm of (0, 20 mm)."""
    from ustb_amd.enums import Dimension
    from ustb_amd.postprocess import CoherentCompounding

    d, bf = _point_target_das(cuda, Dimension.none)
    cc = CoherentCompounding()
    cc.input = bf
    out = cc.go()
    data = out.data.cpu().numpy()
    assert data.shape == (1600, 1, 1, 1) and np.all(np.isfinite(data))
    k = int(np.argmax(np.abs(data[:, 0, 0, 0])))
    assert abs(d.scan.x[k]) <= 2e-3 and abs(d.scan.z[k] - 20e-3) <= 2e-3
    cc.dimension = Dimension.transmit
    cc.window_size = 3
    moving = cc.go().data.cpu().numpy()
    ref = O.coherent_compounding(bf.data.permute(3, 2, 1, 0).contiguous().cpu().numpy(), O.DIM_TX, 3)
    assert moving.shape == ref.transpose(3, 2, 1, 0).shape
    assert _rel(moving, ref.transpose(3, 2, 1, 0)) <= 1e-6
