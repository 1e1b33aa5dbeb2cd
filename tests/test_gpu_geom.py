"""GPU parity of the on-the-fly geometry path.

1. The device's FP32 delay/apodization matrices equal the oracle's
   restatement of das.m / apodization.m (delays bit-identical; apodization
   within one FP32 ulp, which only a last-bit difference of FP64 cos/atan2
   between the device library and NumPy can cause).
2. Beamforming from geometry equals the das_c oracle fed with the oracle's
   matrices: bit-identical for RF in das_c order, else within 1e-5 of the
   peak magnitude.
"""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from tests.cases import channel_noise, geom_cases, oracle_matrices
from ustb_amd.enums import Interpolation
from ustb_amd.geometry import make_plan

pytestmark = pytest.mark.gpu
TOL = 1e-5
CASES = geom_cases()


def _plan(case, dim=3, F=1, **kw):
    return make_plan(case["probe"], case["scan"], case["seq"], n_samples=case["L"], n_frames=F,
                     sampling_frequency=case["fs"], initial_time=0.0, modulation_frequency=case["fd"],
                     sound_speed=1540.0, receive_apodization=case["rx"], transmit_apodization=case["tx"],
                     dimension=dim, tx_model=case["model"], **kw)


def _rel(a, b):
    """max |a - b| / max |b| over finite entries; NaN positions must agree
    (pixels at a focal depth get NaN apodization, as in apodization.m)."""
    na, nb = np.isnan(a), np.isnan(b)
    assert np.array_equal(na, nb), "NaN positions differ"
    a = np.where(na, 0, a).astype(np.complex128)
    b = np.where(nb, 0, b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-30))


@pytest.mark.parametrize("name", sorted(CASES))
def test_materialized_geometry_matches_oracle(cuda, name):
    case = CASES[name]
    print(f"case materialize {name}", flush=True)
    plan = _plan(case)
    ta, ra, td, rd = [t.cpu().numpy() for t in plan.materialize()]
    ota, ora, otd, ord_ = oracle_matrices(case)
    assert np.array_equal(rd, ord_), np.max(np.abs(rd - ord_))
    assert np.array_equal(td, otd), np.max(np.abs(td - otd))
    for got, ref in ((ta, ota), (ra, ora)):
        assert got.shape == ref.shape
        same_nan = np.isnan(got) == np.isnan(ref)
        assert same_nan.all()
        g, r = np.nan_to_num(got), np.nan_to_num(ref)
        assert np.max(np.abs(g - r)) <= 1.2e-7, np.max(np.abs(g - r))
    plan.close()


@pytest.mark.parametrize("name", sorted(CASES))
@pytest.mark.parametrize("dim", [0, 1, 2, 3])
def test_geometry_beamform_dasc_order(cuda, name, dim):
    case = CASES[name]
    F = 2
    print(f"case geom dasc_order {name} dim={dim}", flush=True)
    plan = _plan(case, dim=dim, F=F, separable_phase=False, dasc_order=True)
    ch = channel_noise(case, F=F)
    out = plan.run(torch.from_numpy(ch).to(cuda))
    got = out.cpu().numpy()
    ta, ra, td, rd = [t.cpu().numpy() for t in plan.materialize()]
    ref = O.das_c(ch, case["fs"], 0.0, ta, ra, td, rd, case["fd"], dim)
    if case["fd"] == 0.0:
        assert np.array_equal(got, ref, equal_nan=True), np.nanmax(np.abs(got - ref))
    else:
        assert _rel(got, ref) <= TOL
    plan.close()


@pytest.mark.parametrize("name", sorted(CASES))
@pytest.mark.parametrize("dim", [0, 1, 2, 3])
def test_geometry_beamform_fast(cuda, name, dim):
    """Separable phase on noise data vs the FP64 oracle at das_c's wd.

    das_c's own per-term phase rounding (about 1e-4 rad at these delays) does
    not average down on incoherent noise, so on noise the bar is taken against
    the FP64 oracle; the das_c bar is applied on coherent data below."""
    case = CASES[name]
    F = 3 if dim == 2 else 1
    print(f"case geom fast {name} dim={dim}", flush=True)
    plan = _plan(case, dim=dim, F=F, separable_phase=True)
    ch = channel_noise(case, F=F, seed=5)
    got = plan.run(torch.from_numpy(ch).to(cuda)).cpu().numpy()
    ta, ra, td, rd = oracle_matrices(case)
    ref = O.beamform_fp64(ch, case["fs"], 0.0, ta, ra, td, rd, case["fd"], dim, w0=O.dasc_wd(case["fd"]),
                          index="fp32")
    assert _rel(got, ref) <= TOL
    plan.close()


@pytest.mark.parametrize("name", sorted(CASES))
def test_geometry_fast_vs_dasc_point_targets(cuda, name):
    """The north-star bar: |gpu - das_c| <= 1e-5 * max|das_c| on synthetic
    echoes of point scatterers (coherent data)."""
    from ustb_amd import synth

    case = CASES[name]
    scan = case["scan"]
    rng = np.random.default_rng(17)
    idx = rng.choice(scan.N_pixels, 6, replace=False)
    pts = np.stack([scan.x[idx], scan.y[idx], scan.z[idx]], 1)
    iq = case["fd"] != 0.0
    f0 = case["fd"] if iq else case["fs"] / 4
    data = synth.echoes(case["probe"], case["seq"], pts, rng.standard_normal(6) + 2.0, fs=case["fs"],
                        n_samples=case["L"], f0=f0, bandwidth=0.6, iq=iq)
    ch = data if iq else synth.analytic(data)
    plan = _plan(case, dim=3, separable_phase=True)
    got = plan.run(torch.from_numpy(ch).to(cuda)).cpu().numpy()
    ta, ra, td, rd = oracle_matrices(case)
    ref = O.das_c(ch, case["fs"], 0.0, ta, ra, td, rd, case["fd"], 3)
    exact = O.beamform_fp64(ch, case["fs"], 0.0, ta, ra, td, rd, case["fd"], 3, w0=O.dasc_wd(case["fd"]),
                            index="fp32")
    err_gpu, err_dasc = _rel(got, ref), _rel(ref, exact)
    print(f"{name}: |gpu-das_c| {err_gpu:.2e}, |das_c-exact| {err_dasc:.2e}, |gpu-exact| {_rel(got, exact):.2e}")
    # small apertures: das_c's own FP32 phase rounding can exceed 1e-5 of the peak
    assert err_gpu <= max(TOL, 2.0 * err_dasc)
    assert _rel(got, exact) <= TOL
    plan.close()


@pytest.mark.parametrize("name", ["cpwc_tukey", "dw_sector"])
def test_geometry_cubic_vs_fp64(cuda, name):
    case = CASES[name]
    plan = _plan(case, dim=3, separable_phase=True, interpolation=Interpolation.cubic)
    ch = channel_noise(case, seed=9)
    got = plan.run(torch.from_numpy(ch).to(cuda)).cpu().numpy()
    ta, ra, td, rd = oracle_matrices(case)
    ref = O.beamform_fp64(ch, case["fs"], 0.0, ta, ra, td, rd, case["fd"], 3, interp="cubic",
                          w0=O.dasc_wd(case["fd"]), index="fp32")
    assert _rel(got, ref) <= TOL
    plan.close()


@pytest.mark.parametrize("name", ["cpwc_tukey", "dw_sector"])
def test_geometry_reference_phase(cuda, name):
    case = CASES[name]
    plan = _plan(case, dim=3, separable_phase=True, reference_phase=True)
    ch = channel_noise(case, seed=11)
    got = plan.run(torch.from_numpy(ch).to(cuda)).cpu().numpy()
    ta, ra, td, rd = oracle_matrices(case)
    ref = O.beamform_fp64(ch, case["fs"], 0.0, ta, ra, td, rd, case["fd"], 3, w0=O.dasc_wd(case["fd"]),
                          index="fp32")
    scan = case["scan"]
    ref = ref * O.reference_phase(scan.reference_distance, case["fd"], 1540.0)[None, None, None, :]
    assert _rel(got, ref) <= TOL
    plan.close()


@pytest.mark.parametrize("flavour", ["python", "matlab"])
def test_reference_phase_flavours(cuda, flavour):
    """The plan's reference-phase factor per pixel (the image with the factor
    over the image without it) against the oracle's flavour: das.py's FP64
    argument or das.m's single-precision one. The two flavours differ by up to
    about 1e-4 on this scan (30 mm at 5.2 MHz), so each must match its own."""
    case = CASES["cpwc_tukey"]
    print(f"case reference phase {flavour}", flush=True)
    ch = torch.from_numpy(channel_noise(case, seed=12)).to(cuda)
    plain = _plan(case, dim=3, separable_phase=True)
    with_ref = _plan(case, dim=3, separable_phase=True, reference_phase=flavour)
    a = plain.run(ch).cpu().numpy()[0, 0, 0].astype(np.complex128)
    b = with_ref.run(ch).cpu().numpy()[0, 0, 0].astype(np.complex128)
    keep = np.abs(a) > 1e-3 * np.abs(a).max()
    ratio = b[keep] / a[keep]
    r = case["scan"].reference_distance[keep]
    own = O.reference_phase(r, case["fd"], 1540.0, flavour=flavour)
    other = O.reference_phase(r, case["fd"], 1540.0, flavour="matlab" if flavour == "python" else "python")
    assert np.max(np.abs(ratio - own)) <= 1e-6
    assert np.max(np.abs(own - other)) > 2e-5 and np.max(np.abs(ratio - other)) > 2e-5
    plain.close()
    with_ref.close()


@pytest.mark.parametrize("variant,wc,split", [(0, 8, 1), (0, 8, 4), (0, 16, 1), (0, 16, 4), (1, 8, 4),
                                              (1, 10, 4), (1, 12, 8), (1, 16, 8), (2, 4, 4), (2, 6, 8),
                                              (2, 8, 8), (2, 10, 4), (2, 12, 8), (2, 16, 4), (2, 5, 5)])
@pytest.mark.parametrize("name", ["cpwc_tukey", "cpwc_depth0"])
def test_tile_shapes_agree(cuda, variant, wc, split, name):
    """Every tile shape of variants 0-2 against the FP64 oracle; cpwc_depth0
    has NaN receive weights on its first row (NaN positions must agree)."""
    case = CASES[name]
    print(f"case tiles {name} variant={variant} wc={wc} split={split}", flush=True)
    ch = torch.from_numpy(channel_noise(case, seed=13)).to(cuda)
    plan = _plan(case, dim=3, separable_phase=True)
    plan.set_variant(variant)
    plan.set_tiles(wave_chunk=wc, split=split)
    got = plan.run(ch).cpu().numpy()
    ta, ra, td, rd = oracle_matrices(case)
    ref = O.beamform_fp64(ch.cpu().numpy(), case["fs"], 0.0, ta, ra, td, rd, case["fd"], 3,
                          w0=O.dasc_wd(case["fd"]), index="fp32")
    assert _rel(got, ref) <= TOL
    plan.close()


@pytest.mark.parametrize("name", sorted(CASES))
@pytest.mark.parametrize("interp", [Interpolation.linear, Interpolation.cubic])
def test_receive_table_variant(cuda, name, interp):
    """Variant 2 (resident receive table, branch-free gathers) applies to
    dimension both whether or not every nonzero weight is finite (tiles with
    a non-finite one take the per-term window test); it matches the FP64
    oracle (das_c's wd and FP32 index), NaN positions included, and variant 0
    on two frames, in either tile order."""
    case = CASES[name]
    print(f"case rxtab {name} {interp.name}", flush=True)
    plan = _plan(case, dim=3, F=2, separable_phase=True, interpolation=int(interp))
    assert plan.has_variant(2)
    assert plan.variant[0] == 2  # the default where it applies
    ch = channel_noise(case, F=2, seed=21)
    chd = torch.from_numpy(ch).to(cuda)
    got = plan.run(chd).cpu().numpy()
    plan.set_order(0)
    assert np.array_equal(plan.run(chd).cpu().numpy(), got, equal_nan=True)
    plan.set_order(1)
    if interp == Interpolation.cubic:  # the coefficient form of the same polynomial
        plan.set_variant(3)
        for wc, ws in ((6, 4), (10, 8), (16, 8)):
            plan.set_tiles(wc, ws)
            assert _rel(plan.run(chd).cpu().numpy(), got) <= TOL
    plan.set_variant(0)
    v0 = plan.run(chd).cpu().numpy()
    assert _rel(got, v0) <= TOL
    if interp == Interpolation.linear:
        ta, ra, td, rd = oracle_matrices(case)
        ref = O.beamform_fp64(ch, case["fs"], 0.0, ta, ra, td, rd, case["fd"], 3, w0=O.dasc_wd(case["fd"]),
                              index="fp32")
        assert _rel(got, ref) <= TOL
    plan.close()


@pytest.mark.parametrize("name", sorted(CASES))
@pytest.mark.parametrize("interp", [Interpolation.linear, Interpolation.cubic])
def test_transmit_table_variant(cuda, name, interp):
    """Dimension transmit from the resident tables (frames gathered together):
    9 frames, so the second chunk of 8 holds one frame; against variant 0 and
    the FP64 oracle, over its frame/channel-split shapes."""
    case = CASES[name]
    print(f"case txtab {name} {interp.name}", flush=True)
    plan = _plan(case, dim=2, F=9, separable_phase=True, interpolation=int(interp))
    assert plan.has_variant(2)  # tiles with non-finite weights take the SAFE body
    assert plan.variant[0] == 2
    ch = channel_noise(case, F=9, seed=23)
    chd = torch.from_numpy(ch).to(cuda)
    outs = []
    for fc, cs in ((8, 4), (4, 8), (4, 4), (8, 8)):
        plan.set_tiles(fc, cs)
        outs.append(plan.run(chd).cpu().numpy())
    for o in outs[1:]:
        assert _rel(o, outs[0]) <= TOL
    plan.set_variant(0)
    v0 = plan.run(chd).cpu().numpy()
    assert _rel(outs[0], v0) <= TOL
    if interp == Interpolation.linear:
        ta, ra, td, rd = oracle_matrices(case)
        ref = O.beamform_fp64(ch, case["fs"], 0.0, ta, ra, td, rd, case["fd"], 2, w0=O.dasc_wd(case["fd"]),
                              index="fp32")
        assert _rel(outs[0], ref) <= TOL
    plan.close()


@pytest.mark.parametrize("name", sorted(CASES))
@pytest.mark.parametrize("dasc_order", [False, True])
def test_transmit_table_dasc_phase(cuda, name, dasc_order):
    """Dimension transmit from the resident tables with das_c's phase (and
    with das_c's order): das_c's term arithmetic and summation order, so the
    image equals the exact per-lane kernel's (variant 0) bit for bit, and
    das_c's bit for bit for RF data (within 1e-5 of the peak for IQ, where
    only the device's and the host's sin/cos differ). 9 frames: the second
    chunk of 8 holds one frame."""
    case = CASES[name]
    F = 9
    print(f"case tx dasc phase {name} dasc_order={dasc_order}", flush=True)
    plan = _plan(case, dim=2, F=F, separable_phase=False, dasc_order=dasc_order)
    assert plan.has_variant(2)
    assert plan.variant[0] == 2
    ch = channel_noise(case, F=F, seed=29)
    chd = torch.from_numpy(ch).to(cuda)
    outs = []
    for fc, cs in ((8, 4), (4, 8)):
        plan.set_tiles(fc, cs)
        outs.append(plan.run(chd).cpu().numpy())
    assert np.array_equal(outs[1], outs[0], equal_nan=True)
    ta, ra, td, rd = [t.cpu().numpy() for t in plan.materialize()]
    plan.set_variant(0)
    v0 = plan.run(chd).cpu().numpy()
    ref = O.das_c(ch, case["fs"], 0.0, ta, ra, td, rd, case["fd"], 2)
    assert np.array_equal(outs[0], v0, equal_nan=True)
    if case["fd"] == 0.0:
        assert np.array_equal(outs[0], ref, equal_nan=True)
    else:
        assert _rel(outs[0], ref) <= TOL
    plan.close()


@pytest.mark.parametrize("name", ["cpwc_tukey", "dw_sector"])
def test_transmit_table_dasc_phase_cubic(cuda, name):
    """Cubic interpolation with das_c's phase in dimension transmit: the
    resident-table kernel against variant 0 and the FP64 oracle taking das_c's
    FP32 phase argument (no das_c counterpart: das_c is linear)."""
    case = CASES[name]
    plan = _plan(case, dim=2, F=3, separable_phase=False, interpolation=int(Interpolation.cubic))
    assert plan.variant[0] == 2
    ch = channel_noise(case, F=3, seed=31)
    chd = torch.from_numpy(ch).to(cuda)
    got = plan.run(chd).cpu().numpy()
    plan.set_variant(0)
    assert _rel(got, plan.run(chd).cpu().numpy()) <= TOL
    plan.close()


def test_unified_model_needs_a_2d_scan(cuda):
    """das.m:148-149: the unified model covers 2-D linear and sector scans."""
    from ustb_amd import _lib as L

    case = CASES["matrix3d_dw"]
    seq = [type(w).spherical(0.0, 0.0, 10e-3) for w in case["seq"][:1]]
    with pytest.raises(L.USTBXError, match="Only linear scan and sector scan"):
        make_plan(case["probe"], case["scan"], seq, n_samples=case["L"], n_frames=1,
                  sampling_frequency=case["fs"], initial_time=0.0, modulation_frequency=case["fd"],
                  sound_speed=1540.0, receive_apodization=case["rx"], transmit_apodization=case["tx"],
                  dimension=3, tx_model=2)


def test_receive_table_rejected_modes(cuda):
    """Variants 2/3 cover dimension both and transmit (3: both only), and
    dimension both not in das_c's summation order or with das_c's phase."""
    case = CASES["cpwc_tukey"]
    for dim in (0, 1):
        plan = _plan(case, dim=dim, separable_phase=True)
        assert not plan.has_variant(2)
        plan.close()
    plan = _plan(case, dim=2, separable_phase=True, interpolation=int(Interpolation.cubic))
    assert plan.has_variant(2) and not plan.has_variant(3)
    plan.close()
    plan = _plan(case, dim=3, dasc_order=True)
    assert not plan.has_variant(2)
    assert plan.variant[0] == 0
    plan.close()
    plan = _plan(case, dim=3, separable_phase=False)
    assert not plan.has_variant(2)
    plan.close()


@pytest.mark.parametrize("name", ["dw_sector", "dw_sector_patch", "cpwc_tukey", "pw_rf_hanning", "focused_hybrid",
                                  "cpwc_depth0"])
@pytest.mark.parametrize("separable", [True, False])
@pytest.mark.parametrize("interp", [Interpolation.linear, Interpolation.cubic])
def test_transmit_staging_equals_gathers(cuda, name, separable, interp):
    """Dimension transmit: sample windows staged through LDS give the same
    per-channel images as one gather per term, bit for bit, for every tile
    shape (frames x channels per block), 9 frames (a partial frame chunk);
    NaN-weight tiles (cpwc_depth0, focused_hybrid) run the gather body. On
    the patch tiles of dw_sector_patch the plan's mask stages most (tile,
    channel) pairs; along 64 depths the windows are too wide."""
    case = CASES[name]
    F = 9
    print(f"case staging {name} separable={separable} interp={int(interp)}", flush=True)
    plan = _plan(case, dim=2, F=F, separable_phase=separable, interpolation=interp)
    assert plan.variant[0] == 2
    plan.set_staging(True)
    staged, listed = plan.stage_stats
    if name == "dw_sector_patch":
        assert staged > listed // 2, (staged, listed)
    ch = torch.from_numpy(channel_noise(case, F=F, seed=21)).to(cuda)
    for fc, cs in ((8, 8), (4, 4), (8, 4), (4, 8)):
        plan.set_tiles(fc, cs)
        plan.set_staging(False)
        a = plan.run(ch).cpu().numpy()
        plan.set_staging(True)
        b = plan.run(ch).cpu().numpy()
        assert np.array_equal(a, b, equal_nan=True), (fc, cs, np.nanmax(np.abs(a - b)))
    plan.close()


def _staged_count(ta, ra, td, rd, L, fs, t0, cubic):
    """(tile, channel) pairs the stage mask should mark, from das.m's FP32
    matrices: full 64-pixel tiles whose nonzero weights and delays are all
    finite; per channel with a nonzero weight on the tile, every weight of it
    and of the tile's waves nonzero on all 64 pixels, every term's index
    d = ((td + rd) - t0) * fs (FP32) in [1, lim) and every lane's trunc(d)
    within [that of lane 0, +60] (cubic; +62 linear)."""
    W, P = ta.shape
    lim = np.float32(L - 2 if cubic else L - 1)
    span = 60 if cubic else 62
    t0, fs = np.float32(t0), np.float32(fs)
    count = 0
    for t in range(P // 64):
        sl = slice(64 * t, 64 * t + 64)
        a_t, d_t, a_r, d_r = ta[:, sl], td[:, sl], ra[:, sl], rd[:, sl]
        tx_on, rx_on = a_t != 0, a_r != 0
        if np.any(tx_on & ~(np.isfinite(d_t) & np.isfinite(a_t))) or \
                np.any(rx_on & ~(np.isfinite(d_r) & np.isfinite(a_r))):
            continue
        waves = np.nonzero(tx_on.any(axis=1))[0]
        chans = np.nonzero(rx_on.any(axis=1))[0]
        if waves.size == 0 or chans.size == 0:
            continue
        tw_ok = bool(np.all(tx_on[waves]))
        d = ((d_t[waves][None, :, :] + d_r[chans][:, None, :]) - t0) * fs  # float32 throughout
        n0 = d.astype(np.int32)
        ok = np.all(rx_on[chans], axis=1) & tw_ok
        ok &= np.all((d >= 1) & (d < lim), axis=(1, 2))
        rel = n0 - n0[:, :, :1]
        ok &= np.all((rel >= 0) & (rel <= span), axis=(1, 2))
        count += int(ok.sum())
    return count


def _stage_plan(name, F):
    """Plans for the dimension-both staging test: the geometry cases, and C2's
    probe, waves and depths at a quarter of its elements, waves and scan lines
    (64 lines x 1024 depths, 1024 samples), as IQ (cubic, linear) and as RF
    (fd = 0: the compact receive table). With C2's depth spacing a 64-pixel
    tile spans about 42 samples, so most (tile, channel) pairs are staged."""
    from ustb_amd import configs
    from ustb_amd.uff import LinearScan
    if name in CASES:
        case = CASES[name]
        interp = Interpolation.cubic if name != "pw_rf_hanning" else Interpolation.linear
        return _plan(case, dim=3, F=F, separable_phase=True, interpolation=interp), case["L"], case["fs"], 0.0
    interp = Interpolation.cubic if name == "c2q_cubic" else Interpolation.linear
    acq = configs.c2(scale=0.25, with_data=False)
    scan = LinearScan(np.linspace(-20e-3, 20e-3, 64), np.linspace(5e-3, 55e-3, 1024))
    kw = acq.plan_kwargs(n_samples=1024, n_frames=F, interpolation=interp)
    if name == "c2q_rf":
        kw["modulation_frequency"] = 0.0
    plan = make_plan(acq.probe, scan, acq.sequence, separable_phase=True, **kw)
    return plan, 1024, acq.fs, acq.t0


@pytest.mark.parametrize("name", ["c2q_cubic", "c2q_linear", "c2q_rf", "cpwc_tukey", "dw_sector", "pw_rf_hanning",
                                  "cpwc_depth0", "matrix3d_dw", "focused_hybrid"])
def test_both_staging_equals_gathers(cuda, name):
    """Dimension both: the (tile, channel) pairs the plan marks run from
    LDS-staged sample windows. The mask marks exactly the pairs das.m's FP32
    matrices allow (recounted on the host), and the images equal the gather
    body's bit for bit for every tile shape (full and partial wave rounds,
    channel groups), NaN-weight tiles included; two frames."""
    F = 2
    print(f"case both-staging {name}", flush=True)
    plan, L, fs, t0 = _stage_plan(name, F)
    if plan.variant[0] == 4:
        plan.set_variant(2)
    assert plan.variant[0] == 2
    ta, ra, td, rd = [t.cpu().numpy() for t in plan.materialize()]
    cubic = int(plan.interp) == int(Interpolation.cubic)
    staged, listed = plan.stage_stats
    assert staged == _staged_count(ta, ra, td, rd, L, fs, t0, cubic), (staged, listed)
    if name.startswith("c2q"):
        assert staged > 0.5 * listed, (staged, listed)
    g = torch.Generator(device=cuda).manual_seed(41)
    shape = (F, plan.W, plan.N, L)
    ch = torch.complex(torch.randn(shape, device=cuda, generator=g), torch.randn(shape, device=cuda, generator=g))
    for wc, ws in ((4, 4), (5, 5), (8, 4), (16, 4), (6, 8)):
        plan.set_tiles(wc, ws)
        plan.set_staging(False)
        a = plan.run(ch).cpu().numpy()
        plan.set_staging(True)
        b = plan.run(ch).cpu().numpy()
        assert np.array_equal(a, b, equal_nan=True), (wc, ws, np.nanmax(np.abs(a - b)))
    plan.close()


@pytest.mark.parametrize("name", ["cpwc_tukey", "dw_sector", "pw_rf_hanning", "cpwc_depth0", "matrix3d_dw",
                                  "focused_hybrid", "sta_hybrid"])
@pytest.mark.parametrize("interp", [Interpolation.linear, Interpolation.cubic])
def test_receive_state_on_the_fly(cuda, name, interp):
    """Variant 4 (receive delays and weights computed in the kernel, no
    receive table: das.m's delay from an FP32 square root refined by one FP64
    Newton step, the clamps, ratios and window in FP32) against
    variant 2 (the resident table of das.m's FP32 matrices) within 2e-6 of the
    peak with identical NaN positions, and against the oracle's FP64
    restatement on its matrices (the separable phase's reference; das_c's
    own FP32 phase rounding on noise data is larger) within 1e-5; F = 3."""
    case = CASES[name]
    F = 3
    print(f"case otf {name} interp={int(interp)}", flush=True)
    plan = _plan(case, dim=3, F=F, separable_phase=True, interpolation=interp)
    if any(abs(t) > 0 for t in case["rx"].tilt):  # the table only
        assert not plan.has_variant(4)
        plan.close()
        return
    ch = channel_noise(case, F=F, seed=31)
    chd = torch.from_numpy(ch).to(cuda)
    plan.set_variant(2)
    assert plan.variant[0] == 2
    tab = plan.run(chd).cpu().numpy()
    plan.set_variant(4)
    assert plan.variant[0] == 4
    otf = plan.run(chd).cpu().numpy()
    assert _rel(otf, tab) <= 2e-6
    ta, ra, td, rd = oracle_matrices(case)
    ref = O.beamform_fp64(ch, case["fs"], 0.0, ta, ra, td, rd, case["fd"], 3,
                          interp="cubic" if interp == Interpolation.cubic else "linear", w0=O.dasc_wd(case["fd"]),
                          index="fp32")
    assert _rel(otf, ref) <= TOL
    plan.set_variant(2)
    assert np.array_equal(plan.run(chd).cpu().numpy(), tab, equal_nan=True)  # back to the table
    plan.close()


@pytest.mark.parametrize("kind,mla,overlap", [("linear", (1, 1), (0, 0)), ("linear", (2, 1), (1, 0)),
                                               ("sector", (4, 1), (2, 0))])
def test_scanline_transmit_apodization(cuda, kind, mla, overlap):
    """Scanline (MLA) transmit apodization built by the plan (apodization.m:
    265-306): the plan's FP32 transmit weights equal the oracle's
    scanline_apod exactly, and the image equals das_c on the oracle's
    matrices within 1e-5 (RF, das_c's order: bit for bit)."""
    from ustb_amd.enums import Window
    from ustb_amd.uff import Apodization, LinearArray, LinearScan, SectorScan, Wave

    print(f"case scanline {kind} mla={mla} overlap={overlap}", flush=True)
    prb = LinearArray(16, 0.3e-3)
    if kind == "linear":
        scan = LinearScan(np.linspace(-3e-3, 3e-3, 12), np.linspace(4e-3, 20e-3, 40))
        seq = [Wave.plane(0.0) for _ in range(12 // mla[0])]
        case = dict(probe=prb, scan=scan, seq=seq, model=3, scan_kind="linear")
    else:
        scan = SectorScan(np.deg2rad(np.linspace(-20, 20, 16)), np.linspace(3e-3, 25e-3, 30))
        seq = [Wave.spherical(0.0, 0.0, -5e-3) for _ in range(16 // mla[0])]
        case = dict(probe=prb, scan=scan, seq=seq, model=3, scan_kind="sector")
    case.update(rx=Apodization(Window.hanning, f_number=1.5), tx=Apodization(), fs=20e6, fd=0.0, L=500)
    tx = Apodization(Window.scanline, MLA=mla, MLA_overlap=overlap)
    plan = make_plan(prb, scan, seq, n_samples=case["L"], n_frames=1, sampling_frequency=case["fs"],
                     initial_time=0.0, modulation_frequency=0.0, sound_speed=1540.0,
                     receive_apodization=case["rx"], transmit_apodization=tx, dimension=3, dasc_order=True)
    ta, ra, td, rd = [t.cpu().numpy() for t in plan.materialize()]
    nz, nl = scan.shape[0], scan.shape[1]
    want = O.scanline_apod(nz, nl, 1, len(seq), mla, overlap)
    np.testing.assert_array_equal(ta, want)
    _, ora, otd, ord_ = oracle_matrices(case)
    assert np.array_equal(td, otd) and np.array_equal(rd, ord_)
    ch = channel_noise(case, seed=41)
    got = plan.run(torch.from_numpy(ch).to(cuda)).cpu().numpy()
    ref = O.das_c(ch, case["fs"], 0.0, want, ora, otd, ord_, 0.0, 3)
    assert np.array_equal(got, ref, equal_nan=True)
    plan.close()


def test_scanline_wave_count_mismatch_is_refused(cuda):
    from ustb_amd import _lib as L
    from ustb_amd.enums import Window
    from ustb_amd.uff import Apodization, LinearArray, LinearScan, Wave

    with pytest.raises(L.USTBXError, match="does not match with the number of scanlines"):
        make_plan(LinearArray(8, 0.3e-3), LinearScan(np.linspace(-1e-3, 1e-3, 8), np.linspace(5e-3, 9e-3, 16)),
                  [Wave.plane(0.0)] * 3, n_samples=200, n_frames=1, sampling_frequency=20e6, initial_time=0.0,
                  modulation_frequency=0.0, sound_speed=1540.0,
                  transmit_apodization=Apodization(Window.scanline, MLA=(2, 1)))


def test_frame_of_more_than_2_gib(cuda):
    """A frame of 2.5 GiB (5 waves x 64 channels x 2^20 samples): the plan
    runs it on variant 0, whose gathers address each (frame, wave) through a
    buffer resource of its own; RF, dimension both, against das_c on the
    plan's matrices within 1e-5 (and the receive-table variants refuse it)."""
    from ustb_amd.enums import Window
    from ustb_amd.uff import Apodization, LinearArray, LinearScan, Wave

    N, W, Ls, fs = 64, 5, 1 << 20, 20e6
    probe = LinearArray(N, 0.3e-3)
    scan = LinearScan(np.linspace(-3e-3, 3e-3, 12), np.linspace(5e-3, 25e-3, 20))
    seq = [Wave.plane(a) for a in np.deg2rad(np.linspace(-8, 8, W))]
    print(f"case big frame {W * N * Ls * 8 / 2**30:.2f} GiB", flush=True)
    plan = make_plan(probe, scan, seq, n_samples=Ls, n_frames=1, sampling_frequency=fs, initial_time=0.0,
                     modulation_frequency=0.0, sound_speed=1540.0,
                     receive_apodization=Apodization(Window.hanning, f_number=1.5), dimension=3,
                     separable_phase=True)
    assert plan.variant[0] == 0 and not plan.has_variant(2)
    g = torch.Generator(device=cuda).manual_seed(7)
    ch = torch.complex(torch.randn((1, W, N, Ls), device=cuda, generator=g),
                       torch.randn((1, W, N, Ls), device=cuda, generator=g))
    got = plan.run(ch).cpu().numpy()
    ta, ra, td, rd = [t.cpu().numpy() for t in plan.materialize()]
    ref = O.das_c(ch.cpu().numpy(), fs, 0.0, ta, ra, td, rd, 0.0, 3)
    del ch
    assert _rel(got, ref) <= TOL
    plan.close()
