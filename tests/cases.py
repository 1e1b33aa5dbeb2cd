"""Seeded problem generators shared by the CPU and GPU tests."""
from __future__ import annotations

import math

import numpy as np

from ustb_amd.enums import Wavefront, Window
from ustb_amd.uff import (Apodization, CurvilinearArray, LinearArray, LinearScan, MatrixArray, SectorScan, Wave,
                          probe_radius)


def matrix_problem(seed=0, P=700, N=12, W=5, F=3, L=96, zero_frac=0.3, edges=True):
    """Random das_c-ABI inputs in sample units (fs = 1, t0 = 0) so that exact
    integer sample positions, including both trace edges, are hit."""
    rng = np.random.default_rng(seed)
    ch = (rng.standard_normal((F, W, N, L)) + 1j * rng.standard_normal((F, W, N, L))).astype(np.complex64)
    td = rng.uniform(-3, L * 0.5, (W, P)).astype(np.float32)
    rd = rng.uniform(-3, L * 0.5 + 3, (N, P)).astype(np.float32)
    if edges:
        # exact integer delays: d = 0, 1, L-2, L-1, L
        k = min(P, 40)
        td[:, :k] = 0.0
        rd[:, :k] = np.resize(np.array([0.0, 1.0, L - 2.0, L - 1.0, L, 0.5, L - 1.5], np.float32), (N, k))
    ta = rng.uniform(-0.2, 1.0, (W, P)).astype(np.float32)
    ra = rng.uniform(-0.2, 1.0, (N, P)).astype(np.float32)
    ta[rng.random((W, P)) < zero_frac] = 0.0
    ra[rng.random((N, P)) < zero_frac] = 0.0
    return dict(ch=ch, fs=1.0, t0=0.0, tx_apo=ta, rx_apo=ra, tx_delay=td, rx_delay=rd)


def wave_dict(w: Wave) -> dict:
    """Flatten a Wave for the oracle's geometry restatement."""
    return dict(wavefront=int(w.wavefront), src=w.source.xyz if not math.isinf(w.source.distance) else (0.0, 0.0, 0.0),
                distance=w.source.distance, azimuth=w.source.azimuth, elevation=w.source.elevation,
                sound_speed=w.sound_speed, delay=w.delay, origin=w.origin.xyz)


def geom_cases():
    """Small acquisitions covering the geometry kinds of das.m / apodization.m."""
    c = 1540.0
    cases = {}
    # CPWC, linear array, linear scan, tx/rx Tukey-25 (C2 in small)
    prb = LinearArray(32, 0.3e-3)
    scan = LinearScan(np.linspace(-5e-3, 5e-3, 24), np.linspace(4e-3, 30e-3, 40))
    seq = [Wave.plane(a, sound_speed=c) for a in np.deg2rad(np.linspace(-16, 16, 7))]
    cases["cpwc_tukey"] = dict(probe=prb, scan=scan, seq=seq, rx=Apodization(Window.tukey25, f_number=1.7),
                               tx=Apodization(Window.tukey25, f_number=1.7), fs=10.4e6, fd=5.2e6, L=512,
                               model=3, scan_kind="linear")
    # CPWC on a scan that starts at depth 0: apodization.m's windows give NaN
    # receive weights on the whole first row (x / 0 after the aperture clamp)
    cases["cpwc_depth0"] = dict(cases["cpwc_tukey"], scan=LinearScan(np.linspace(-5e-3, 5e-3, 24),
                                                                      np.linspace(0.0, 30e-3, 40)),
                                rx=Apodization(Window.hanning, f_number=1.7))
    # RF plane wave 0 deg, hanning rx, no tx apodization (C1 in small)
    cases["pw_rf_hanning"] = dict(probe=LinearArray(24, 0.3e-3), scan=LinearScan(np.linspace(-4e-3, 4e-3, 20),
                                                                                   np.linspace(5e-3, 25e-3, 33)),
                                  seq=[Wave.plane(0.0, sound_speed=c)], rx=Apodization(Window.hanning, f_number=1.7),
                                  tx=Apodization(), fs=20.8e6, fd=0.0, L=600, model=3, scan_kind="linear")
    # phased array, diverging waves, sector scan (C3 in small)
    prb = LinearArray(16, 0.2e-3)
    ang = np.deg2rad(np.linspace(-45, 45, 5))
    # origin at the virtual source, as wave.fix_origin_from_source sets it (wave.m:87-113)
    seq = [Wave.spherical(-10e-3 * math.sin(a), 0.0, -10e-3 * math.cos(a), sound_speed=c,
                          origin=(-10e-3 * math.sin(a), 0.0, -10e-3 * math.cos(a))) for a in ang]
    scan = SectorScan(np.deg2rad(np.linspace(-45, 45, 21)), np.linspace(1e-3, 40e-3, 37))
    cases["dw_sector"] = dict(probe=prb, scan=scan, seq=seq, rx=Apodization(Window.hamming, f_number=1.2),
                              tx=Apodization(Window.tukey50, f_number=1.0), fs=10e6, fd=2.5e6, L=600, model=3,
                              scan_kind="sector")
    # the same on 48 depths x 12 lines of a narrow sector: axes that the
    # transmit kernel's 16-depth x 4-line patch tiles divide (the other cases
    # keep 64-pixel runs), sampled finely enough (1.6 samples per depth step)
    # that a tile's sample windows fit the staged form's 64-sample rows
    cases["dw_sector_patch"] = dict(cases["dw_sector"], scan=SectorScan(np.deg2rad(np.linspace(-5, 5, 12)),
                                                                        np.linspace(20e-3, 26e-3, 48)))
    # CPWC from depth 0 (NaN receive weights on the first row) on 48 depths x
    # 24 lines: patch tiles that take the SAFE body in dimension transmit
    cases["cpwc_depth0_patch"] = dict(cases["cpwc_depth0"], scan=LinearScan(np.linspace(-5e-3, 5e-3, 24),
                                                                              np.linspace(0.0, 30e-3, 48)))
    # RF plane wave on a linear scan of 32 depths x 8 lines (patch tiles)
    cases["pw_rf_patch"] = dict(cases["pw_rf_hanning"], scan=LinearScan(np.linspace(-4e-3, 4e-3, 8),
                                                                         np.linspace(5e-3, 25e-3, 32)))
    # STA: sources at elements, wave.delay = |src|/c (TE_ps_sta_rf.m:33-34), hybrid model
    prb = LinearArray(12, 0.3e-3)
    seq = [Wave.spherical(x, 0.0, 0.0, sound_speed=c, delay=abs(x) / c) for x in prb.x]
    scan = LinearScan(np.linspace(-3e-3, 3e-3, 17), np.linspace(0.5e-3, 20e-3, 29))
    cases["sta_hybrid"] = dict(probe=prb, scan=scan, seq=seq, rx=Apodization(Window.boxcar, f_number=1.0),
                               tx=Apodization(Window.none), fs=20e6, fd=0.0, L=700, model=3, scan_kind="linear")
    # focused transmits: spherical / blended models, triangle window, tilt
    prb = LinearArray(16, 0.3e-3)
    seq = [Wave.spherical(x, 0.0, 15e-3, sound_speed=c, origin=(x, 0.0, 0.0)) for x in np.linspace(-2e-3, 2e-3, 4)]
    scan = LinearScan(np.linspace(-3e-3, 3e-3, 15), np.linspace(5e-3, 25e-3, 31))
    cases["focused_blended"] = dict(probe=prb, scan=scan, seq=seq,
                                    rx=Apodization(Window.triangle, f_number=1.5, tilt=0.1),
                                    tx=Apodization(Window.hanning, f_number=2.0), fs=20e6, fd=5e6, L=700,
                                    model=4, scan_kind="linear")
    cases["focused_spherical"] = dict(cases["focused_blended"], model=1)
    cases["focused_hybrid"] = dict(cases["focused_blended"], model=3)
    cases["focused_unified"] = dict(cases["focused_blended"], model=2)
    # phased array, focused steered transmits, sector scan, unified model
    prb = LinearArray(16, 0.2e-3)
    ang = np.deg2rad(np.linspace(-30, 30, 5))
    seq = [Wave.spherical(20e-3 * math.sin(a), 0.0, 20e-3 * math.cos(a), sound_speed=c) for a in ang]
    scan = SectorScan(np.deg2rad(np.linspace(-35, 35, 21)), np.linspace(2e-3, 40e-3, 37))
    cases["focused_unified_sector"] = dict(probe=prb, scan=scan, seq=seq, rx=Apodization(Window.hanning, f_number=1.4),
                                           tx=Apodization(Window.none, f_number=1.5), fs=10e6, fd=2.5e6, L=600,
                                           model=2, scan_kind="sector")
    # 3-D: matrix array, diverging waves, 3-D linear scan (C4c in small)
    prb = MatrixArray(6, 5, 0.3e-3)
    seq = [Wave.spherical(x, y, -8e-3, sound_speed=c, origin=(x, y, -8e-3)) for x in (-2e-3, 2e-3)
           for y in (-1e-3, 1e-3)]
    seq.append(Wave.spherical(0.0, 0.0, -8e-3, sound_speed=c))  # origin at 0: grating-lobe mask
    scan = LinearScan(np.linspace(-2e-3, 2e-3, 7), np.linspace(5e-3, 15e-3, 19), np.linspace(-1e-3, 1e-3, 5))
    cases["matrix3d_dw"] = dict(probe=prb, scan=scan, seq=seq, rx=Apodization(Window.tukey75, f_number=(1.5, 1.2)),
                                tx=Apodization(Window.none), fs=12e6, fd=3e6, L=400, model=3, scan_kind="linear")
    # curvilinear array (radius 20 mm), diverging waves from behind the arc,
    # sector scan about the arc's centre: the curvilinear receive aperture
    # (azimuth around (0, 0, -radius)) and the transmit incidence of finite
    # sources
    R = 20e-3
    prb = CurvilinearArray(32, 0.3e-3, R)
    ang = np.deg2rad(np.linspace(-20, 20, 4))
    seq = [Wave.spherical(-(5e-3) * math.sin(a), 0.0, -5e-3 * math.cos(a), sound_speed=c,
                          origin=(-(5e-3) * math.sin(a), 0.0, -5e-3 * math.cos(a))) for a in ang]
    scan = SectorScan(np.deg2rad(np.linspace(-25, 25, 19)), np.linspace(R + 2e-3, R + 35e-3, 41), origin=(0.0, 0.0, -R))
    cases["curv_dw"] = dict(probe=prb, scan=scan, seq=seq, rx=Apodization(Window.hanning, f_number=2.5),
                            tx=Apodization(Window.tukey50, f_number=1.0), fs=12e6, fd=3e6, L=700, model=3,
                            scan_kind="sector")
    # curvilinear STA (origin at each source): the transmit axis from the
    # arc's centre to the source; boxcar receive with a minimum aperture
    prb = CurvilinearArray(12, 0.3e-3, 15e-3)
    seq = [Wave.spherical(x, 0.0, z, sound_speed=c, origin=(x, 0.0, z)) for x, z in zip(prb.x, prb.z)]
    scan = LinearScan(np.linspace(-3e-3, 3e-3, 17), np.linspace(1e-3, 20e-3, 29))
    cases["curv_sta"] = dict(probe=prb, scan=scan, seq=seq,
                             rx=Apodization(Window.boxcar, f_number=6.0, minimum_aperture=2e-3),
                             tx=Apodization(Window.hanning, f_number=1.5), fs=20e6, fd=0.0, L=700, model=3,
                             scan_kind="linear")
    return cases


def oracle_matrices(case, lens_delay=0.0, pw_margin=1e-3, blending_power=0.5):
    """FP32 matrices from the oracle's restatement of das.m / apodization.m."""
    from oracle import oracle as O

    prb, scan, seq = case["probe"], case["scan"], case["seq"]
    px, py, pz = scan.x, scan.y, scan.z
    waves = [wave_dict(w) for w in seq]
    rx = case["rx"]
    radius = probe_radius(prb)
    origin = scan.origin if case["scan_kind"] == "sector" else None
    if rx.origin is not None:
        origin = rx.origin.xyz
    elif radius is not None:
        origin = (0.0, 0.0, -radius)
    ra = O.rx_apod(px, py, pz, prb.x, prb.y, prb.z, int(rx.window), rx.f_number, rx.minimum_aperture,
                   rx.maximum_aperture, rx.tilt, origin, curvilinear=radius is not None)
    rd = O.rx_delay(px, py, pz, prb.x, prb.y, prb.z, 1540.0, lens_delay)
    tx = case["tx"]
    ta = O.tx_apod(px, py, pz, waves, int(tx.window), tx.f_number, tx.minimum_aperture, tx.maximum_aperture,
                   tx.tilt, tx.grating_lobe_angle, probe_radius=radius)
    td = O.tx_delay(px, py, pz, waves, case["scan_kind"], case["model"], pw_margin, blending_power,
                    scan_shape=scan.shape[:2], tx_f_number=tx.f_number)
    return ta, ra, td, rd


def channel_noise(case, F=1, seed=1):
    rng = np.random.default_rng(seed)
    N, W, L = case["probe"].N_elements, len(case["seq"]), case["L"]
    return (rng.standard_normal((F, W, N, L)) + 1j * rng.standard_normal((F, W, N, L))).astype(np.complex64)
