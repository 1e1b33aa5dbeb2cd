"""The benchmark acquisitions (seeded, synthetic), at full size or scaled.

C1  plain DAS, CPWC RF, linear array (128 el, 0.3 mm, f0 5.2 MHz, fs 20.8 MHz),
    one plane wave at 0 deg, 2048 samples, 8 point scatterers + noise
    N(0, 1e-3 peak); 256 x 256 scan (x +-19 mm, z 5-50 mm); Tukey-25 rx
    apodization at f# 1.7; linear; dimension both.
C2  plain DAS, CPWC IQ compounded: same array, 75 plane waves -16..16 deg,
    IQ at 5.2 MHz sampled at 10.4 MHz, 1024 samples; 20k speckle scatterers
    N(0,1) in 40 x 50 mm + 10 point targets; 512 (x) x 1024 (z) scan,
    x +-20 mm, z 5-55 mm; cubic; phase rotation; tx/rx Tukey-25 f# 1.7; both.
C3  phased array (64 el, 0.2 mm, f0 2.5 MHz), 31 diverging waves from
    virtual sources 10 mm behind the array steered +-45 deg, IQ at 2.5 MHz
    sampled at 10 MHz, 2048 samples, 32 frames with scatterers moving
    0.1 mm/frame; sector scan 256 beams x 1024 depths (1-120 mm);
    dimension transmit.

``scale`` shrinks pixels/elements/waves/samples for tests.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from . import synth
from .enums import Dimension, Interpolation, Window
from .uff import Apodization, LinearArray, LinearScan, Probe, Scan, SectorScan, Wave


@dataclass
class Acquisition:
    name: str
    probe: Probe
    scan: Scan
    sequence: list
    fs: float
    fd: float
    n_samples: int
    n_frames: int
    rx_apod: Apodization
    tx_apod: Apodization
    dimension: int
    interpolation: int
    c: float = 1540.0
    t0: float = 0.0
    data: np.ndarray | None = None          # (F, W, N, L) complex64
    targets: np.ndarray | None = None       # point targets (K, 3)
    meta: dict = field(default_factory=dict)

    @property
    def P(self) -> int:
        return self.scan.N_pixels

    @property
    def units(self) -> int:
        """pixel x channel x wave x frame samples (the Gsamples unit)."""
        return self.P * self.probe.N_elements * len(self.sequence) * self.n_frames

    def plan_kwargs(self, **over):
        kw = dict(n_samples=self.n_samples, n_frames=self.n_frames, sampling_frequency=self.fs,
                  initial_time=self.t0, modulation_frequency=self.fd, sound_speed=self.c,
                  receive_apodization=self.rx_apod, transmit_apodization=self.tx_apod,
                  dimension=self.dimension, interpolation=self.interpolation)
        kw.update(over)
        return kw


def c1(seed: int = 1, scale: float = 1.0, with_data: bool = True) -> Acquisition:
    rng = np.random.default_rng(seed)
    N = max(8, int(round(128 * scale)))
    nx = nz = max(8, int(round(256 * scale)))
    L = max(64, int(round(2048 * scale)))
    fs, f0, c = 20.8e6, 5.2e6, 1540.0
    prb = LinearArray(N, 0.3e-3)
    scan = LinearScan(np.linspace(-19e-3, 19e-3, nx), np.linspace(5e-3, 50e-3, nz))
    seq = [Wave.plane(0.0, sound_speed=c)]
    zmax = (L / fs) * c / 2 * 0.9
    pts = np.stack([rng.uniform(-15e-3, 15e-3, 8), np.zeros(8), rng.uniform(10e-3, min(45e-3, zmax), 8)], 1)
    acq = Acquisition("C1 CPWC RF 1 PW, 256x256, linear, both", prb, scan, seq, fs, 0.0, L, 1,
                      Apodization(Window.tukey25, f_number=1.7), Apodization(), Dimension.both,
                      Interpolation.linear, targets=pts)
    if with_data:
        bw = synth.pulse_bandwidth_for_cycles(2.5)
        rf = synth.echoes(prb, seq, pts, np.ones(8), fs=fs, n_samples=L, f0=f0, bandwidth=bw, iq=False)
        peak = float(np.max(np.abs(rf)))
        rf = rf + rng.normal(0.0, 1e-3 * peak, rf.shape).astype(np.float32)
        acq.data = synth.analytic(rf)          # the RF branch of das.m:211-214
        acq.meta["rf"] = rf
    return acq


def c2(seed: int = 2, scale: float = 1.0, with_data: bool = True, n_speckle: int = 20000) -> Acquisition:
    rng = np.random.default_rng(seed)
    N = max(8, int(round(128 * scale)))
    W = max(3, int(round(75 * scale)))
    nx = max(8, int(round(512 * scale)))
    nz = max(8, int(round(1024 * scale)))
    L = max(64, int(round(1024 * scale)))
    fs, fd, c = 10.4e6, 5.2e6, 1540.0
    prb = LinearArray(N, 0.3e-3)
    scan = LinearScan(np.linspace(-20e-3, 20e-3, nx), np.linspace(5e-3, 55e-3, nz))
    seq = [Wave.plane(a, sound_speed=c) for a in np.deg2rad(np.linspace(-16.0, 16.0, W))]
    S = max(100, int(round(n_speckle * scale)))
    zmax = min(55e-3, (L / fs) * c / 2 * 0.8)
    spk = np.stack([rng.uniform(-20e-3, 20e-3, S), np.zeros(S), rng.uniform(5e-3, zmax, S)], 1)
    tgt = np.stack([np.linspace(-15e-3, 15e-3, 10), np.zeros(10),
                    np.linspace(10e-3, zmax * 0.95, 10)], 1)
    acq = Acquisition("C2 CPWC IQ 75 PW, 512x1024, cubic, both", prb, scan, seq, fs, fd, L, 1,
                      Apodization(Window.tukey25, f_number=1.7), Apodization(Window.tukey25, f_number=1.7),
                      Dimension.both, Interpolation.cubic, targets=tgt)
    if with_data:
        bw = synth.pulse_bandwidth_for_cycles(2.5)
        amps = np.concatenate([rng.standard_normal(S), np.full(10, 20.0)])
        acq.data = synth.echoes(prb, seq, np.concatenate([spk, tgt]), amps, fs=fs, n_samples=L, f0=fd,
                                bandwidth=bw, iq=True)
    return acq


def c3(seed: int = 3, scale: float = 1.0, with_data: bool = True, n_speckle: int = 5000,
       frames: int | None = None) -> Acquisition:
    rng = np.random.default_rng(seed)
    N = max(8, int(round(64 * scale)))
    W = max(3, int(round(31 * scale)))
    F = frames if frames is not None else max(2, int(round(32 * scale)))
    nb = max(8, int(round(256 * scale)))
    nd = max(8, int(round(1024 * scale)))
    L = max(64, int(round(2048 * scale)))
    fs, fd, c = 10e6, 2.5e6, 1540.0
    prb = LinearArray(N, 0.2e-3)
    ang = np.deg2rad(np.linspace(-45.0, 45.0, W))
    # virtual sources 10 mm behind the array; origin at the source, as
    # wave.fix_origin_from_source sets it (wave.m:87-113)
    seq = []
    for a in ang:
        s = (-10e-3 * math.sin(a), 0.0, -10e-3 * math.cos(a))
        seq.append(Wave.spherical(*s, sound_speed=c, origin=s))
    dmax = min(120e-3, (L / fs) * c / 2 * 0.8)
    scan = SectorScan(np.deg2rad(np.linspace(-45.0, 45.0, nb)), np.linspace(1e-3, dmax, nd))
    S = max(100, int(round(n_speckle * scale)))
    r = rng.uniform(5e-3, dmax * 0.95, S)
    th = rng.uniform(-math.pi / 4, math.pi / 4, S)
    spk = np.stack([r * np.sin(th), np.zeros(S), r * np.cos(th)], 1)
    acq = Acquisition("C3 phased DW 31 waves, sector 256x1024, 32 frames, transmit", prb, scan, seq, fs, fd,
                      L, F, Apodization(), Apodization(), Dimension.transmit, Interpolation.linear)
    if with_data:
        bw = synth.pulse_bandwidth_for_cycles(2.5)
        offs = np.stack([np.zeros(F), np.zeros(F), 0.1e-3 * np.arange(F)], 1)
        acq.data = synth.echoes(prb, seq, spk, rng.standard_normal(S), fs=fs, n_samples=L, f0=fd,
                                bandwidth=bw, iq=True, n_frames=F, frame_offsets=offs)
    return acq


BUILDERS = {"C1": c1, "C2": c2, "C3": c3}
