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
C5  serving: requests of one frame each, Poisson arrivals, drawn uniformly
    from linear CPWC 11 waves (C4b on 256^2 pixels), phased DW 31 waves
    (C3's probe and scan, dimension both, rx Tukey-25 f# 1.7) and matrix
    3-D 25 waves (C4c); complex64 IQ.
C4  ragged batch of 64 acquisitions drawn (seeded) from
    (a) STA, 128 el linear array: 128 transmits (sources at the elements)
        x 128 channels x 2048 RF samples at 20.8 MHz (f0 5.2 MHz), beamformed
        from the analytic signal (das.m's hilbert step, applied when the
        data is made);
    (b) CPWC, same array: 11 plane waves -16..16 deg x 128 channels x 1024 IQ
        samples (10.4 MHz, fd 5.2 MHz);
    (c) 3-D diverging waves: 32 x 32 matrix array (0.3 mm) x 25 virtual
        sources on a 5 x 5 grid 10 mm behind it x 512 IQ samples (7.5 MHz,
        fd 3 MHz), onto a 128 x 128 x 256 voxel volume (x, y +-10 mm,
        z 5-35 mm);
    (a) and (b) on square scans of 128^2, 256^2 or 512^2 pixels (x +-19 mm,
    z 5-50 mm). Speckle plus point targets; rx Tukey-25 f# 1.7; tx none for
    (a) and (c), Tukey-25 f# 1.7 for (b); linear interpolation; dimension
    both. Acquisitions of one configuration share its plan; each draws one of
    a few seeded datasets of its kind.
RTB retrospective transmit beamforming of focused transmits, the setup of
    examples/fresnel/linear_array/RTB_linear_array.m: 128-element linear
    array (0.3 mm), 31 focused waves with sources at x in -15..15 mm,
    z = 20 mm and origins below them on the array, RF at 50 MHz (f0 5.2 MHz,
    fractional bandwidth 0.6) beamformed from its analytic signal, 3500
    samples; 512 x 512 scan (x +-19.2 mm, z 0-45 mm); rx Hamming f# 2.5 and
    tx Hamming f# 1.5, both with a 3 mm minimum aperture; linear; dimension
    both; transmit delays by the hybrid model (das.m's default), or another
    spherical_transmit_delay_model.

``scale`` shrinks pixels/elements/waves/samples for tests.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from . import synth
from .enums import Dimension, Interpolation, SphericalTransmitDelayModel, Window
from .uff import Apodization, LinearArray, LinearScan, MatrixArray, Probe, Scan, SectorScan, Wave


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
    tx_model: int = SphericalTransmitDelayModel.hybrid
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
                  dimension=self.dimension, interpolation=self.interpolation, tx_model=int(self.tx_model))
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


@dataclass
class BatchWorkload:
    """Acquisitions of a ragged batch: configs[k] are the distinct probe
    configurations (an Acquisition without data), datasets[k] the seeded
    channel data of configuration k, and acquisitions the (configuration,
    dataset) pair of every acquisition."""
    name: str
    configs: list
    datasets: list
    acquisitions: list

    @property
    def units(self) -> int:
        return sum(self.configs[k].units for k, _ in self.acquisitions)


def _c4_sta(P_side: int, scale: float) -> Acquisition:
    N = max(8, int(round(128 * scale)))
    n = max(8, int(round(P_side * scale)))
    L = max(128, int(round(2048 * scale)))
    c = 1540.0
    prb = LinearArray(N, 0.3e-3)
    seq = [Wave.spherical(x, 0.0, 0.0, sound_speed=c, delay=abs(x) / c) for x in prb.x]
    scan = LinearScan(np.linspace(-19e-3, 19e-3, n), np.linspace(5e-3, 50e-3, n))
    return Acquisition(f"C4a STA RF {N}x{N}, {n}^2", prb, scan, seq, 20.8e6, 0.0, L, 1,
                       Apodization(Window.tukey25, f_number=1.7), Apodization(), Dimension.both,
                       Interpolation.linear, meta={"f0": 5.2e6, "rf": True})


def _c4_cpwc(P_side: int, scale: float) -> Acquisition:
    N = max(8, int(round(128 * scale)))
    n = max(8, int(round(P_side * scale)))
    L = max(128, int(round(1024 * scale)))
    prb = LinearArray(N, 0.3e-3)
    seq = [Wave.plane(a, sound_speed=1540.0) for a in np.deg2rad(np.linspace(-16.0, 16.0, 11))]
    scan = LinearScan(np.linspace(-19e-3, 19e-3, n), np.linspace(5e-3, 50e-3, n))
    return Acquisition(f"C4b CPWC IQ 11 PW, {n}^2", prb, scan, seq, 10.4e6, 5.2e6, L, 1,
                       Apodization(Window.tukey25, f_number=1.7), Apodization(Window.tukey25, f_number=1.7),
                       Dimension.both, Interpolation.linear, meta={"f0": 5.2e6})


def _c4_3d(scale: float) -> Acquisition:
    m = max(4, int(round(32 * scale)))
    nxy = max(8, int(round(128 * scale)))
    nz = max(16, int(round(256 * scale)))
    L = max(128, int(round(512 * scale)))
    c = 1540.0
    prb = MatrixArray(m, m, 0.3e-3)
    seq = []
    for x in np.linspace(-3e-3, 3e-3, 5):
        for y in np.linspace(-3e-3, 3e-3, 5):
            seq.append(Wave.spherical(x, y, -10e-3, sound_speed=c, origin=(x, y, -10e-3)))
    scan = LinearScan(np.linspace(-10e-3, 10e-3, nxy), np.linspace(5e-3, 35e-3, nz), np.linspace(-10e-3, 10e-3, nxy))
    return Acquisition(f"C4c 3-D DW {m}x{m}, {nxy}x{nxy}x{nz}", prb, scan, seq, 7.5e6, 3e6, L, 1,
                       Apodization(Window.tukey25, f_number=(1.7, 1.7)), Apodization(), Dimension.both,
                       Interpolation.linear, meta={"f0": 3e6})


def _c4_data(acq: Acquisition, rng, n_speckle: int) -> np.ndarray:
    """Seeded speckle + 5 point targets in the acquisition's scan volume."""
    sc = acq.scan
    x0, x1 = float(sc.x.min()), float(sc.x.max())
    y0, y1 = float(sc.y.min()), float(sc.y.max())
    z0, z1 = float(sc.z.min()), float(sc.z.max())
    S = n_speckle
    spk = np.stack([rng.uniform(x0, x1, S), rng.uniform(y0, y1, S) if y1 > y0 else np.zeros(S),
                    rng.uniform(z0, z1, S)], 1)
    tgt = np.stack([np.linspace(x0, x1, 7)[1:-1], np.zeros(5), np.linspace(z0, z1, 7)[1:-1]], 1)
    amps = np.concatenate([rng.standard_normal(S), np.full(5, 20.0)])
    bw = synth.pulse_bandwidth_for_cycles(2.5)
    f0 = acq.meta["f0"]
    if acq.meta.get("rf"):
        rf = synth.echoes(acq.probe, acq.sequence, np.concatenate([spk, tgt]), amps, fs=acq.fs,
                          n_samples=acq.n_samples, f0=f0, bandwidth=bw, iq=False)
        return synth.analytic(rf)
    return synth.echoes(acq.probe, acq.sequence, np.concatenate([spk, tgt]), amps, fs=acq.fs,
                        n_samples=acq.n_samples, f0=acq.fd, bandwidth=bw, iq=True)


def c4(seed: int = 4, scale: float = 1.0, with_data: bool = True, n_acq: int = 64, datasets_per_kind: int = 2,
       n_speckle: int = 2000) -> BatchWorkload:
    rng = np.random.default_rng(seed)
    sides = (128, 256, 512)
    kinds = rng.integers(0, 3, n_acq)
    side = rng.integers(0, 3, n_acq)
    data_idx = rng.integers(0, datasets_per_kind, n_acq)
    keys, configs = {}, []
    acqs = []
    for k, sd, d in zip(kinds, side, data_idx):
        key = (int(k), int(sd) if k < 2 else 0)
        if key not in keys:
            keys[key] = len(configs)
            configs.append(_c4_sta(sides[key[1]], scale) if k == 0 else
                           _c4_cpwc(sides[key[1]], scale) if k == 1 else _c4_3d(scale))
        acqs.append((keys[key], int(d)))
    datasets = [[] for _ in configs]
    if with_data:
        # one dataset family per kind: configurations of one kind share data
        # (their probes, sequences and samples agree; only the scan differs)
        by_kind = {}
        for (k, _), ci in keys.items():
            if k not in by_kind:
                by_kind[k] = [_c4_data(configs[ci], np.random.default_rng(seed * 100 + 10 * k + i),
                                       max(50, int(round(n_speckle * scale)))) for i in range(datasets_per_kind)]
            datasets[ci] = by_kind[k]
    return BatchWorkload(f"C4 ragged batch of {n_acq} (STA RF / CPWC IQ / 3-D DW), linear, both",
                         configs, datasets, acqs)


def _c5_dw(scale: float) -> Acquisition:
    """C3's phased array and diverging waves, one frame, dimension both."""
    acq = c3(scale=scale, with_data=False, frames=1)
    acq.name = "C5 phased DW 31 waves, sector 256x1024"
    acq.dimension = Dimension.both
    acq.tx_apod = Apodization()
    acq.rx_apod = Apodization(Window.tukey25, f_number=1.7)
    acq.meta = {"f0": 2.5e6}
    return acq


def c5(seed: int = 5, scale: float = 1.0, with_data: bool = True, datasets_per_config: int = 2,
       n_speckle: int = 2000):
    """C5 serving configurations: linear CPWC 11 waves (C4b, 256^2 pixels),
    phased DW 31 waves (C3's probe and sector scan, one frame, dimension
    both) and matrix 3-D 25 waves (C4c); complex64 IQ from seeded speckle.
    Returns (configs, datasets)."""
    configs = [_c4_cpwc(256, scale), _c5_dw(scale), _c4_3d(scale)]
    datasets = [[] for _ in configs]
    if with_data:
        for k, acq in enumerate(configs):
            datasets[k] = [_c4_data(acq, np.random.default_rng(seed * 100 + 10 * k + i),
                                    max(50, int(round(n_speckle * scale)))) for i in range(datasets_per_config)]
    return configs, datasets


def rtb(seed: int = 6, scale: float = 1.0, with_data: bool = True, n_speckle: int = 5000,
        tx_model: int = SphericalTransmitDelayModel.hybrid) -> Acquisition:
    rng = np.random.default_rng(seed)
    N = max(8, int(round(128 * scale)))
    W = max(3, int(round(31 * scale)))
    n = max(8, int(round(512 * scale)))
    L = max(128, int(round(3500 * scale)))
    fs, f0, c = 50e6, 5.2e6, 1540.0
    prb = LinearArray(N, 0.3e-3)
    seq = [Wave.spherical(x, 0.0, 20e-3, sound_speed=c, origin=(x, 0.0, 0.0))
           for x in np.linspace(-15e-3, 15e-3, W)]
    scan = LinearScan(np.linspace(-19.2e-3, 19.2e-3, n), np.linspace(0.0, 45e-3, n))
    ap = (3e-3, 3e-3)
    name = SphericalTransmitDelayModel(int(tx_model)).name
    acq = Acquisition(f"RTB focused 31 waves RF, 512x512, linear, both, {name} model", prb, scan, seq, fs, 0.0, L,
                      1, Apodization(Window.hamming, f_number=2.5, minimum_aperture=ap),
                      Apodization(Window.hamming, f_number=1.5, minimum_aperture=ap), Dimension.both,
                      Interpolation.linear, c=c, tx_model=int(tx_model), meta={"f0": f0, "rf": True})
    if with_data:
        zmax = min(45e-3, (L / fs) * c / 2 * 0.8)
        S = max(100, int(round(n_speckle * scale)))
        spk = np.stack([rng.uniform(-19e-3, 19e-3, S), np.zeros(S), rng.uniform(2e-3, zmax, S)], 1)
        tgt = np.array([(0.0, 0.0, z) for z in (5e-3, 10e-3, 20e-3, 30e-3, 40e-3)] +
                       [(10e-3, 0.0, 20e-3), (-10e-3, 0.0, 20e-3)])
        tgt = tgt[tgt[:, 2] <= zmax]
        amps = np.concatenate([0.1 * rng.standard_normal(S), np.ones(len(tgt))])
        rf = synth.echoes(prb, seq, np.concatenate([spk, tgt]), amps, fs=fs, n_samples=L, f0=f0, bandwidth=0.6,
                          iq=False)
        acq.data = synth.analytic(rf)          # the RF branch of das.m:211-214
        acq.targets = tgt
    return acq


def c4c(seed: int = 4, scale: float = 1.0, with_data: bool = True, n_speckle: int = 2000) -> Acquisition:
    """C4's 3-D class on its own: one frame of the 32 x 32 matrix array, 25
    diverging waves, 128 x 128 x 256 voxels (the first dataset of C4's 3-D
    kind)."""
    acq = _c4_3d(scale)
    if with_data:
        acq.data = _c4_data(acq, np.random.default_rng(seed * 100 + 20), max(50, int(round(n_speckle * scale))))
    return acq


BUILDERS = {"C1": c1, "C2": c2, "C3": c3, "C4C": c4c, "RTB": rtb}
