"""Acquisition and image descriptors used by the delay-and-sum path.

Minimal counterparts of the reference's UFF classes, limited to what
midprocess.das reads: points (+uff/point.m), waves (+uff/wave.m), probes
(+uff/linear_array.m, +uff/matrix_array.m), scans (+uff/linear_scan.m,
+uff/sector_scan.m and their 3-D forms), apodization settings
(+uff/apodization.m), channel data (+uff/channel_data.m) and beamformed data
(+uff/beamformed_data.m). Array conventions are the reference's: pixel order
is MATLAB's ndgrid order (depth fastest).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from .enums import Wavefront, Window


def _cosd(deg: float) -> float:
    """MATLAB cosd: exact zeros/ones at multiples of 90 degrees."""
    r = math.fmod(deg, 360.0)
    if r % 90.0 == 0.0:
        return [1.0, 0.0, -1.0, 0.0][int(round(r / 90.0)) % 4]
    return math.cos(math.radians(r))


@dataclass
class Point:
    """Point in spherical form (point.m:27-31) with the Cartesian getters of
    point.m:121-144."""

    distance: float = 0.0
    azimuth: float = 0.0
    elevation: float = 0.0

    @classmethod
    def from_xyz(cls, x: float, y: float, z: float) -> "Point":
        """set.xyz (point.m:67-80)."""
        d = math.sqrt(x * x + y * y + z * z)
        az = math.atan2(x, z)
        if d > 0:
            el = (math.pi / 2 * math.copysign(1.0, y)) if math.isinf(y) else math.asin(y / d)
        else:
            el = 0.0
        return cls(d, az, el)

    @property
    def x(self) -> float:
        if self.azimuth == math.pi or self.azimuth == 0:
            return 0.0
        return self.distance * math.sin(self.azimuth) * math.cos(self.elevation)

    @property
    def y(self) -> float:
        if self.elevation == 0:
            return 0.0
        return self.distance * math.sin(self.elevation)

    @property
    def z(self) -> float:
        if abs(self.azimuth) == math.pi / 2 or abs(self.elevation) == math.pi / 2:
            return 0.0
        return self.distance * _cosd(self.azimuth * 180 / math.pi) * _cosd(self.elevation * 180 / math.pi)

    @property
    def xyz(self) -> tuple:
        return (self.x, self.y, self.z)


@dataclass
class Wave:
    """One transmit event (wave.m:43-57)."""

    wavefront: Wavefront = Wavefront.spherical
    source: Point = field(default_factory=Point)
    origin: Point = field(default_factory=Point)
    delay: float = 0.0
    sound_speed: float = 1540.0

    @classmethod
    def plane(cls, azimuth: float, elevation: float = 0.0, sound_speed: float = 1540.0,
              delay: float = 0.0) -> "Wave":
        return cls(Wavefront.plane, Point(math.inf, azimuth, elevation), Point(), delay, sound_speed)

    @classmethod
    def spherical(cls, x: float, y: float, z: float, sound_speed: float = 1540.0, delay: float = 0.0,
                  origin: tuple | None = None) -> "Wave":
        o = Point.from_xyz(*origin) if origin is not None else Point()
        return cls(Wavefront.spherical, Point.from_xyz(x, y, z), o, delay, sound_speed)


class Probe:
    """Element positions (probe.m:42: geometry rows [x y z theta phi w h])."""

    def __init__(self, x, y=None, z=None):
        self.x = np.asarray(x, dtype=np.float64).ravel()
        self.y = np.zeros_like(self.x) if y is None else np.asarray(y, dtype=np.float64).ravel()
        self.z = np.zeros_like(self.x) if z is None else np.asarray(z, dtype=np.float64).ravel()

    @property
    def N_elements(self) -> int:
        return self.x.size


class LinearArray(Probe):
    """x0 = (1:N)*pitch - mean (linear_array.m:56-61)."""

    def __init__(self, N: int, pitch: float):
        x0 = np.arange(1, N + 1, dtype=np.float64) * pitch
        super().__init__(x0 - np.sum(x0) / N)
        self.pitch = pitch


class MatrixArray(Probe):
    """[X, Y] = ndgrid(x0, y0) (matrix_array.m:60-70)."""

    def __init__(self, N_x: int, N_y: int, pitch_x: float, pitch_y: float | None = None):
        pitch_y = pitch_x if pitch_y is None else pitch_y
        x0 = np.arange(1, N_x + 1, dtype=np.float64) * pitch_x
        y0 = np.arange(1, N_y + 1, dtype=np.float64) * pitch_y
        x0 = x0 - np.sum(x0) / N_x
        y0 = y0 - np.sum(y0) / N_y
        X, Y = np.meshgrid(x0, y0, indexing="ij")
        super().__init__(X.ravel(order="F"), Y.ravel(order="F"))
        self.N_x, self.N_y, self.pitch_x, self.pitch_y = N_x, N_y, pitch_x, pitch_y


class CurvilinearArray(Probe):
    """Elements on an arc of radius R, the arc's top at the origin
    (curvilinear_array.m:update): dtheta = 2*asin(pitch/2/R),
    theta = (0:N-1)*dtheta - mean, x = R*sin(theta), z = R*cos(theta) - R.
    Its apodization geometry is curvilinear (apodization.m:377-400, 506-518)."""

    def __init__(self, N: int, pitch: float, radius: float):
        dtheta = 2.0 * math.asin(pitch / 2.0 / radius)
        theta = np.arange(N, dtype=np.float64) * dtheta
        theta = theta - np.sum(theta) / N
        super().__init__(radius * np.sin(theta), None, radius * np.cos(theta) - radius)
        self.N, self.pitch, self.radius, self.theta = N, pitch, radius, theta


class CurvilinearMatrixArray(Probe):
    """A curvilinear array in azimuth (radius_x) repeated along y
    (curvilinear_matrix_array.m:update): x0 = R*sin(theta), y0 = (0:N_y-1)*
    pitch_y - mean, [X, Y] = meshgrid(x0, y0), Z = R*cos(theta) - R, elements
    in X(:) order (y fastest)."""

    def __init__(self, N_x: int, N_y: int, pitch_x: float, pitch_y: float, radius_x: float):
        dtheta = 2.0 * math.asin(pitch_x / 2.0 / radius_x)
        theta = np.arange(N_x, dtype=np.float64) * dtheta
        theta = theta - np.sum(theta) / N_x
        x0 = radius_x * np.sin(theta)
        y0 = np.arange(N_y, dtype=np.float64) * pitch_y
        y0 = y0 - np.sum(y0) / N_y
        X, Y = np.meshgrid(x0, y0)                                   # (N_y, N_x)
        Z = np.ones((N_y, 1)) * (radius_x * np.cos(theta))[None, :] - radius_x
        super().__init__(X.ravel(order="F"), Y.ravel(order="F"), Z.ravel(order="F"))
        self.N_x, self.N_y, self.pitch_x, self.pitch_y, self.radius_x = N_x, N_y, pitch_x, pitch_y, radius_x
        self.radius = radius_x


def probe_radius(probe: Probe) -> float | None:
    """The radius of a curvilinear probe (its apodization geometry), else None."""
    if isinstance(probe, (CurvilinearArray, CurvilinearMatrixArray)):
        return float(probe.radius)
    return None


class Scan:
    kind = "xyz"

    @property
    def N_pixels(self) -> int:
        return self.x.size

    @property
    def reference_distance(self):
        return self.z


class XYZScan(Scan):
    """Explicit pixel list (uff.scan with xyz)."""

    def __init__(self, x, y, z):
        self.x = np.asarray(x, dtype=np.float64).ravel()
        self.y = np.asarray(y, dtype=np.float64).ravel()
        self.z = np.asarray(z, dtype=np.float64).ravel()


class LinearScan(Scan):
    """[Z, X, Y] = ndgrid(z_axis, x_axis, y_axis), z fastest (linear_scan.m:64-82);
    reference_distance = z (linear_scan.m:128-130). Covers linear_scan_3D."""

    kind = "linear"

    def __init__(self, x_axis, z_axis, y_axis=(0.0,)):
        self.x_axis = np.asarray(x_axis, dtype=np.float64).ravel()
        self.z_axis = np.asarray(z_axis, dtype=np.float64).ravel()
        self.y_axis = np.asarray(y_axis, dtype=np.float64).ravel()
        Z, X, Y = np.meshgrid(self.z_axis, self.x_axis, self.y_axis, indexing="ij")
        self.x, self.y, self.z = X.ravel(order="F"), Y.ravel(order="F"), Z.ravel(order="F")

    @property
    def shape(self):
        return (self.z_axis.size, self.x_axis.size, self.y_axis.size)


class SectorScan(Scan):
    """ndgrid(depth, azimuth, elevation) + sph2cart + origin (sector_scan.m:73-121);
    reference_distance = depth (sector_scan.m:167-169). Covers sector_scan_3D."""

    kind = "sector"

    def __init__(self, azimuth_axis, depth_axis, elevation_axis=(0.0,), origin=(0.0, 0.0, 0.0)):
        self.azimuth_axis = np.asarray(azimuth_axis, dtype=np.float64).ravel()
        self.depth_axis = np.asarray(depth_axis, dtype=np.float64).ravel()
        self.elevation_axis = np.asarray(elevation_axis, dtype=np.float64).ravel()
        self.origin = tuple(float(v) for v in origin)
        rho, ia, ie = np.meshgrid(self.depth_axis, np.arange(self.azimuth_axis.size),
                                  np.arange(self.elevation_axis.size), indexing="ij")
        rho, ia, ie = rho.ravel(order="F"), ia.ravel(order="F"), ie.ravel(order="F")
        st, ct = np.sin(self.azimuth_axis)[ia], np.cos(self.azimuth_axis)[ia]
        sp, cp = np.sin(self.elevation_axis)[ie], np.cos(self.elevation_axis)[ie]
        rc = rho * ct
        self.x = rho * st + self.origin[0]
        self.y = rc * sp + self.origin[1]
        self.z = rc * cp + self.origin[2]
        self.rho = rho

    @property
    def reference_distance(self):
        return self.rho

    @property
    def shape(self):
        return (self.depth_axis.size, self.azimuth_axis.size, self.elevation_axis.size)


@dataclass
class Apodization:
    """Apodization settings (apodization.m:40-60)."""

    window: Window = Window.none
    f_number: tuple = (1.0, 1.0)
    minimum_aperture: tuple = (1e-3, 1e-3)
    maximum_aperture: tuple = (10.0, 10.0)
    tilt: tuple = (0.0, 0.0)
    grating_lobe_angle: tuple = (math.pi, math.pi)
    apodization_vector: np.ndarray | None = None
    origin: Point | None = None
    MLA: tuple = (1, 1)             # scan lines per wave (window.scanline)
    MLA_overlap: tuple = (0, 0)     # lines each wave's block is widened by (window.scanline)

    def __post_init__(self):
        def pair(v):
            v = np.atleast_1d(np.asarray(v, dtype=np.float64))
            return (float(v[0]), float(v[0])) if v.size == 1 else (float(v[0]), float(v[1]))

        self.f_number = pair(self.f_number)
        self.minimum_aperture = pair(self.minimum_aperture)
        self.maximum_aperture = pair(self.maximum_aperture)
        t = np.atleast_1d(np.asarray(self.tilt, dtype=np.float64))
        self.tilt = (float(t[0]), 0.0) if t.size == 1 else (float(t[0]), float(t[1]))
        self.grating_lobe_angle = pair(self.grating_lobe_angle)
        ipair = lambda v: tuple(int(x) for x in (np.resize(np.atleast_1d(v), 2)))
        self.MLA = ipair(self.MLA)
        self.MLA_overlap = ipair(self.MLA_overlap)


@dataclass
class ChannelData:
    """Channel data (channel_data.m): ``data`` in the reference layout
    [time, channel, wave, frame] (NumPy (L, N, W, F)), or a device tensor in the
    MEX memory layout (F, W, N, L) via ``device_data``."""

    probe: Probe
    sequence: list
    sampling_frequency: float
    initial_time: float = 0.0
    sound_speed: float = 1540.0
    modulation_frequency: float = 0.0
    data: np.ndarray | None = None
    device_data: object = None

    @property
    def N_samples(self) -> int:
        return self.data.shape[0] if self.data is not None else self.device_data.shape[3]

    @property
    def N_channels(self) -> int:
        return self.probe.N_elements

    @property
    def N_waves(self) -> int:
        return len(self.sequence)

    @property
    def N_frames(self) -> int:
        if self.data is not None:
            return self.data.shape[3] if self.data.ndim > 3 else 1
        return self.device_data.shape[0]

    @property
    def time(self):
        """initial_time + (0:L-1)/fs (das.py:151)."""
        return (np.float32(self.initial_time)
                + np.arange(self.N_samples, dtype=np.float32) / np.float32(self.sampling_frequency))


@dataclass
class BeamformedData:
    """Beamformed data (beamformed_data.m:43): data [pixel, channel, wave, frame]."""

    scan: Scan
    data: object
    probe: object = None      # optional properties (beamformed_data.m:47-54)
    sequence: object = None
