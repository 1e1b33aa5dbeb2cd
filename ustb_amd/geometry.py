"""Conversion of acquisition descriptors into the C-ABI geometry descriptor
(include/ustbx_das.h: ustbx_geometry) and creation of device-resident plans."""
from __future__ import annotations

import ctypes
import math

import numpy as np

from . import _lib as L
from .enums import Dimension, Interpolation, SphericalTransmitDelayModel, Wavefront, Window
from .uff import Apodization, LinearScan, Probe, Scan, SectorScan, Wave, probe_radius


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(L.P_double)


def _apod_struct(a: Apodization, keep: list) -> L.Apod:
    s = L.Apod()
    s.window = int(a.window)
    s.f_number[:] = a.f_number
    s.minimum_aperture[:] = a.minimum_aperture
    s.maximum_aperture[:] = a.maximum_aperture
    s.tilt[:] = a.tilt
    s.grating_lobe_angle[:] = a.grating_lobe_angle
    s.mla[:] = a.MLA
    s.mla_overlap[:] = a.MLA_overlap
    if a.origin is not None:
        s.has_origin = 1
        s.origin[:] = a.origin.xyz
    if a.apodization_vector is not None:
        v = np.ascontiguousarray(np.asarray(a.apodization_vector, dtype=np.float32).ravel())
        keep.append(v)
        s.vector = v.ctypes.data_as(L.P_float)
    return s


def wave_struct(w: Wave) -> L.Wave:
    s = L.Wave()
    s.wavefront = int(w.wavefront)
    src = w.source
    s.src_distance = src.distance
    s.azimuth, s.elevation = src.azimuth, src.elevation
    if math.isinf(src.distance):
        s.src[:] = (0.0, 0.0, 0.0)
    else:
        s.src[:] = src.xyz
    s.cos_az, s.sin_az = math.cos(src.azimuth), math.sin(src.azimuth)
    s.cos_el, s.sin_el = math.cos(src.elevation), math.sin(src.elevation)
    s.sound_speed = w.sound_speed
    s.delay = w.delay
    s.origin[:] = w.origin.xyz
    return s


def build_geometry(probe: Probe, scan: Scan, sequence: list, *, n_samples: int, n_frames: int,
                   sampling_frequency: float, initial_time: float, modulation_frequency: float,
                   sound_speed: float, receive_apodization: Apodization | None = None,
                   transmit_apodization: Apodization | None = None,
                   dimension: int = Dimension.both, interpolation: int = Interpolation.linear,
                   edge: int = L.EDGE_DASC, separable_phase: bool = True,
                   reference_phase: bool | str = False, dasc_order: bool = False,
                   lens_delay: float = 0.0, pw_margin: float = 1e-3, blending_power: float = 0.5,
                   tx_model: int = SphericalTransmitDelayModel.hybrid,
                   tx_delay_matrix=None, rx_apod_matrix=None, tx_apod_matrix=None):
    """Returns (Geometry struct, keepalive list). Matrices, when given, are
    device tensors ([W, P] / [N, P] float32) that stay owned by the caller.

    reference_phase: False, True or "python" (the factor exp(-1j*2*w0*r/c)
    with its argument in FP64, python das.py:169-171), or "matlab" (the
    argument in single precision, das.m:308)."""
    if reference_phase not in (False, True, "python", "matlab"):
        raise ValueError(f"reference_phase must be a bool, 'python' or 'matlab', not {reference_phase!r}")
    keep: list = []
    g = L.Geometry()
    g.N = probe.N_elements
    g.W = len(sequence)
    g.L = int(n_samples)
    g.F = int(n_frames)
    for name, arr in (("elem_x", probe.x), ("elem_y", probe.y), ("elem_z", probe.z)):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        keep.append(a)
        setattr(g, name, _dptr(a))
    sc = L.Scan()
    if isinstance(scan, LinearScan):
        sc.kind = L.SCAN_LINEAR
        sc.n0, sc.n1, sc.n2 = scan.z_axis.size, scan.x_axis.size, scan.y_axis.size
        axes = (scan.z_axis, scan.x_axis, scan.y_axis)
        sc.ref_kind = 0
    elif isinstance(scan, SectorScan):
        sc.kind = L.SCAN_SECTOR
        sc.n0, sc.n1, sc.n2 = scan.depth_axis.size, scan.azimuth_axis.size, scan.elevation_axis.size
        az, el = scan.azimuth_axis, scan.elevation_axis
        axes = (scan.depth_axis, np.stack([np.sin(az), np.cos(az)], 1).ravel(),
                np.stack([np.sin(el), np.cos(el)], 1).ravel())
        sc.origin[:] = scan.origin
        sc.ref_kind = 1
    else:
        sc.kind = L.SCAN_XYZ
        sc.n0, sc.n1, sc.n2 = scan.N_pixels, 1, 1
        axes = (scan.x, scan.y, scan.z)
        sc.ref_kind = 0
    for name, arr in zip(("a0", "a1", "a2"), axes):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        keep.append(a)
        setattr(sc, name, _dptr(a))
    g.scan = sc
    waves = (L.Wave * g.W)(*[wave_struct(w) for w in sequence])
    keep.append(waves)
    g.waves = ctypes.cast(waves, ctypes.POINTER(L.Wave))
    g.rx_apod = _apod_struct(receive_apodization or Apodization(), keep)
    g.tx_apod = _apod_struct(transmit_apodization or Apodization(), keep)
    if rx_apod_matrix is not None:
        g.rx_apod.matrix = ctypes.cast(ctypes.c_void_p(rx_apod_matrix.data_ptr()), L.P_float)
        keep.append(rx_apod_matrix)
    if tx_apod_matrix is not None:
        g.tx_apod.matrix = ctypes.cast(ctypes.c_void_p(tx_apod_matrix.data_ptr()), L.P_float)
        keep.append(tx_apod_matrix)
    if tx_delay_matrix is not None:
        g.tx_delay_matrix = ctypes.cast(ctypes.c_void_p(tx_delay_matrix.data_ptr()), L.P_float)
        keep.append(tx_delay_matrix)
    g.sound_speed = sound_speed
    g.lens_delay = lens_delay
    g.pw_margin = pw_margin
    g.blending_power = blending_power
    g.tx_model = int(tx_model)
    g.fs = sampling_frequency
    g.t0 = initial_time
    g.fd = modulation_frequency
    g.dim = int(dimension)
    g.interp = int(interpolation)
    g.edge = int(edge)
    radius = probe_radius(probe)
    if radius is not None:
        g.probe_kind = L.PROBE_CURVILINEAR
        g.probe_radius = radius
    g.flags = ((L.FLAG_SEPARABLE_PHASE if separable_phase else 0)
               | (L.FLAG_REF_PHASE if reference_phase else 0)
               | (L.FLAG_REF_PHASE_SINGLE if reference_phase == "matlab" else 0)
               | (L.FLAG_DASC_ORDER if dasc_order else 0))
    return g, keep


def make_plan(*args, device=None, **kw):
    """Build a device-resident GeometryPlan (see build_geometry)."""
    from .ops import GeometryPlan

    g, keep = build_geometry(*args, **kw)
    return GeometryPlan(g, keepalive=keep, device=device)
