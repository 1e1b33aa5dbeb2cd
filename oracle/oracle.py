"""CPU oracle for the delay-and-sum path.

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() as
the checker, and by bench.py's CPU-baseline leg. The product never imports it.

Contents
- ``das_c``: ctypes front of ``das_c_oracle.c``, the C restatement of the
  reference das_c MEX (+mex/source/das_c.cpp:46-418), FP32.
- ``beamform_fp64``: FP64 restatement of tools.matlab_beamformer
  (+tools/matlab_beamformer.m:1-51) with selectable sample-window rule and a
  cubic option, evaluated on all pixels or a pixel subset.
- ``python_beamform``: restatement of the Python reference's
  DAS._beamform (python/src/ustb/midprocess/das.py:203-261) in NumPy FP32.
- geometry in FP64 rounded to FP32 where das.m applies single():
  ``rx_delay`` (das.m:103-106), ``tx_delay`` (das.m:110-206) with
  ``unified_delay`` (calculate_unified_delay_model.m:14-100),
  ``window`` (apodization.m:172-186), ``rx_apod`` (apodization.m:323-468),
  ``tx_apod`` (apodization.m:249-320, 472-555), ``scanline_apod``
  (apodization.m:265-306) with ``filter2_same``, pixel grids
  (linear_scan.m:64-82, sector_scan.m:73-121), element positions
  (linear_array.m:56-61, matrix_array.m:60-70), ``reference_phase``
  (das.m:306-309).
- postprocess in FP64: ``coherent_compounding`` (coherent_compounding.m:41-73),
  ``coherence_factor`` with ``coherence_dimension`` and ``active_elements``
  (coherence_factor.m:58-151; python/src/ustb/postprocess/coherence_factor.py:19-50).
- RF -> IQ in FP64: ``fast_demodulation`` with ``kaiser_lowpass`` and
  ``fir_causal`` (fast_demodulation.m:58-183, 267-275),
  ``estimate_modulation_frequency`` with ``power_spectrum``
  (fast_demodulation.m:57-73, power_spectrum.m:1-22),
  ``python_fast_demodulation`` with ``fir_zero_phase``
  (python/src/ustb/preprocess/fast_demodulation.py:24-85).

Array layout: MATLAB column-major [L,N,W,F] is held as NumPy C-order
(F, W, N, L); [P,N] as (N, P); [P,W] as (W, P); outputs (F, W', N', P).
"""
from __future__ import annotations

import ctypes
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

DIM_NONE, DIM_RX, DIM_TX, DIM_BOTH = 0, 1, 2, 3
WIN_NONE, WIN_BOXCAR, WIN_HANNING, WIN_HAMMING = 0, 1, 2, 3
WIN_TUKEY25, WIN_TUKEY50, WIN_TUKEY75, WIN_STA, WIN_SCANLINE, WIN_TRIANGLE = 4, 5, 6, 7, 8, 9
WAVE_PLANE, WAVE_SPHERICAL, WAVE_PHOTOACOUSTIC = 0, 1, 2
MODEL_SPHERICAL, MODEL_UNIFIED, MODEL_HYBRID, MODEL_BLENDED = 1, 2, 3, 4
EPS = np.finfo(np.float64).eps

_lib = None


def _load():
    global _lib
    if _lib is None:
        from . import build as _b  # noqa: WPS433 (package-relative recipe)

        path = _b.build(force=False)
        lib = ctypes.CDLL(path)
        fp = ctypes.POINTER(ctypes.c_float)
        lib.oracle_das_c.argtypes = [fp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                     ctypes.c_int, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                     ctypes.c_int, fp, fp, fp, fp, fp, ctypes.c_int]
        lib.oracle_das_c.restype = ctypes.c_int
        _lib = lib
    return _lib


def _fptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def out_shape(dim, P, N, W, F):
    return {DIM_NONE: (F, W, N, P), DIM_RX: (F, W, 1, P), DIM_TX: (F, 1, N, P),
            DIM_BOTH: (F, 1, 1, P)}[dim]


# ---------------------------------------------------------------------------
# das_c restatement (C)
# ---------------------------------------------------------------------------
def das_c(ch, fs, t0, tx_apo, rx_apo, tx_delay, rx_delay, fd, dim, nthreads=None):
    """Run the C restatement of das_c. ch: complex64 (F, W, N, L)."""
    lib = _load()
    ch = np.ascontiguousarray(ch, dtype=np.complex64)
    F, W, N, L = ch.shape
    tx_apo = np.ascontiguousarray(tx_apo, dtype=np.float32)
    rx_apo = np.ascontiguousarray(rx_apo, dtype=np.float32)
    tx_delay = np.ascontiguousarray(tx_delay, dtype=np.float32)
    rx_delay = np.ascontiguousarray(rx_delay, dtype=np.float32)
    P = tx_delay.shape[-1]
    assert tx_apo.shape == (W, P) and tx_delay.shape == (W, P)
    assert rx_apo.shape == (N, P) and rx_delay.shape == (N, P)
    out = np.zeros(out_shape(dim, P, N, W, F), dtype=np.complex64)
    if nthreads is None:
        nthreads = os.cpu_count() or 1
    rc = lib.oracle_das_c(_fptr(ch.view(np.float32)), L, N, W, F, P, np.float32(fs), np.float32(t0),
                          np.float32(fd), int(dim), _fptr(tx_apo), _fptr(rx_apo), _fptr(tx_delay),
                          _fptr(rx_delay), _fptr(out.view(np.float32)), int(nthreads))
    if rc != 0:
        raise ValueError("Unknown sum mode. Available: 0 -> NONE, 1->RX, 2->TX, 3->BOTH")
    return out


# ---------------------------------------------------------------------------
# FP64 matlab_beamformer restatement
# ---------------------------------------------------------------------------
def _keys(b):
    b2 = b * b
    b3 = b2 * b
    return (-0.5 * b3 + b2 - 0.5 * b, 1.5 * b3 - 2.5 * b2 + 1.0,
            -1.5 * b3 + 2.0 * b2 + 0.5 * b, 0.5 * b3 - 0.5 * b2)


def _sample(trace, d, interp, edge):
    """Interpolate trace (F, L) at fractional indices d (P,) -> (F, P)."""
    L = trace.shape[-1]
    i0 = np.floor(d)
    if interp == "linear":
        if edge == "dasc":
            valid = (d >= 1.0) & (d < L - 1)
        elif edge == "python":
            valid = (d >= 0.0) & (d < L - 1)
        else:  # interp1 closed interval [t_0, t_{L-1}]
            valid = (d >= 0.0) & (d <= L - 1)
        i0c = np.clip(np.where(valid, i0, 0), 0, L - 2).astype(np.int64)
        b = np.where(valid, d - i0c, 0.0)
        v = trace[:, i0c] * (1.0 - b) + trace[:, i0c + 1] * b
    else:
        valid = (d >= 1.0) & (d < L - 2)
        i0c = np.where(valid, i0, 1).astype(np.int64)
        b = np.where(valid, d - i0c, 0.0)
        c0, c1, c2, c3 = _keys(b)
        v = (trace[:, i0c - 1] * c0 + trace[:, i0c] * c1 + trace[:, i0c + 1] * c2
             + trace[:, i0c + 2] * c3)
    return v, valid


def dasc_wd(fd):
    """das_c's angular frequency: 2*(float)PI*fd in FP32 (das_c.cpp:33,240-244)."""
    fd32 = np.float32(fd)
    if abs(float(fd32)) <= 1e-6:
        return 0.0
    return float(np.float32(np.float32(2) * np.float32(3.14159265359)) * fd32)


def beamform_fp64(ch, fs, t0, tx_apo, rx_apo, tx_delay, rx_delay, fd, dim,
                  interp="linear", edge="dasc", pixels=None, w0=None, index="fp64"):
    """FP64 generalized DAS (matlab_beamformer.m:19-51) on FP32 inputs.

    w0 = 2*pi*fd unless given (e.g. dasc_wd(fd)); the phase exp(1i*w0*delay)
    is applied when |w0| > eps (matlab_beamformer.m:33-35). ``pixels`` selects
    a subset (indices). ``index="fp32"`` forms the fractional sample index
    with das_c's FP32 arithmetic, d = (float(tx + rx) - t0) * Fs
    (das_c.cpp:372-378), everything else staying FP64. Returns complex128 of
    shape out_shape(dim, len(pixels), N, W, F).
    """
    ch = np.asarray(ch)
    F, W, N, L = ch.shape
    tx_apo = np.asarray(tx_apo, dtype=np.float64)
    rx_apo = np.asarray(rx_apo, dtype=np.float64)
    tx_delay = np.asarray(tx_delay, dtype=np.float64)
    rx_delay = np.asarray(rx_delay, dtype=np.float64)
    if pixels is not None:
        pixels = np.asarray(pixels)
        tx_apo, rx_apo = tx_apo[:, pixels], rx_apo[:, pixels]
        tx_delay, rx_delay = tx_delay[:, pixels], rx_delay[:, pixels]
    P = tx_delay.shape[1]
    if w0 is None:
        w0 = 2.0 * math.pi * float(fd) if abs(float(fd)) > 1e-6 else 0.0
    out = np.zeros(out_shape(dim, P, N, W, F), dtype=np.complex128)
    ch64 = ch.astype(np.complex128)
    for w in range(W):
        if not np.any(tx_apo[w]):
            continue
        for n in range(N):
            if not np.any(rx_apo[n]):
                continue
            apo = rx_apo[n] * tx_apo[w]
            delay = rx_delay[n] + tx_delay[w]
            if index == "fp32":
                c32 = rx_delay[n].astype(np.float32) + tx_delay[w].astype(np.float32)
                d = ((c32 - np.float32(t0)) * np.float32(fs)).astype(np.float64)
            else:
                d = (delay - float(t0)) * float(fs)
            v, valid = _sample(ch64[:, w, n, :], d, interp, edge)
            # out-of-window terms and terms with a zero tx or rx weight are
            # skipped, not multiplied (das_c.cpp:366,371,381); only there does
            # das_c differ from matlab_beamformer (which gives 0*NaN = NaN)
            valid = valid & (tx_apo[w] != 0) & (rx_apo[n] != 0)
            temp = np.where(valid[None, :], v * apo[None, :], 0.0)
            if abs(w0) > EPS:
                temp = np.where(valid[None, :], temp * np.exp(1j * w0 * delay)[None, :], 0.0)
            if dim == DIM_NONE:
                out[:, w, n, :] = temp
            elif dim == DIM_RX:
                out[:, w, 0, :] += temp
            elif dim == DIM_TX:
                out[:, 0, n, :] += temp
            else:
                out[:, 0, 0, :] += temp
    return out


def python_beamform(ch_data, ch_data_time, tx_apo, rx_apo, tx_delay, rx_delay, w0, dim):
    """Restatement of DAS._beamform (python das.py:203-261), NumPy FP32.

    Reference shapes: ch_data (L, N, W, F) complex64; apodization/delays
    (P, N) / (P, W). Returns (P, N', W', F) complex64.
    """
    L, N, W, F = ch_data.shape
    P = tx_apo.shape[0]
    shape = {DIM_NONE: (P, N, W, F), DIM_RX: (P, 1, W, F), DIM_TX: (P, N, 1, F),
             DIM_BOTH: (P, 1, 1, F)}[int(dim)]
    bf = np.zeros(shape, dtype=np.complex64)
    t0 = float(ch_data_time[0])
    dt = float(ch_data_time[1] - ch_data_time[0]) if len(ch_data_time) > 1 else 1.0
    for w in range(W):
        if not np.any(tx_apo[:, w]):
            continue
        for n in range(N):
            if not np.any(rx_apo[:, n]):
                continue
            apo = rx_apo[:, n] * tx_apo[:, w]
            delay = rx_delay[:, n] + tx_delay[:, w]
            sample_idx = (delay - t0) / dt
            idx = np.floor(sample_idx).astype(np.int32)
            frac = sample_idx - idx
            valid = (idx >= 0) & (idx < L - 1)
            for f in range(F):
                trace = ch_data[:, n, w, f]
                iv = np.zeros(P, dtype=trace.dtype)
                iv[valid] = trace[idx[valid]] * (1.0 - frac[valid]) + trace[idx[valid] + 1] * frac[valid]
                temp = apo * iv
                if abs(w0) > np.finfo(np.float32).eps:
                    temp = np.exp(1j * w0 * delay) * temp
                if dim == DIM_NONE:
                    bf[:, n, w, f] = temp
                elif dim == DIM_RX:
                    bf[:, 0, w, f] += temp
                elif dim == DIM_TX:
                    bf[:, n, 0, f] += temp
                else:
                    bf[:, 0, 0, f] += temp
    return bf


# ---------------------------------------------------------------------------
# Geometry restatement (FP64 -> FP32 where das.m applies single())
# ---------------------------------------------------------------------------
def linear_scan_xyz(x_axis, z_axis, y_axis=(0.0,)):
    """[Z, X, Y] = ndgrid(z_axis, x_axis, y_axis) (linear_scan.m:71-81)."""
    z_axis = np.asarray(z_axis, dtype=np.float64)
    x_axis = np.asarray(x_axis, dtype=np.float64)
    y_axis = np.asarray(y_axis, dtype=np.float64)
    Z, X, Y = np.meshgrid(z_axis, x_axis, y_axis, indexing="ij")
    return (X.ravel(order="F"), Y.ravel(order="F"), Z.ravel(order="F"))


def sector_scan_xyz(azimuth_axis, depth_axis, elevation_axis=(0.0,), origin=(0.0, 0.0, 0.0)):
    """ndgrid(depth, az, el) + sph2cart(phi, theta, rho) + origin (sector_scan.m:90-112).

    sph2cart(az=phi, elev=theta, r=rho): z = r*sin(theta) -> X;
    rcoselev = r*cos(theta); x = rcoselev*cos(phi) -> Z; y = rcoselev*sin(phi) -> Y.
    Returns x, y, z, rho.
    """
    rho, th, ph = np.meshgrid(np.asarray(depth_axis, float), np.asarray(azimuth_axis, float),
                              np.asarray(elevation_axis, float), indexing="ij")
    rho = rho.ravel(order="F")
    th = th.ravel(order="F")
    ph = ph.ravel(order="F")
    X = rho * np.sin(th)
    rc = rho * np.cos(th)
    Z = rc * np.cos(ph)
    Y = rc * np.sin(ph)
    return X + origin[0], Y + origin[1], Z + origin[2], rho


def sector_tables(azimuth_axis, elevation_axis=(0.0,)):
    """{sin, cos} per angle, the device tables of a sector scan."""
    az = np.asarray(azimuth_axis, float)
    el = np.asarray(elevation_axis, float)
    return (np.stack([np.sin(az), np.cos(az)], 1).ravel(), np.stack([np.sin(el), np.cos(el)], 1).ravel())


def sector_scan_xyz_tables(azimuth_axis, depth_axis, elevation_axis=(0.0,), origin=(0.0, 0.0, 0.0)):
    """Same grid from precomputed sin/cos tables (the device's arithmetic)."""
    rho, th, ph = np.meshgrid(np.asarray(depth_axis, float), np.arange(len(azimuth_axis)),
                              np.arange(len(elevation_axis)), indexing="ij")
    rho = rho.ravel(order="F")
    ia = th.ravel(order="F")
    ie = ph.ravel(order="F")
    az = np.asarray(azimuth_axis, float)
    el = np.asarray(elevation_axis, float)
    st, ct = np.sin(az)[ia], np.cos(az)[ia]
    sp, cp = np.sin(el)[ie], np.cos(el)[ie]
    rc = rho * ct
    return rho * st + origin[0], rc * sp + origin[1], rc * cp + origin[2], rho


def linear_array_x(N, pitch):
    """x0 = (1:N)*pitch; x0 = x0 - mean(x0) (linear_array.m:56-58)."""
    x0 = np.arange(1, N + 1, dtype=np.float64) * pitch
    return x0 - np.sum(x0) / N


def matrix_array_xy(Nx, Ny, pitch_x, pitch_y):
    """[X, Y] = ndgrid(x0, y0), x fastest (matrix_array.m:60-70)."""
    x0 = np.arange(1, Nx + 1, dtype=np.float64) * pitch_x
    y0 = np.arange(1, Ny + 1, dtype=np.float64) * pitch_y
    x0 = x0 - np.sum(x0) / Nx
    y0 = y0 - np.sum(y0) / Ny
    X, Y = np.meshgrid(x0, y0, indexing="ij")
    return X.ravel(order="F"), Y.ravel(order="F")


def rx_delay(px, py, pz, ex, ey, ez, c, lens_delay=0.0):
    """single(sqrt(xm.^2+ym.^2+zm.^2)/c + lens_delay), xm = probe.x.' - scan.x
    (das.m:103-106). Returns (N, P) float32."""
    xm = ex[None, :] - px[:, None]
    ym = ey[None, :] - py[:, None]
    zm = ez[None, :] - pz[:, None]
    d = np.sqrt((xm * xm + ym * ym) + zm * zm) / c + lens_delay
    return np.ascontiguousarray(d.astype(np.float32).T)


def plane_delay(px, py, pz, az, el):
    """z*cos(az)*cos(el) + x*sin(az)*cos(el) + y*sin(el) (das.m:193)."""
    ca, sa, ce, se = math.cos(az), math.sin(az), math.cos(el), math.sin(el)
    return ((pz * ca) * ce + (px * sa) * ce) + py * se


def unified_delay(tx_delay_col, mask, x_matrix, z_matrix, src_x, src_z, src_distance, dirtest):
    """tools.calculate_unified_delay_model (calculate_unified_delay_model.m:14-100),
    FP64. tx_delay_col: the signed source distance per pixel, (-1)^(z<src.z)
    * |src - x| (das.m:120); mask: the wave's boxcar transmit mask; x_matrix,
    z_matrix: pixel coordinates. All (N_depth, N_lines), depth fastest in the
    scan's pixel order. Returns the (N_depth, N_lines) delay distance, the
    source distance added.

    The branch order and the quick fix are the reference's: the mask's last row
    is set true before the per-line loop (lines 34-37), so the first branch
    (line 50) always holds and the two elseif branches (lines 74-82) and their
    fills (lines 85-90) never run; they are kept as written."""
    tx_delay = np.asarray(tx_delay_col, dtype=np.float64)
    mask = np.array(mask, dtype=bool)
    n_depth, n_lines = tx_delay.shape
    # lines 34-37: quick fix where the beam does not cover the bottom row
    botinds = ~mask[-1, :]
    maxval = np.max(tx_delay[-1, :] * mask[-1, :])
    masked_delays = mask * tx_delay
    mask[-1, :] = True
    masked_delays[-1, botinds] = maxval
    interpolated = np.zeros_like(tx_delay)
    mask_before = np.zeros(tx_delay.shape, dtype=bool)
    mask_after = np.zeros(tx_delay.shape, dtype=bool)
    last_before_idx = None
    first_after_idx = None
    for x in range(n_lines):
        z_idx_focus = int(np.argmin(np.abs(z_matrix[:, x] - src_z)))  # line 49
        if np.sum(mask[z_idx_focus:, x] > 0):
            ray = masked_delays[:, x]
            pos = ray.copy()                      # lines 52-55
            pos[ray > 0] = 0
            pos[pos == 0] = -np.inf
            neg = ray.copy()                      # lines 56-58
            neg[ray < 0] = 0
            neg[neg == 0] = np.inf
            idx_a = int(np.argmax(pos))           # lines 60-61 (first index of the max)
            idx_b = int(np.argmin(neg))
            dx = x_matrix[idx_b, x] - x_matrix[idx_a, x]
            dz = z_matrix[idx_b, x] - z_matrix[idx_a, x]
            nrm = math.sqrt(dx * dx + dz * dz)    # norm(pos_b - pos_a)
            with np.errstate(divide="ignore", invalid="ignore"):
                w1 = np.sqrt((z_matrix[idx_a, x] - z_matrix[:, x]) ** 2) / nrm   # lines 68-71
                w1 = w1 / w1[idx_b]
                w2 = np.sqrt((z_matrix[idx_b, x] - z_matrix[:, x]) ** 2) / nrm
                w2 = w2 / w2[idx_a]
                interpolated[:, x] = w1 * tx_delay[idx_b, x] + w2 * tx_delay[idx_a, x]   # line 72
        elif np.sum(masked_delays[z_idx_focus:, x] == 0) and (x_matrix[z_idx_focus, x] - src_x) * dirtest < 0:
            last_before_idx = x
            mask_before[:, x] = True
        elif np.sum(masked_delays[z_idx_focus:, x] == 0) and (x_matrix[z_idx_focus, x] - src_x) * dirtest > 0:
            if first_after_idx is None:
                first_after_idx = x
            mask_after[:, x] = True
    if mask_before.any():                         # lines 85-90
        interpolated[:, mask_before[0]] = interpolated[:, [last_before_idx + 1]]
    if mask_after.any():
        interpolated[:, mask_after[0]] = interpolated[:, [first_after_idx - 1]]
    interpolated[mask] = tx_delay[mask]           # line 93
    interpolated[np.isinf(interpolated)] = 0      # line 94
    return interpolated + src_distance            # line 96


def tx_delay(px, py, pz, waves, scan_kind="linear", model=MODEL_HYBRID, pw_margin=1e-3,
             blending_power=0.5, scan_shape=None, tx_f_number=(1.0, 1.0), dirtest=1.0):
    """Transmit delay (das.m:110-206). waves: dicts with wavefront, src (xyz),
    distance, azimuth, elevation, sound_speed, delay. Returns (W, P) float32.
    The unified model needs scan_shape = (N_depth, N_lines) of a 2-D linear or
    sector scan and the transmit apodization's f-number (das.m:134-152)."""
    out = np.zeros((len(waves), px.size), dtype=np.float32)
    unified_mask = None
    if model == MODEL_UNIFIED:
        if scan_kind not in ("linear", "sector") or scan_shape is None:
            raise ValueError("Only linear scan and sector scan in 2D is supported for the unified spherical "
                             "transmit delay model.")
        # das.m:136-149: a boxcar transmit apodization with no minimum aperture
        m = tx_apod(px, py, pz, waves, WIN_BOXCAR, f_number=tx_f_number, minimum_aperture=(0.0, 0.0))
        unified_mask = np.nan_to_num(m) != 0
    for w, wv in enumerate(waves):
        sx, sy, sz = wv["src"]
        dist = wv["distance"]
        if wv["wavefront"] == WAVE_PLANE or (wv["wavefront"] == WAVE_SPHERICAL and math.isinf(dist)):
            v = plane_delay(px, py, pz, wv["azimuth"], wv["elevation"])
        elif wv["wavefront"] == WAVE_PHOTOACOUSTIC:
            v = np.zeros_like(px)
        else:
            dx, dy, dz = sx - px, sy - py, sz - pz
            sgn = np.where(pz < sz, -1.0, 1.0)
            d = sgn * np.sqrt((dx * dx + dy * dy) + dz * dz)
            if sz < 0:
                v = d - abs(dist)
            elif model == MODEL_SPHERICAL:
                v = d + dist
            elif model in (MODEL_HYBRID, MODEL_BLENDED):
                if scan_kind == "linear":
                    pd = sgn * np.sqrt(dz * dz) + dist
                elif scan_kind == "sector":
                    pd = plane_delay(px, py, pz, wv["azimuth"], wv["elevation"])
                else:
                    raise ValueError("Only linear scan and sector scan in 2D is supported")
                if model == MODEL_HYBRID:
                    mask = (pz < (sz + pw_margin)) & (pz > (sz - pw_margin))
                    v = np.where(mask, pd, d + dist)
                else:
                    rr = np.sqrt((px * px + py * py) + pz * pz)
                    nd = np.minimum(np.abs(dist - rr) / dist, 1.0)
                    npw = nd ** blending_power
                    v = (d + dist) * npw + pd * (1.0 - npw)
            else:  # MODEL_UNIFIED (das.m:133-152)
                nd, nl = scan_shape
                shape = lambda a: np.reshape(a, (nd, nl), order="F")
                v = unified_delay(shape(d), shape(unified_mask[w]), shape(px), shape(pz), sx, sz, dist,
                                  dirtest).ravel(order="F")
        out[w] = (v / wv["sound_speed"] - wv["delay"]).astype(np.float32)
    return out


def window(win, r):
    """apodization.m:172-186, FP64 (a logical mask times NaN stays NaN)."""
    r = np.asarray(r, dtype=np.float64)
    with np.errstate(invalid="ignore"):
        m = (r <= 0.5).astype(np.float64)
        if win == WIN_BOXCAR:
            return m
        if win == WIN_HANNING:
            return m * (0.5 + 0.5 * np.cos(2 * np.pi * r))
        if win == WIN_HAMMING:
            return m * (0.53836 + 0.46164 * np.cos(2 * np.pi * r))
        if win in (WIN_TUKEY25, WIN_TUKEY50, WIN_TUKEY75):
            roll = {WIN_TUKEY25: 0.25, WIN_TUKEY50: 0.50, WIN_TUKEY75: 0.75}[win]
            th = 0.5 * (1.0 - roll)
            return ((r <= th).astype(float)
                    + (((r > th).astype(float) * (r < 0.5).astype(float)) * 0.5)
                    * (1.0 + np.cos(((2 * np.pi) / roll) * ((r - roll / 2) - 0.5))))
        if win == WIN_TRIANGLE:
            return m * (1.0 - 2.0 * r)
    raise ValueError("Unknown apodization type!")


def _clamp(z, min_af, max_af):
    zd = z.copy()
    m = np.abs(z) <= min_af
    zd[m] = np.sign(zd[m]) * min_af
    m = np.abs(z) >= max_af
    zd[m] = np.sign(zd[m]) * max_af
    return zd


def rx_apod(px, py, pz, ex, ey, ez, window_kind, f_number=(1.0, 1.0), minimum_aperture=(1e-3, 1e-3),
            maximum_aperture=(10.0, 10.0), tilt=(0.0, 0.0), origin=None, curvilinear=False):
    """Receive apodization (apodization.m:323-468).
    Flat probes: origin=None centres the aperture on each pixel (linear
    scan), else on the origin (sector scan or a user origin).
    curvilinear=True (uff.curvilinear_array / curvilinear_matrix_array,
    apodization.m:388-398): origin is the arc's centre (0, 0, -radius) unless
    the user set one; x_dist = origin.z * (pixel azimuth - element azimuth),
    both azimuths atan2(x - origin.x, z - origin.z); y_dist = origin.y - y;
    z_dist = hypot(pixel.x - origin.x, pixel.z - origin.z) - origin.z.
    Returns (N, P) float32."""
    N, P = ex.size, px.size
    if window_kind == WIN_NONE:
        return np.ones((N, P), dtype=np.float32)
    if curvilinear:
        x0, y0, z0 = origin
        el_az = np.arctan2(ex - x0, ez - z0)
        px_az = np.arctan2(px - x0, pz - z0)
        dx, dz = px - x0, pz - z0
        pdist = np.sqrt(dx * dx + dz * dz)
        xd = z0 * (px_az[:, None] - el_az[None, :])
        yd = np.broadcast_to(y0 - ey[None, :], (P, N)).copy()
        zd = np.broadcast_to(pdist[:, None] - z0, (P, N)).copy()
    elif origin is None:
        xd = px[:, None] - ex[None, :]
        yd = py[:, None] - ey[None, :]
        zd = pz[:, None] - ez[None, :]
    else:
        x0, y0, z0 = origin
        dx, dy, dz = px - x0, py - y0, pz - z0
        pdist = np.sqrt((dx * dx + dy * dy) + dz * dz)
        xd = np.broadcast_to(ex[None, :] - x0, (P, N)).copy()
        yd = np.broadcast_to(ey[None, :] - y0, (P, N)).copy()
        zd = np.broadcast_to(pdist[:, None], (P, N)).copy()
    if abs(tilt[0]) > 0:
        ct, st = math.cos(tilt[0]), math.sin(tilt[0])
        xd, zd = xd * ct - zd * st, xd * st + zd * ct
    if abs(tilt[1]) > 0:
        ct, st = math.cos(tilt[1]), math.sin(tilt[1])
        yd, zd = yd * ct - zd * st, yd * st + zd * ct
    zx = _clamp(zd, minimum_aperture[0] * f_number[0], maximum_aperture[0] * f_number[0])
    zy = _clamp(zd, minimum_aperture[1] * f_number[1], maximum_aperture[1] * f_number[1])
    with np.errstate(divide="ignore", invalid="ignore"):
        rt = np.abs(f_number[0] * (xd / zx))
        rp = np.abs(f_number[1] * (yd / zy))
        a = window(window_kind, rt) * window(window_kind, rp)
    return np.ascontiguousarray(a.astype(np.float32).T)


def tx_apod(px, py, pz, waves, window_kind, f_number=(1.0, 1.0), minimum_aperture=(1e-3, 1e-3),
            maximum_aperture=(10.0, 10.0), tilt=(0.0, 0.0), grating_lobe_angle=(math.pi, math.pi),
            probe_radius=None):
    """Transmit apodization (apodization.m:249-320, 472-555). waves: dicts with
    wavefront, src, origin, distance, azimuth, elevation. probe_radius: a
    curvilinear probe's radius, whose STA waves (source at the origin) take
    the axis from (0, 0, -radius) to the source (apodization.m:506-518)
    instead of [0, 0, 1]. Returns (W, P) float32."""
    W, P = len(waves), px.size
    if window_kind == WIN_NONE:
        return np.ones((W, P), dtype=np.float32)
    out = np.zeros((W, P), dtype=np.float32)
    for w, wv in enumerate(waves):
        if wv["wavefront"] == WAVE_PLANE or math.isinf(wv["distance"]):
            rt = abs(f_number[0] * math.tan(wv["azimuth"] - tilt[0]))
            rp = abs(f_number[1] * math.tan(wv["elevation"] - tilt[1]))
            out[w] = np.float32(window(window_kind, rt) * window(window_kind, rp))
            continue
        s = np.asarray(wv["src"], float)
        o = np.asarray(wv["origin"], float)
        SP = (px - s[0], py - s[1], pz - s[2])
        OP = (px - o[0], py - o[1], pz - o[2])
        OS = s - o
        if math.sqrt((OS[0] * OS[0] + OS[1] * OS[1]) + OS[2] * OS[2]) <= EPS:
            if probe_radius is not None:
                OS = np.array([s[0], s[1], s[2] + probe_radius])
            else:
                OS = np.array([0.0, 0.0, 1.0])
        ns = math.sqrt((OS[0] * OS[0] + OS[1] * OS[1]) + OS[2] * OS[2])
        zu = OS / ns
        e1 = np.array([1.0, 0.0, 0.0])
        yu = np.array([zu[1] * e1[2] - zu[2] * e1[1], zu[2] * e1[0] - zu[0] * e1[2],
                       zu[0] * e1[1] - zu[1] * e1[0]])
        xu = np.array([zu[1] * yu[2] - zu[2] * yu[1], zu[2] * yu[0] - zu[0] * yu[2],
                       zu[0] * yu[1] - zu[1] * yu[0]])
        zd = (SP[0] * zu[0] + SP[1] * zu[1]) + SP[2] * zu[2]
        xd = (SP[0] * xu[0] + SP[1] * xu[1]) + SP[2] * xu[2]
        yd = (SP[0] * yu[0] + SP[1] * yu[1]) + SP[2] * yu[2]
        zx = _clamp(zd, minimum_aperture[0] * f_number[0], maximum_aperture[0] * f_number[0])
        zy = _clamp(zd, minimum_aperture[1] * f_number[1], maximum_aperture[1] * f_number[1])
        zx[np.abs(np.arctan2(OP[0], OP[2]) - math.atan2(OS[0], OS[2])) > 0.5 * grating_lobe_angle[0]] = EPS
        zy[np.abs(np.arctan2(OP[1], OP[2]) - math.atan2(OS[1], OS[2])) > 0.5 * grating_lobe_angle[1]] = EPS
        with np.errstate(divide="ignore", invalid="ignore"):
            rt = np.abs(f_number[0] * (xd / zx))
            rp = np.abs(f_number[1] * (yd / zy))
            out[w] = (window(window_kind, rt) * window(window_kind, rp)).astype(np.float32)
    return out


def filter2_same(h, x):
    """MATLAB filter2(h, x, 'same') = conv2(x, rot90(h, 2), 'same'): the full
    2-D convolution with zero padding, central part starting at offset
    floor(size(h)/2) (MATLAB's centring, which differs from scipy's 'same'
    for even kernel sizes)."""
    from scipy.signal import convolve2d

    h = np.asarray(h, float)
    full = convolve2d(np.asarray(x, float), h[::-1, ::-1], mode="full")
    o0, o1 = h.shape[0] // 2, h.shape[1] // 2
    return full[o0:o0 + x.shape[0], o1:o1 + x.shape[1]]


def scanline_apod(n_depth, n_lines, n_lines_y, n_waves, mla=(1, 1), mla_overlap=(0, 0)):
    """Scanline (MLA) transmit apodization (apodization.m:265-306) of a linear
    scan ([N_z, N_x, N_y]) or sector scan ([N_depth, N_az, N_el]): wave
    i + j*N_x/MLA(1) covers the MLA(1) x MLA(2) block of scan lines starting
    at ((i-1) MLA(1), (j-1) MLA(2)) (circshift), smoothed by filter2 with a
    box of (MLA_overlap+1) elements on each axis, and is constant along
    depth. Returns (W, P) float32 in the scan's pixel order (depth fastest)."""
    m1, m2 = int(mla[0]), int(mla[1])
    if n_waves * m1 * m2 != n_lines * n_lines_y:
        raise ValueError("The number of waves in the sequence does not match with the number of scanlines and set "
                         "MLA.")
    A = np.zeros((n_lines, n_lines_y))
    A[:m1, :m2] = 1.0
    k = (int(mla_overlap[0]) + 1, int(mla_overlap[1]) + 1)
    h = np.ones(k) / (k[0] * k[1])
    B = np.zeros((n_waves, n_lines, n_lines_y))
    for i in range(n_lines // m1):
        for j in range(n_lines_y // m2):
            B[i + j * (n_lines // m1)] = filter2_same(h, np.roll(A, (i * m1, j * m2), axis=(0, 1)))
    # reshape(permute(B .* ones(1,1,1,N_z), [4 2 3 1]), [P, W]): depth fastest, then x, then y
    out = np.repeat(B.transpose(0, 2, 1).reshape(n_waves, n_lines * n_lines_y)[:, :, None], n_depth, axis=2)
    return np.ascontiguousarray(out.reshape(n_waves, -1).astype(np.float32))


def reference_phase(r_ref, fd, c, flavour="python"):
    """exp(-1j*2*w0*r_ref/c) (das.m:306-309; das.py:164-173), complex128.

    w0 = 2*pi*single(fd) is a single-precision value in both references
    (das.m:88-89; das.py:54-55 under NumPy 2's scalar promotion), equal to
    das_c's wd. The two references then differ:
    - "python" (das.py:169-171): the complex64 factor -2j*w0 times the FP64
      distance is complex128, so the argument is formed in FP64 and the
      factor rounded to complex64 at the end;
    - "matlab" (das.m:308): -1j*2*w0 is complex single, and MATLAB's products
      and quotients of single with double operands are single, so the
      argument is fl32(fl32(-2*w0 * single(r)) / single(c)) and exp is taken
      in single precision (here: the exact FP64 exp of that FP32 argument,
      rounded to complex64)."""
    w0 = dasc_wd(fd)
    r = np.asarray(r_ref, float)
    if flavour == "python":
        return np.exp(-1j * 2.0 * w0 * r / c)
    if flavour != "matlab":
        raise ValueError(f"unknown flavour {flavour!r}")
    w32, r32, c32 = np.float32(w0), r.astype(np.float32), np.float32(c)
    arg = np.float32(np.float32(np.float32(-2.0) * w32) * r32) / c32
    return np.exp(1j * np.asarray(arg, np.float64)).astype(np.complex64).astype(np.complex128)


def hilbert(x, axis=-1):
    """Analytic signal (das.m:211-214 / das.py:144-145) via scipy.signal.hilbert."""
    from scipy.signal import hilbert as _h

    return _h(x, axis=axis)


# ---------------------------------------------------------------------------
# Postprocess on beamformed data, layout (F, W, N, P) = MATLAB [P, N, W, F]
# ---------------------------------------------------------------------------

def coherent_compounding(data, dim=DIM_BOTH, window_size=None):
    """postprocess.coherent_compounding (coherent_compounding.m:41-73), FP64.

    Without a window: the sum over channels and waves (both), waves
    (transmit) or channels (receive). With one, frames and waves are merged
    (wave fastest) and a moving sum of window_size consecutive entries with
    the incomplete ends discarded (movsum 'Endpoints','discard') runs along
    them: both after summing channels, transmit per channel; receive ignores
    the window. Returns (F', W', N', P)."""
    x = np.asarray(data).astype(np.complex128)
    F, W, N, P = x.shape
    if window_size is None:
        if dim == DIM_BOTH:
            return x.sum(axis=(1, 2), keepdims=True)
        if dim == DIM_TX:
            return x.sum(axis=1, keepdims=True)
        if dim == DIM_RX:
            return x.sum(axis=2, keepdims=True)
        raise ValueError("Unknown dimension mode; check HELP dimension")
    if window_size > W:
        raise ValueError("Cannot have a window size greater that the available number of transmit events.")
    merged = x.reshape(F * W, N, P)                  # reshape to [P, N, 1, W*F] (line 58)
    if dim == DIM_RX:
        return x.sum(axis=2, keepdims=True)
    if dim == DIM_BOTH:
        merged = merged.sum(axis=1, keepdims=True)
    elif dim != DIM_TX:
        raise ValueError("Unknown dimension mode; check HELP dimension")
    T = F * W - window_size + 1
    out = np.stack([merged[t:t + window_size].sum(axis=0) for t in range(T)])
    return out[:, None, :, :]                        # [P, N', 1, T] -> (T, 1, N', P)


def coherence_dimension(dim, N, W):
    """The dimension coherence_factor.m actually uses (lines 62-79): dimension
    both falls back to transmit (one channel) or receive (one wave); errors
    as the reference raises them."""
    if dim == DIM_RX and N < 2:
        raise ValueError("Not enough channels to compute factor")
    if dim == DIM_TX and W < 2:
        raise ValueError("Not enough waves to compute factor")
    if dim == DIM_BOTH:
        if N < 2 and W > 1:
            return DIM_TX
        if W < 2 and N > 1:
            return DIM_RX
        if W < 2 and N < 2:
            raise ValueError("Not enough waves and channels to compute factor")
    return dim


def active_elements(dim, tx_apod, rx_apod, criterium=0.16):
    """M per pixel (coherence_factor.m:82-131): the number of summed elements
    whose apodization exceeds the criterium; both uses the product of the
    transmit and receive apodization. tx_apod (W, P), rx_apod (N, P), FP64."""
    if dim == DIM_BOTH:
        prod = np.asarray(tx_apod, np.float64)[:, None, :] * np.asarray(rx_apod, np.float64)[None, :, :]
        return (prod > criterium).sum(axis=(0, 1)).astype(np.float64)
    if dim == DIM_TX:
        return (np.asarray(tx_apod, np.float64) > criterium).sum(axis=0).astype(np.float64)
    return (np.asarray(rx_apod, np.float64) > criterium).sum(axis=0).astype(np.float64)


def coherence_factor(data, dim=DIM_BOTH, M=None, flavour="matlab"):
    """Mallart-Fink coherence factor. flavour "matlab": postprocess.
    coherence_factor (coherence_factor.m:97-143): CF = |sum|^2 / sum|.|^2 *
    (1/M), NaN -> 0, M per pixel (None: every summed element active).
    flavour "python": the Python reference's CoherenceFactor
    (postprocess/coherence_factor.py:19-50): CF = |sum|^2 / (sum|.|^2 * M +
    1e-30) with M the number of summed elements, NaN -> 0. FP64. Returns
    (image = CF * coherent sum, CF), each (F, W', N', P)."""
    x = np.asarray(data).astype(np.complex128)
    F, W, N, P = x.shape
    axes = {DIM_BOTH: (1, 2), DIM_TX: (1,), DIM_RX: (2,)}[dim]
    count = {DIM_BOTH: N * W, DIM_TX: W, DIM_RX: N}[dim]
    coherent = x.sum(axis=axes, keepdims=True)
    incoherent = (np.abs(x) ** 2).sum(axis=axes, keepdims=True)
    with np.errstate(divide="ignore", invalid="ignore"):
        if flavour == "python":
            cf = np.abs(coherent) ** 2 / (incoherent * float(count) + 1e-30)
        else:
            Mp = np.full(P, float(count)) if M is None else np.asarray(M, np.float64)
            cf = (np.abs(coherent) ** 2 / incoherent) * (1.0 / Mp)[None, None, None, :]
    cf[np.isnan(cf)] = 0
    return cf * coherent, cf


# ---------------------------------------------------------------------------
# RF -> IQ front end: preprocess.fast_demodulation, layout (F, W, N, L)
# ---------------------------------------------------------------------------

def kaiser_lowpass(f_pass, f_stop, fs, dev=1e-3):
    """The low-pass FIR of fast_demodulation.m:267-275: kaiserord([f_pass
    f_stop], [1 0], [dev dev], fs) and fir1(N, Wn, 'low', kaiser(N+1, beta),
    'noscale'), restated from Kaiser's published design formulas (MATLAB
    Signal Processing Toolbox, not in the reference): A = -20 log10(dev);
    beta = 0.1102 (A - 8.7) for A > 50, 0.5842 (A - 21)^0.4 + 0.07886 (A - 21)
    for 21 <= A <= 50, else 0; order N = ceil((A - 7.95) / (2.285 dw)), dw =
    2 pi (f_stop - f_pass) / fs; Wn = (f_pass + f_stop) / fs (normalised to
    Nyquist); taps h[k] = Wn sinc(Wn (k - N/2)) I0(beta sqrt(1 - (2k/N -
    1)^2)) / I0(beta), k = 0..N, unscaled. FP64."""
    A = -20.0 * math.log10(dev)
    if A > 50:
        beta = 0.1102 * (A - 8.7)
    elif A >= 21:
        beta = 0.5842 * (A - 21.0) ** 0.4 + 0.07886 * (A - 21.0)
    else:
        beta = 0.0
    dw = 2.0 * math.pi * (f_stop - f_pass) / fs
    N = int(math.ceil((A - 7.95) / (2.285 * dw)))
    Wn = (f_pass + f_stop) / fs
    k = np.arange(N + 1, dtype=np.float64)
    h = Wn * np.sinc(Wn * (k - N / 2.0))
    w = np.i0(beta * np.sqrt(np.clip(1.0 - (2.0 * k / N - 1.0) ** 2, 0.0, None))) / np.i0(beta)
    return h * w


def fir_causal(b, x):
    """filter(b, 1, x, [], dim) with zero initial state along the last axis:
    y[n] = sum_k b[k] x[n - k], x[n < 0] = 0."""
    x = np.asarray(x)
    y = np.zeros(x.shape, dtype=np.result_type(x, np.float64))
    L = x.shape[-1]
    for k, bk in enumerate(np.asarray(b, dtype=np.float64)):
        if k < L:
            y[..., k:] += bk * x[..., :L - k]
    return y


def fir_zero_phase(b, x):
    """scipy.signal.filtfilt(b, 1, x) along the last axis with its defaults
    (padtype 'odd', padlen 3 * len(b), steady-state initial conditions),
    restated for a FIR: odd extension by padlen samples at both ends, a
    forward pass whose input before the start holds the first extended
    sample, a backward pass likewise from the end, then the padding cut."""
    x = np.asarray(x)
    K = len(b)
    pad = 3 * K
    L = x.shape[-1]
    if L <= pad:
        raise ValueError("The length of the input vector x must be greater than padlen")
    left = 2.0 * x[..., :1] - x[..., pad:0:-1]
    right = 2.0 * x[..., -1:] - x[..., -2:-pad - 2:-1]
    e = np.concatenate([left, x, right], axis=-1)
    Lp = e.shape[-1]
    bb = np.asarray(b, dtype=np.float64)
    y = np.zeros(e.shape, dtype=np.result_type(e, np.float64))
    for k, bk in enumerate(bb):          # forward: e[n < 0] = e[0]
        idx = np.maximum(np.arange(Lp) - k, 0)
        y += bk * e[..., idx]
    z = np.zeros_like(y)
    for k, bk in enumerate(bb):          # backward: y[n >= Lp] = y[Lp - 1]
        idx = np.minimum(np.arange(Lp) + k, Lp - 1)
        z += bk * y[..., idx]
    return z[..., pad:pad + L]


def fast_demodulation(rf, fs, t0, fmod, fdown=None, lowpass=(0.5, 1.0)):
    """preprocess.fast_demodulation (fast_demodulation.m:58-183), FP64: down-mix
    by exp(-1j 2 pi fmod t), t = t0 + (0:L-1)/fs (line 136), the Kaiser
    low-pass with zero initial state (line 149), decimation by Ndown =
    floor(fs / fdown) from the first sample (line 171). fdown defaults to
    2 fmod (line 81). Returns (iq, fs / Ndown, t0 - (len(b) - 1) / 2 / fs, b)
    (lines 90-91, 177-179)."""
    if fdown is None:
        fdown = 2.0 * fmod
    Ndown = int(math.floor(fs / fdown))
    b = kaiser_lowpass(lowpass[0] * fmod, lowpass[1] * fmod, fs)
    delay = (len(b) - 1) / 2.0 / fs
    L = np.asarray(rf).shape[-1]
    t = t0 + np.arange(L, dtype=np.float64) / fs
    x = np.asarray(rf, dtype=np.float64) * np.exp(-1j * 2.0 * np.pi * fmod * t)
    y = fir_causal(b, x)
    return y[..., ::Ndown], fs / Ndown, t0 - delay, b


def power_spectrum(data, fs, normalised=True, n_frames=None):
    """tools.power_spectrum (power_spectrum.m:1-22): the mean over channels,
    waves and the first min(100, F) frames of |fftshift(fft(trace))|^2,
    divided by its maximum; fx = linspace(-fs/2, fs/2, L+1) without its last
    point. data: (F, W, N, L) real."""
    x = np.asarray(data, dtype=np.float64)
    F = x.shape[0]
    n = min(100, F) if n_frames is None else n_frames
    X = np.fft.fftshift(np.fft.fft(x[:n], axis=-1), axes=-1)
    pw = np.mean(np.abs(X) ** 2, axis=(0, 1, 2))
    if normalised:
        pw = pw / np.max(pw)
    L = x.shape[-1]
    fx = np.linspace(-fs / 2, fs / 2, L + 1)[:-1]
    return fx, pw


def _interp1_unsorted(x, v, xq):
    """MATLAB interp1(x, v, xq) for distinct, unsorted sample points x: the
    points are sorted first, then linear interpolation (NaN outside)."""
    x = np.asarray(x, float)
    v = np.asarray(v, float)
    o = np.argsort(x, kind="stable")
    xs, vs = x[o], v[o]
    if np.any(np.diff(xs) == 0):
        raise ValueError("The sample points must be unique.")
    if xq < xs[0] or xq > xs[-1]:
        return float("nan")
    return float(np.interp(xq, xs, vs))


def estimate_modulation_frequency(data, fs):
    """The modulation frequency fast_demodulation estimates when none is set
    (fast_demodulation.m:57-73): the power spectrum's maximum (the first, in
    the negative half), the -6 dB points below it (interp1 over pw(1:ic)) and
    above it up to the middle (pw(ic:round(end/2))), their centre fc, and
    modulation_frequency = -fc."""
    fx, pw = power_spectrum(data, fs)
    ic = int(np.argmax(pw))                      # [dc, ic] = max(pw) (first maximum)
    dc = pw[ic]
    L = pw.size
    mid = int(math.floor(L / 2 + 0.5))           # round(end/2), 1-based
    lo = _interp1_unsorted(pw[:ic + 1], fx[:ic + 1], dc / 2)
    up = _interp1_unsorted(pw[ic:mid], fx[ic:mid], dc / 2)
    return -(lo + up) / 2


def python_fast_demodulation(rf, fs, t0, fc, fdown=None, lowpass=(0.5, 1.0)):
    """The Python reference's FastDemodulation.go
    (python/src/ustb/preprocess/fast_demodulation.py:24-85), FP64: the same
    down-mix, scipy.signal.firwin(numtaps, fc * lowpass[0], fs=fs) with
    numtaps = min(65, L // 3 * 2 - 1) (at least 3, odd), zero-phase
    filtfilt, decimation by floor(fs / fdown) (default floor(fs / (4 fc)));
    t0 is kept. Returns (iq, fs_out, t0, b)."""
    from scipy.signal import firwin

    L = np.asarray(rf).shape[-1]
    t = t0 + np.arange(L, dtype=np.float64) / fs
    x = np.asarray(rf, dtype=np.float64) * np.exp(-1j * 2.0 * np.pi * fc * t)
    numtaps = min(65, L // 3 * 2 - 1)
    if numtaps < 3:
        numtaps = 3
    if numtaps % 2 == 0:
        numtaps += 1
    b = firwin(numtaps, fc * lowpass[0], fs=fs)
    y = fir_zero_phase(b, x)
    dec = max(1, int(math.floor(fs / fdown))) if fdown is not None else max(1, int(math.floor(fs / (4 * fc))))
    if dec > 1:
        return y[..., ::dec], fs / dec, t0, b
    return y, fs, t0, b
