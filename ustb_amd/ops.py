"""Device operators over the C-ABI, on torch tensors (torch supplies device
memory and the current HIP stream; the arithmetic is the library's).

Layouts follow the reference MEX arrays (MATLAB column-major) held as torch
row-major tensors with reversed dimensions:
    ch        (F, W, N, L) complex64      <- channel_data [L, N, W, F]
    tx_*      (W, P) float32              <- [P, W]
    rx_*      (N, P) float32              <- [P, N]
    output    (F, W', N', P) complex64    <- [P, N', W', F]
"""
from __future__ import annotations

import ctypes

import numpy as np

import torch

from . import _lib as L


def _stream_ptr(device: torch.device) -> int:
    return torch.cuda.current_stream(device).cuda_stream


def out_shape(dim: int, P: int, N: int, W: int, F: int) -> tuple:
    return {L.DIM_NONE: (F, W, N, P), L.DIM_RX: (F, W, 1, P), L.DIM_TX: (F, 1, N, P),
            L.DIM_BOTH: (F, 1, 1, P)}[int(dim)]


def _require_cuda(*ts):
    for t in ts:
        if not (isinstance(t, torch.Tensor) and t.is_cuda):
            raise TypeError("device tensors are required (the HIP path has no CPU fallback)")
        if not t.is_contiguous():
            raise ValueError("tensors must be contiguous")


def das(ch: torch.Tensor, fs: float, t0: float, tx_apo: torch.Tensor, rx_apo: torch.Tensor,
        tx_delay: torch.Tensor, rx_delay: torch.Tensor, fd: float, dim: int, *,
        interp: int = L.INTERP_LINEAR, edge: int = L.EDGE_DASC, dasc_order: bool = False,
        python_arithmetic: bool = False, split: int = 0, out: torch.Tensor | None = None) -> torch.Tensor:
    """mex.das_c / mex.das_cuda contract on device tensors (all four modes).
    python_arithmetic: the Python reference's DAS._beamform arithmetic
    (linear, Python edge rule); then fs is the sample period dt [s] and fd is
    w0 [rad/s]."""
    _require_cuda(ch, tx_apo, rx_apo, tx_delay, rx_delay)
    if ch.dtype != torch.complex64:
        raise TypeError("The channel data should be single precision complex")
    for t in (tx_apo, rx_apo, tx_delay, rx_delay):
        if t.dtype != torch.float32:
            raise TypeError("delays and apodization should be single precision")
    if ch.dim() != 4:
        raise ValueError("channel data must be (F, W, N, L)")
    F, W, N, Ls = ch.shape
    P = tx_delay.shape[-1]
    if tuple(tx_delay.shape) != (W, P) or tuple(tx_apo.shape) != (W, P):
        raise ValueError("Number of waves in channel data & transmit delays/apodization do not match.")
    if tuple(rx_delay.shape) != (N, P) or tuple(rx_apo.shape) != (N, P):
        raise ValueError("Number of channels in channel data & receive delays/apodization do not match.")
    shape = out_shape(dim, P, N, W, F)
    if out is None:
        out = torch.empty(shape, dtype=torch.complex64, device=ch.device)
    elif tuple(out.shape) != shape or out.dtype != torch.complex64 or not out.is_contiguous():
        raise ValueError(f"out must be a contiguous complex64 tensor of shape {shape}")
    flags = (L.FLAG_DASC_ORDER if dasc_order else 0) | (L.FLAG_PYTHON_ARITHMETIC if python_arithmetic else 0)
    d = L.DasDesc(L=Ls, N=N, W=W, F=F, P=P, fs=fs, t0=t0, fd=fd, dim=int(dim), interp=int(interp),
                  edge=int(edge), flags=flags, split=int(split))
    st = L.Status()
    rc = L.lib().ustbx_das_run(ctypes.byref(d), ch.data_ptr(), tx_apo.data_ptr(), rx_apo.data_ptr(),
                               tx_delay.data_ptr(), rx_delay.data_ptr(), out.data_ptr(),
                               _stream_ptr(ch.device), ctypes.byref(st))
    L.check(rc, st)
    return out


def _summed_shape(dim: int, F: int, W: int, N: int, P: int) -> tuple:
    return {L.DIM_BOTH: (F, 1, 1, P), L.DIM_TX: (F, 1, N, P), L.DIM_RX: (F, W, 1, P)}[dim]


def coherent_compounding(x: torch.Tensor, dim: int, window_size: int = 0) -> torch.Tensor:
    """postprocess.coherent_compounding on a device tensor (F, W, N, P)
    complex64; returns (F', W', N', P)."""
    _require_cuda(x)
    if x.dtype != torch.complex64 or x.dim() != 4:
        raise ValueError("beamformed data must be a (F, W, N, P) complex64 tensor")
    x = x.contiguous()
    F, W, N, P = x.shape
    if window_size and dim != L.DIM_RX:
        shape = (F * W - window_size + 1, 1, 1 if dim == L.DIM_BOTH else N, P)
    else:
        shape = _summed_shape(dim, F, W, N, P)
    out = torch.empty(tuple(max(0, s) for s in shape), dtype=torch.complex64, device=x.device)
    st = L.Status()
    L.check(L.lib().ustbx_coherent_compounding(x.data_ptr(), out.data_ptr(), P, N, W, F, int(dim),
                                               int(window_size), _stream_ptr(x.device), ctypes.byref(st)), st)
    return out


def active_elements(tx_apo: torch.Tensor | None, rx_apo: torch.Tensor | None, dim: int, P: int, N: int, W: int,
                    criterium: float = 0.16) -> torch.Tensor:
    """M[p] of the coherence factor from FP32 apodization matrices ([W, P],
    [N, P]; None = ones)."""
    ref = tx_apo if tx_apo is not None else rx_apo
    dev = ref.device if ref is not None else torch.device("cuda", torch.cuda.current_device())
    for t in (tx_apo, rx_apo):
        if t is not None:
            _require_cuda(t)
            if t.dtype != torch.float32 or not t.is_contiguous():
                raise ValueError("apodization must be contiguous float32")
    M = torch.empty(P, dtype=torch.float32, device=dev)
    st = L.Status()
    L.check(L.lib().ustbx_active_elements(tx_apo.data_ptr() if tx_apo is not None else None,
                                          rx_apo.data_ptr() if rx_apo is not None else None, P, N, W, int(dim),
                                          float(criterium), M.data_ptr(), _stream_ptr(dev), ctypes.byref(st)), st)
    return M


def coherence_factor(x: torch.Tensor, dim: int, M: torch.Tensor | None = None, flavour: int = L.CF_MATLAB):
    """Mallart-Fink coherence factor of a device tensor (F, W, N, P)
    complex64: (image, cf), each (F, W', N', P)."""
    _require_cuda(x)
    if x.dtype != torch.complex64 or x.dim() != 4:
        raise ValueError("beamformed data must be a (F, W, N, P) complex64 tensor")
    x = x.contiguous()
    F, W, N, P = x.shape
    shape = _summed_shape(dim, F, W, N, P)
    image = torch.empty(shape, dtype=torch.complex64, device=x.device)
    cf = torch.empty(shape, dtype=torch.float32, device=x.device)
    if M is not None and (M.dtype != torch.float32 or M.numel() != P or not M.is_contiguous()):
        raise ValueError(f"M must be a contiguous float32 tensor of {P} elements")
    st = L.Status()
    L.check(L.lib().ustbx_coherence_factor(x.data_ptr(), image.data_ptr(), cf.data_ptr(),
                                           M.data_ptr() if M is not None else None, P, N, W, F, int(dim),
                                           int(flavour), _stream_ptr(x.device), ctypes.byref(st)), st)
    return image, cf


def demodulate(rf: torch.Tensor, fs: float, t0: float, fmod: float, b, down: int,
               mode: int = L.FIR_CAUSAL) -> torch.Tensor:
    """RF -> IQ (down-mix, FIR with taps b, keep every down-th sample) of a
    float32 device tensor whose last axis is time; returns complex64 with the
    last axis ceil(L / down)."""
    _require_cuda(rf)
    if rf.dtype != torch.float32:
        raise TypeError("RF data must be float32")
    rf = rf.contiguous()
    Ls = rf.shape[-1]
    traces = rf.numel() // max(Ls, 1)
    taps = np.ascontiguousarray(np.asarray(b, dtype=np.float64))
    out = torch.empty(tuple(rf.shape[:-1]) + ((Ls + down - 1) // down,), dtype=torch.complex64, device=rf.device)
    st = L.Status()
    L.check(L.lib().ustbx_demodulate(rf.data_ptr(), out.data_ptr(), Ls, traces, float(fs), float(t0), float(fmod),
                                     taps.ctypes.data, int(taps.size), int(down), int(mode),
                                     _stream_ptr(rf.device), ctypes.byref(st)), st)
    return out


def estimate_modulation_frequency(rf: torch.Tensor, fs: float, return_spectrum: bool = False):
    """The modulation frequency preprocess.fast_demodulation estimates from
    RF data (F, W, N, L) float32 on the device when none is set
    (fast_demodulation.m:57-73): the power spectrum averaged over channels,
    waves and the first min(100, F) frames, the centre of its -6 dB points
    around the maximum, negated. return_spectrum: also (fx, pw) as
    tools.power_spectrum gives them."""
    _require_cuda(rf)
    if rf.dtype != torch.float32 or rf.dim() != 4:
        raise ValueError("RF data must be a (F, W, N, L) float32 tensor")
    F, W, N, Ls = rf.shape
    traces = min(100, F) * W * N
    fmod = ctypes.c_double(0.0)
    pw = np.zeros(Ls, dtype=np.float64)
    st = L.Status()
    L.check(L.lib().ustbx_estimate_modulation_frequency(rf.data_ptr(), Ls, traces, float(fs), ctypes.byref(fmod),
                                                        pw.ctypes.data, _stream_ptr(rf.device), ctypes.byref(st)),
            st)
    if return_spectrum:
        fx = np.linspace(-fs / 2, fs / 2, Ls + 1)[:-1]
        return fmod.value, fx, pw
    return fmod.value


def hilbert(rf: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
    """Analytic signal along the last axis (MATLAB/scipy hilbert) of a
    float32 device tensor; returns complex64 of the same shape."""
    _require_cuda(rf)
    if rf.dtype != torch.float32:
        raise TypeError("RF data must be float32")
    Ls = rf.shape[-1]
    traces = rf.numel() // max(Ls, 1)
    if not rf.is_contiguous():
        raise ValueError("RF data must be contiguous")
    if out is None:
        out = torch.empty(rf.shape, dtype=torch.complex64, device=rf.device)
    elif (not isinstance(out, torch.Tensor) or out.dtype != torch.complex64 or tuple(out.shape) != tuple(rf.shape)
          or not out.is_contiguous() or out.device != rf.device):
        raise ValueError(f"out must be a contiguous complex64 tensor of shape {tuple(rf.shape)} on {rf.device}")
    st = L.Status()
    L.check(L.lib().ustbx_hilbert(rf.data_ptr(), out.data_ptr(), Ls, traces, _stream_ptr(rf.device),
                                  ctypes.byref(st)), st)
    return out


class GeometryPlan:
    """Device-resident geometry (probe, scan, sequence, apodization) for one
    acquisition configuration; delays and apodization are computed on the fly."""

    def __init__(self, geometry: L.Geometry, keepalive=(), device: torch.device | None = None):
        self.device = device or torch.device("cuda", torch.cuda.current_device())
        self._keepalive = keepalive
        self.handle = ctypes.c_void_p()
        st = L.Status()
        with torch.cuda.device(self.device):
            rc = L.lib().ustbx_plan_create(ctypes.byref(geometry), ctypes.byref(self.handle),
                                           ctypes.byref(st))
        L.check(rc, st)
        self.N, self.W, self.L, self.F = geometry.N, geometry.W, geometry.L, geometry.F
        self.dim = geometry.dim
        self.interp = geometry.interp
        self.separable = bool(geometry.flags & L.FLAG_SEPARABLE_PHASE) and not (geometry.flags & L.FLAG_DASC_ORDER)
        self.fd = geometry.fd
        self.P = int(L.lib().ustbx_plan_pixels(self.handle))
        self._keepalive = None  # host arrays were copied into the plan

    @property
    def out_shape(self):
        return out_shape(self.dim, self.P, self.N, self.W, self.F)

    @property
    def device_bytes(self) -> int:
        return int(L.lib().ustbx_plan_device_bytes(self.handle))

    def set_variant(self, variant: int):
        """0: per-lane wave chunks; 1: receive state shared in LDS (dimension
        both, IQ, separable phase, one frame); 2: resident receive table
        (dimension both with the separable or RF phase; dimension
        transmit)."""
        st = L.Status()
        L.check(L.lib().ustbx_plan_set_variant(self.handle, int(variant), ctypes.byref(st)), st)

    def has_variant(self, variant: int) -> bool:
        """Whether the kernel variant applies to this plan (selecting it
        builds its tables); the current selection is kept."""
        cur = self.variant
        st = L.Status()
        rc = L.lib().ustbx_plan_set_variant(self.handle, int(variant), ctypes.byref(st))
        if rc == 0:
            self.set_variant(cur[0])
            self.set_tiles(cur[1], cur[2])
            return True
        if rc != L.E_UNSUPPORTED:
            L.check(rc, st)
        return False

    @property
    def batchable(self) -> bool:
        """Whether batches and the batcher take this plan: dimension-both
        receive-table state."""
        return int(L.lib().ustbx_plan_positions(self.handle)) >= 0

    @property
    def variant(self) -> tuple[int, int, int]:
        """(variant, waves per chunk or split, split) currently selected."""
        wc, sp = ctypes.c_int32(0), ctypes.c_int32(0)
        v = L.lib().ustbx_plan_variant(self.handle, ctypes.byref(wc), ctypes.byref(sp))
        return int(v), int(wc.value), int(sp.value)

    def set_order(self, order: int):
        """Tile visiting order of variants 2/3: 0 scan order, 1 band-major
        per XCD (default)."""
        st = L.Status()
        L.check(L.lib().ustbx_plan_set_order(self.handle, int(order), ctypes.byref(st)), st)

    def set_staging(self, on: bool):
        """Stage sample windows through LDS or gather each term's samples;
        results are the same. Default: on for dimension both (the channels
        the plan marked), off for dimension transmit."""
        st = L.Status()
        L.check(L.lib().ustbx_plan_set_staging(self.handle, int(bool(on)), ctypes.byref(st)), st)

    @property
    def stage_stats(self) -> tuple[int, int]:
        """(staged, listed) (tile, channel) pairs of a receive-table plan
        (dimension both or transmit)."""
        a, b, st = ctypes.c_int64(), ctypes.c_int64(), L.Status()
        with torch.cuda.device(self.device):
            L.check(L.lib().ustbx_plan_stage_stats(self.handle, ctypes.byref(a), ctypes.byref(b),
                                                   ctypes.byref(st)), st)
        return a.value, b.value

    def set_tiles(self, wave_chunk: int = 0, split: int = 0):
        st = L.Status()
        L.check(L.lib().ustbx_plan_set_tiles(self.handle, wave_chunk, split, ctypes.byref(st)), st)

    def run(self, ch: torch.Tensor, out: torch.Tensor | None = None,
            stream: torch.cuda.Stream | None = None) -> torch.Tensor:
        _require_cuda(ch)
        if ch.dtype != torch.complex64 or tuple(ch.shape) != (self.F, self.W, self.N, self.L):
            raise ValueError(f"channel data must be complex64 {(self.F, self.W, self.N, self.L)}")
        if out is None:
            out = torch.empty(self.out_shape, dtype=torch.complex64, device=ch.device)
        elif tuple(out.shape) != self.out_shape or out.dtype != torch.complex64 or not out.is_contiguous():
            raise ValueError(f"out must be contiguous complex64 {self.out_shape}")
        sp = stream.cuda_stream if stream is not None else _stream_ptr(ch.device)
        st = L.Status()
        L.check(L.lib().ustbx_plan_run(self.handle, ch.data_ptr(), out.data_ptr(), sp,
                                       ctypes.byref(st)), st)
        return out

    def run_cf(self, ch: torch.Tensor, M: torch.Tensor | None = None, flavour: int = L.CF_MATLAB,
               image: torch.Tensor | None = None, cf: torch.Tensor | None = None,
               stream: torch.cuda.Stream | None = None):
        """Dimension transmit with the coherence factor over the receive
        channels in the kernel's epilogue: (image [F, P] complex64, cf [F, P]
        float32), equal to run() followed by coherence_factor(dim receive) up
        to the FP32 rounding of the per-group sums. The partial-sum workspace
        is kept with the plan (calls on one plan must not overlap)."""
        _require_cuda(ch)
        if ch.dtype != torch.complex64 or tuple(ch.shape) != (self.F, self.W, self.N, self.L):
            raise ValueError(f"channel data must be complex64 {(self.F, self.W, self.N, self.L)}")
        if image is None:
            image = torch.empty((self.F, self.P), dtype=torch.complex64, device=ch.device)
        if cf is None:
            cf = torch.empty((self.F, self.P), dtype=torch.float32, device=ch.device)
        for t, dt in ((image, torch.complex64), (cf, torch.float32)):
            if tuple(t.shape) != (self.F, self.P) or t.dtype != dt or not t.is_contiguous():
                raise ValueError(f"image / cf must be contiguous ({self.F}, {self.P}) complex64 / float32")
        if M is not None and (M.dtype != torch.float32 or M.numel() != self.P or not M.is_contiguous()):
            raise ValueError(f"M must be a contiguous float32 tensor of {self.P} elements")
        need = int(L.lib().ustbx_plan_cf_workspace_bytes(self.handle))
        ws = getattr(self, "_cf_workspace", None)
        if ws is None or ws.numel() < need or ws.device != ch.device:
            ws = torch.empty(max(need, 1), dtype=torch.uint8, device=ch.device)
            self._cf_workspace = ws
        sp = stream.cuda_stream if stream is not None else _stream_ptr(ch.device)
        st = L.Status()
        L.check(L.lib().ustbx_plan_run_cf(self.handle, ch.data_ptr(), image.data_ptr(), cf.data_ptr(),
                                          M.data_ptr() if M is not None else None, int(flavour), ws.data_ptr(),
                                          ws.numel(), sp, ctypes.byref(st)), st)
        return image, cf

    def active_elements(self, dim: int, criterium: float = 0.16) -> torch.Tensor:
        """M[p] of the coherence factor from the plan's FP32 apodization."""
        M = torch.empty(self.P, dtype=torch.float32, device=self.device)
        st = L.Status()
        L.check(L.lib().ustbx_plan_active_elements(self.handle, int(dim), float(criterium), M.data_ptr(),
                                                   _stream_ptr(self.device), ctypes.byref(st)), st)
        return M

    def active_terms(self) -> int:
        """(pixel, channel, wave, frame) terms with nonzero tx and rx weight."""
        v = ctypes.c_int64(0)
        st = L.Status()
        L.check(L.lib().ustbx_plan_active_terms(self.handle, ctypes.byref(v), ctypes.byref(st)), st)
        return int(v.value)

    def materialize(self):
        """FP32 (tx_apo, rx_apo, tx_delay, rx_delay) as das.m hands them to the MEX."""
        kw = dict(dtype=torch.float32, device=self.device)
        ta = torch.empty((self.W, self.P), **kw)
        td = torch.empty((self.W, self.P), **kw)
        ra = torch.empty((self.N, self.P), **kw)
        rd = torch.empty((self.N, self.P), **kw)
        st = L.Status()
        L.check(L.lib().ustbx_plan_materialize(self.handle, ta.data_ptr(), ra.data_ptr(), td.data_ptr(),
                                               rd.data_ptr(), _stream_ptr(self.device),
                                               ctypes.byref(st)), st)
        return ta, ra, td, rd

    def close(self):
        if self.handle and self.handle.value:
            L.lib().ustbx_plan_destroy(self.handle)
            self.handle = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
