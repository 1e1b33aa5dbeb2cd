"""Many acquisitions beamformed by one persistent launch (dimension both).

A Batch lists acquisitions, each beamformed with a GeometryPlan (plans may
repeat: every frame of one probe configuration uses the same plan). The
device keeps per-XCD queues of (acquisition, tile, frame) work items; one
persistent kernel drains them. Images equal each plan's own run with the same
tile shape bit for bit. `replay` runs the launch through a hipGraph captured
on first use; `set_io` points the batch at new buffers without recapturing.
"""
from __future__ import annotations

import ctypes

import torch

from . import _lib as L
from .ops import GeometryPlan, _require_cuda, _stream_ptr


class Batch:
    def __init__(self, plans: list[GeometryPlan]):
        if not plans:
            raise ValueError("a batch needs at least one plan")
        for p in plans:
            if p.variant[0] < 2:
                p.set_variant(2)  # variant 4 (the receive state on the fly) is kept
        self.plans = list(plans)
        self.device = plans[0].device
        arr = (ctypes.c_void_p * len(plans))(*[p.handle.value for p in plans])
        self.handle = ctypes.c_void_p()
        st = L.Status()
        with torch.cuda.device(self.device):
            rc = L.lib().ustbx_batch_create(arr, len(plans), ctypes.byref(self.handle), ctypes.byref(st))
        L.check(rc, st)
        self._io = None

    @property
    def shape(self) -> dict:
        wc, ws, blocks, items = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int64()
        L.lib().ustbx_batch_shape(self.handle, ctypes.byref(wc), ctypes.byref(ws), ctypes.byref(blocks),
                                  ctypes.byref(items))
        return {"wave_chunk": wc.value, "split": ws.value, "blocks": blocks.value, "items": items.value}

    def set_tiles(self, wave_chunk: int = 0, split: int = 0):
        st = L.Status()
        L.check(L.lib().ustbx_batch_set_tiles(self.handle, wave_chunk, split, ctypes.byref(st)), st)

    def set_io(self, chs: list[torch.Tensor], outs: list[torch.Tensor] | None = None) -> list[torch.Tensor]:
        """Binds device buffers (kept referenced by the batch); returns the outputs."""
        if len(chs) != len(self.plans):
            raise ValueError("one channel-data tensor per acquisition")
        if outs is None:
            outs = [torch.empty(p.out_shape, dtype=torch.complex64, device=c.device) for p, c in zip(self.plans, chs)]
        for p, c, o in zip(self.plans, chs, outs):
            _require_cuda(c)
            if c.dtype != torch.complex64 or tuple(c.shape) != (p.F, p.W, p.N, p.L) or not c.is_contiguous():
                raise ValueError(f"channel data must be contiguous complex64 {(p.F, p.W, p.N, p.L)}")
            if tuple(o.shape) != p.out_shape or o.dtype != torch.complex64 or not o.is_contiguous():
                raise ValueError(f"out must be contiguous complex64 {p.out_shape}")
        n = len(chs)
        ca = (ctypes.c_void_p * n)(*[c.data_ptr() for c in chs])
        oa = (ctypes.c_void_p * n)(*[o.data_ptr() for o in outs])
        st = L.Status()
        L.check(L.lib().ustbx_batch_set_io(self.handle, ca, oa, ctypes.byref(st)), st)
        self._io = (list(chs), list(outs))
        return outs

    def launch(self, stream: torch.cuda.Stream | None = None):
        sp = stream.cuda_stream if stream is not None else _stream_ptr(self.device)
        st = L.Status()
        L.check(L.lib().ustbx_batch_launch(self.handle, sp, ctypes.byref(st)), st)

    def replay(self, stream: torch.cuda.Stream | None = None):
        sp = stream.cuda_stream if stream is not None else _stream_ptr(self.device)
        st = L.Status()
        L.check(L.lib().ustbx_batch_replay(self.handle, sp, ctypes.byref(st)), st)

    def close(self):
        if self.handle:
            L.lib().ustbx_batch_destroy(self.handle)
            self.handle = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
