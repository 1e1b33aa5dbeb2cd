"""postprocess.coherent_compounding and postprocess.coherence_factor on the
device.

Mirrors the reference classes (+postprocess/coherent_compounding.m,
+postprocess/coherence_factor.m and the Python reference's
python/src/ustb/postprocess/coherence_factor.py): an ``input`` beamformed
data set, the reference's properties, ``go()`` returning beamformed data in the
reference layout [pixel, channel, wave, frame]. The data stays on the device;
host arrays are uploaded once.
"""
from __future__ import annotations

import warnings

import numpy as np
import torch

from . import _lib as L
from . import ops
from .enums import Dimension, Window
from .uff import Apodization, BeamformedData, Wave


def _device_fwnp(data) -> torch.Tensor:
    """[P, N, W, F] (trailing axes may be missing) -> contiguous device
    complex64 (F, W, N, P)."""
    if isinstance(data, torch.Tensor):
        t = data
        if not t.is_cuda:
            t = t.to("cuda")
    else:
        t = torch.from_numpy(np.asarray(data)).to("cuda")
    while t.dim() < 4:
        t = t.unsqueeze(-1)
    if t.dtype != torch.complex64:
        t = t.to(torch.complex64)
    return t.permute(3, 2, 1, 0).contiguous()


def _reference_layout(x: torch.Tensor) -> torch.Tensor:
    return x.permute(3, 2, 1, 0)


class CoherentCompounding:
    """postprocess.coherent_compounding (coherent_compounding.m:1-80)."""

    def __init__(self):
        self.input: BeamformedData | None = None
        self.dimension = Dimension.both
        self.window_size: int | None = None
        self.output: BeamformedData | None = None

    def go(self) -> BeamformedData:
        x = _device_fwnp(self.input.data)
        out = ops.coherent_compounding(x, int(self.dimension), int(self.window_size or 0))
        self.output = BeamformedData(scan=self.input.scan, data=_reference_layout(out), probe=self.input.probe,
                                     sequence=self.input.sequence)
        return self.output


class CoherenceFactor:
    """Mallart-Fink coherence factor (coherence_factor.m:1-151).

    flavour "matlab" (default): the MATLAB class, M = the active elements
    (apodization above active_element_criterium, from the receive and
    transmit apodization when given); "python": the Python reference's
    CoherenceFactor (M = the number of summed elements, + 1e-30)."""

    def __init__(self):
        self.input: BeamformedData | None = None
        self.dimension = Dimension.both
        self.active_element_criterium = 0.16
        self.receive_apodization: Apodization | None = None
        self.transmit_apodization: Apodization | None = None
        self.flavour = "matlab"
        self.CF: BeamformedData | None = None
        self.output: BeamformedData | None = None

    def _dimension(self, N: int, W: int) -> int:
        """coherence_factor.m:62-79."""
        dim = int(self.dimension)
        if dim == L.DIM_RX and N < 2:
            raise ValueError("Not enough channels to compute factor")
        if dim == L.DIM_TX and W < 2:
            raise ValueError("Not enough waves to compute factor")
        if dim == L.DIM_BOTH:
            if N < 2 and W > 1:
                warnings.warn("Not enough channels to compute factor. Changing dimension to dimension.transmit")
                dim = L.DIM_TX
            elif W < 2 and N > 1:
                warnings.warn("Not enough waves to compute factor. Changing dimension to dimension.receive")
                dim = L.DIM_RX
            elif W < 2 and N < 2:
                raise ValueError("Not enough waves and channels to compute factor")
        self.dimension = Dimension(dim)
        return dim

    def _active_elements(self, dim: int, P: int, N: int, W: int) -> torch.Tensor | None:
        """M from the apodization (coherence_factor.m:82-131); None when every
        summed element counts (no apodization given or window none)."""
        rx = self.receive_apodization if dim != L.DIM_TX else None
        tx = self.transmit_apodization if dim != L.DIM_RX else None
        rx_on = rx is not None and rx.window != Window.none
        tx_on = tx is not None and tx.window != Window.none and self.input.sequence is not None
        if not rx_on and not tx_on:
            return None
        if self.input.probe is None:
            raise ValueError("the coherence factor's apodization needs the beamformed data's probe")
        from .geometry import make_plan

        seq = self.input.sequence if self.input.sequence else [Wave.plane(0.0)]
        plan = make_plan(self.input.probe, self.input.scan, seq, n_samples=2, n_frames=1,
                         sampling_frequency=1.0, initial_time=0.0, modulation_frequency=0.0, sound_speed=1540.0,
                         receive_apodization=rx if rx_on else Apodization(),
                         transmit_apodization=tx if tx_on else Apodization(), dimension=3)
        try:
            M = plan.active_elements(dim, self.active_element_criterium)
            torch.cuda.synchronize(M.device)  # the plan's memory is read until then
            return M
        finally:
            plan.close()

    def go(self) -> BeamformedData:
        x = _device_fwnp(self.input.data)
        F, W, N, P = x.shape
        dim = self._dimension(N, W)
        if self.flavour == "python":
            image, cf = ops.coherence_factor(x, dim, None, L.CF_PYTHON)
        else:
            image, cf = ops.coherence_factor(x, dim, self._active_elements(dim, P, N, W), L.CF_MATLAB)
        kw = dict(scan=self.input.scan, probe=self.input.probe, sequence=self.input.sequence)
        self.CF = BeamformedData(data=_reference_layout(cf), **kw)
        self.output = BeamformedData(data=_reference_layout(image), **kw)
        return self.output
