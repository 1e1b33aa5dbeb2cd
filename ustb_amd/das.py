"""DAS: the midprocess.das facade over the MI355X library.

Mirrors the Python reference's ``ustb.midprocess.DAS``
(python/src/ustb/midprocess/das.py:15-261) and the MATLAB class
(+midprocess/das.m:1-319): same properties, ``go()`` returning beamformed
data, and the static ``_beamform`` with the reference's signature.

``go()`` computes delays and apodization on the device (a resident geometry
plan, cached per configuration, including the unified transmit delay model)
and fuses the IQ reference phase (das.m:306-309) into the kernel epilogue.
Scanline (MLA) transmit apodization is built by the plan from the scan's
lines (apodization.m:265-306).
"""
from __future__ import annotations

import hashlib
import time
from collections import OrderedDict

import numpy as np

from .enums import Code, Dimension, Interpolation, SphericalTransmitDelayModel, Window
from .uff import Apodization, BeamformedData, ChannelData, Scan, probe_radius


class DAS:
    """Generalized delay-and-sum beamformer (das.m:1-47; das.py:15-41).

    go() keeps the device plans of its last MAX_PLANS configurations (probe,
    scan, sequence, apodization, sampling, mode, GPU) and releases the least
    recently used one beyond that."""

    MAX_PLANS = 4

    def __init__(self):
        self.channel_data: ChannelData | None = None
        self.scan: Scan | None = None
        self.dimension = Dimension.both
        self.receive_apodization = Apodization()
        self.transmit_apodization = Apodization()
        self.lens_delay = 0.0
        self.pw_margin = 1e-3
        self.blending_power = 0.5
        self.spherical_transmit_delay_model = SphericalTransmitDelayModel.hybrid
        self.code = Code.mex_hip
        self.interpolation = Interpolation.linear
        self.phase = "separable"       # or "dasc": das_c's per-term phase arithmetic
        # the IQ reference phase exp(-1j*2*w0*r/c): its argument in FP64 as das.py
        # forms it ("python") or in single precision as das.m does ("matlab")
        self.reference_phase = "python"
        self.gpu_id = 0
        self.beamformed_data = None
        self.elapsed_time = None
        self._plans: OrderedDict = OrderedDict()  # most recently used last

    # ------------------------------------------------------------------
    @staticmethod
    def _apodization_key(h, apo: Apodization):
        """Every field of an Apodization into the hash, arrays by their bytes
        (a repr would elide the middle of long vectors)."""
        h.update(repr((int(apo.window), apo.f_number, apo.minimum_aperture, apo.maximum_aperture, apo.tilt,
                       apo.grating_lobe_angle, None if apo.origin is None else apo.origin.xyz, apo.MLA,
                       apo.MLA_overlap)).encode())
        v = apo.apodization_vector
        if v is None:
            h.update(b"no-vector")
        else:
            v = np.asarray(v)
            h.update(repr(v.shape).encode())
            h.update(np.ascontiguousarray(v, dtype=np.float64).tobytes())

    def _plan_key(self, n_samples, n_frames):
        cd, sc = self.channel_data, self.scan
        h = hashlib.sha1()
        for a in (cd.probe.x, cd.probe.y, cd.probe.z, sc.x, sc.y, sc.z):
            h.update(np.ascontiguousarray(a, dtype=np.float64).tobytes())
        h.update(repr([(int(w.wavefront), w.source, w.origin, w.delay, w.sound_speed) for w in cd.sequence]).encode())
        self._apodization_key(h, self.receive_apodization)
        self._apodization_key(h, self.transmit_apodization)
        h.update(repr((n_samples, n_frames, cd.sampling_frequency, cd.initial_time, cd.modulation_frequency,
                       cd.sound_speed, int(self.dimension), int(self.interpolation), self.lens_delay,
                       self.pw_margin, self.blending_power, int(self.spherical_transmit_delay_model),
                       self.phase, self.reference_phase, type(sc).__name__, int(self.gpu_id),
                       probe_radius(cd.probe))).encode())
        return h.hexdigest()

    def _needs_hilbert(self) -> bool:
        """das.m:212 / das.py:143: RF data (|w0| < eps, w0 = 2*pi*single(fd))
        is made analytic, whatever its storage type."""
        w0 = np.float32(2.0 * np.pi * np.float32(self.channel_data.modulation_frequency))
        return abs(float(w0)) < float(np.finfo(np.float32).eps)

    def _device_channel_data(self, torch, dev):
        """Channel data as a complex64 (F, W, N, L) device tensor; RF data is
        made analytic first (das.m:211-214)."""
        cd = self.channel_data
        if cd.device_data is not None:
            x = cd.device_data.to(dev)
        else:
            data = np.asarray(cd.data)
            while data.ndim < 4:
                data = data[..., None]
            x = torch.from_numpy(np.ascontiguousarray(data.transpose(3, 2, 1, 0))).to(dev)
        if self._needs_hilbert():
            from .ops import hilbert

            x = hilbert(x.real.float().contiguous() if x.is_complex() else x.float().contiguous())
        elif not x.is_complex():
            raise ValueError("IQ channel data (modulation frequency > 0) must be complex")
        return x.to(torch.complex64).contiguous()

    def go(self) -> BeamformedData:
        """das.m:77-317 / das.py:43-178 with the device backend."""
        import torch

        from .geometry import make_plan

        if self.code != Code.mex_hip:
            raise ValueError("this backend implements code.mex_hip only")
        dev = torch.device("cuda", self.gpu_id)
        cd = self.channel_data
        ch = self._device_channel_data(torch, dev)
        F, W, N, Ls = ch.shape
        key = self._plan_key(Ls, F)
        plan = self._plans.get(key)
        if plan is not None:
            self._plans.move_to_end(key)
        else:
            # keep at most MAX_PLANS configurations' device tables; go()
            # synchronizes after each run, so a released plan has no work queued
            while len(self._plans) >= self.MAX_PLANS:
                self._plans.popitem(last=False)[1].close()
            with torch.cuda.device(dev):
                plan = make_plan(cd.probe, self.scan, cd.sequence, n_samples=Ls, n_frames=F,
                                 sampling_frequency=cd.sampling_frequency, initial_time=cd.initial_time,
                                 modulation_frequency=cd.modulation_frequency, sound_speed=cd.sound_speed,
                                 receive_apodization=self.receive_apodization,
                                 transmit_apodization=self.transmit_apodization, dimension=self.dimension,
                                 interpolation=self.interpolation, separable_phase=(self.phase == "separable"),
                                 reference_phase=self.reference_phase, lens_delay=self.lens_delay,
                                 pw_margin=self.pw_margin,
                                 blending_power=self.blending_power,
                                 tx_model=self.spherical_transmit_delay_model, device=dev)
            self._plans[key] = plan
        torch.cuda.synchronize(dev)
        t0 = time.perf_counter()
        out = plan.run(ch)
        torch.cuda.synchronize(dev)
        self.elapsed_time = time.perf_counter() - t0
        # (F, W', N', P) -> reference layout [pixel, channel, wave, frame]
        data = out.permute(3, 2, 1, 0)
        self.beamformed_data = BeamformedData(scan=self.scan, data=data, probe=cd.probe, sequence=cd.sequence)
        return self.beamformed_data

    # ------------------------------------------------------------------
    @staticmethod
    def _beamform(ch_data, ch_data_time, tx_apodization, rx_apodization, transmit_delay, receive_delay, w0, dim,
                  N_pixels, N_channels, N_waves, N_frames):
        """The reference's DAS._beamform contract (das.py:203-261) on the HIP
        matrix kernel with its arithmetic (USTBX_FLAG_PYTHON_ARITHMETIC) and
        loop order: ch_data (L, N, W, F) complex64; apodization and delays
        (P, N) / (P, W) float32; the time vector (t0 = t[0], dt = t[1] - t[0]
        in float32, das.py:217-218); w0 in rad/s; the Python edge rule
        0 <= floor(d) < L-1. Returns (P, N', W', F) complex64 (NumPy)."""
        import torch

        from . import torch_ops  # noqa: F401  (registers torch.ops.ustbx.das_python)

        dev = torch.device("cuda", 0)
        t = np.asarray(ch_data_time)
        t0 = float(t[0])
        dt = float(t[1] - t[0]) if len(t) > 1 else 1.0
        ch = np.asarray(ch_data)
        while ch.ndim < 4:
            ch = ch[..., None]
        tch = torch.from_numpy(np.ascontiguousarray(ch.transpose(3, 2, 1, 0).astype(np.complex64))).to(dev)
        T = lambda a: torch.from_numpy(np.ascontiguousarray(np.asarray(a, np.float32).T)).to(dev)
        out = torch.ops.ustbx.das_python(tch, t0, dt, T(tx_apodization), T(rx_apodization), T(transmit_delay),
                                         T(receive_delay), float(np.float32(w0)), int(dim))
        return out.permute(3, 2, 1, 0).cpu().numpy()
