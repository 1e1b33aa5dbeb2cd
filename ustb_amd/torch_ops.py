"""The das_c / das_cuda contract as a PyTorch operator, ``torch.ops.ustbx.das``.

Registered with the dispatcher through ``torch.library`` over the C-ABI
(``ustbx_das_run``), so the Python package reaches the HIP kernels the way a
PyTorch-ROCm extension would: by operator name, on device tensors, with a fake
(meta) kernel that gives the output's shape and dtype without running it
(``torch.compile`` tracing, ``FakeTensorMode``).

    torch.ops.ustbx.das(ch, fs, t0, tx_apo, rx_apo, tx_delay, rx_delay, fd, dim,
                        interp=0) -> Tensor

Arguments are the MEX call's, in its order (das.m:251-259; das_c.cpp:46-244):
    ch        (F, W, N, L) complex64, MATLAB's [L, N, W, F] column-major
    fs, t0, fd  sampling frequency [Hz], initial time [s], modulation
              frequency [Hz] (the MEX's single scalars)
    tx_apo, tx_delay  (W, P) float32, MATLAB's [P, W]
    rx_apo, rx_delay  (N, P) float32, MATLAB's [P, N]
    dim       the dimension enum 0..3 (none, receive, transmit, both)
    interp    0 = linear (das_c's), 1 = Keys cubic
Returns (F, W', N', P) complex64, MATLAB's [P, N', W', F], with N' = N for
none / transmit and W' = W for none / receive (das_c.cpp:257-298).

Errors are raised as das_c raises them (its checks and messages, with its
"Toolbox:SRP_SRC:*" identifier in front where das_c gives one). There is only a CUDA (HIP)
kernel: CPU tensors raise, there is no fallback.

    torch.ops.ustbx.das_python(ch, t0, dt, tx_apo, rx_apo, tx_delay, rx_delay,
                               w0, dim) -> Tensor

is the Python reference's ``DAS._beamform`` (python das.py:203-261) on the
same layouts: its NumPy arithmetic term by term (FP32 ``(tau - t0) / dt``,
the exact fraction, FP64 interpolation rounded to complex64, ``apo * v``,
``exp(1j * fl32(w0 * tau))``), its loop order and its edge rule
``0 <= floor(d) < L - 1``; t0 and dt are the first time sample and the
sample period, w0 = 2*pi*fd [rad/s].
"""
from __future__ import annotations

import torch

from . import _lib as L
from . import ops


def _check(ch, fs, t0, tx_apo, rx_apo, tx_delay, rx_delay, fd, dim, interp):
    """das_c's argument checks on tensors (das_c.cpp:53-244), in its order."""
    def fail(code, msg):
        ident = L.lib().ustbx_status_id(code).decode()
        raise RuntimeError(f"{ident}: {msg}" if ident else msg)

    if dim not in (0, 1, 2, 3):
        fail(L.E_ARG, "Unknown sum mode. Available: 0 -> NONE, 1->RX, 2->TX, 3->BOTH")
    if ch.dim() != 4:
        fail(L.E_DIMENSIONS, "Unknown channel data format. Expected (F, W, N, L) (MATLAB [time, channel, wave, "
                             "frame])")
    if ch.dtype == torch.complex128:
        fail(L.E_FLOAT, "The channel data should be single precision")
    if ch.dtype != torch.complex64:
        fail(L.E_ARG, "The channel data should be complex64 (run hilbert on RF data first, das.m:211-214)")
    F, W, N, Ls = ch.shape
    for t, name in ((tx_delay, "transmit delays"), (rx_delay, "receive delays"), (tx_apo, "transmit apodization"),
                    (rx_apo, "receive apodization")):
        if t.dim() != 2:
            fail(L.E_DIMENSIONS, f"Unknown {name} format. Expected 2 dimensions")
        if t.dtype == torch.float64:
            fail(L.E_FLOAT, f"The {name} should be single precision")
        if t.dtype != torch.float32:
            fail(L.E_ARG, f"The {name} should be float32")
    P = tx_delay.shape[1]
    if tx_delay.shape[0] != W or tx_apo.shape[0] != W:
        fail(L.E_DIMENSIONS, "Number of waves in channel data & transmit delays/apodization do not match.")
    if rx_delay.shape[0] != N or rx_apo.shape[0] != N:
        fail(L.E_DIMENSIONS, "Number of channels in channel data & receive delays/apodization do not match.")
    if rx_delay.shape[1] != P or tx_apo.shape[1] != P or rx_apo.shape[1] != P:
        fail(L.E_DIMENSIONS, "Number of pixels in the delays and apodization do not match.")
    if interp not in (L.INTERP_LINEAR, L.INTERP_CUBIC):
        fail(L.E_ARG, f"unknown interpolation {interp}")
    return P, N, W, F


@torch.library.custom_op("ustbx::das", mutates_args=(), device_types="cuda")
def das(ch: torch.Tensor, fs: float, t0: float, tx_apo: torch.Tensor, rx_apo: torch.Tensor,
        tx_delay: torch.Tensor, rx_delay: torch.Tensor, fd: float, dim: int, interp: int = 0) -> torch.Tensor:
    """mex.das_c(ch, fs, t0, tx_apo, rx_apo, tx_delay, rx_delay, fd, dim) on
    the device (das_c.cpp:46-418); see the module docstring."""
    _check(ch, fs, t0, tx_apo, rx_apo, tx_delay, rx_delay, fd, dim, interp)
    return ops.das(ch.contiguous(), float(fs), float(t0), tx_apo.contiguous(), rx_apo.contiguous(),
                   tx_delay.contiguous(), rx_delay.contiguous(), float(fd), int(dim), interp=int(interp))


@das.register_fake
def _das_fake(ch, fs, t0, tx_apo, rx_apo, tx_delay, rx_delay, fd, dim, interp=0):
    P, N, W, F = _check(ch, fs, t0, tx_apo, rx_apo, tx_delay, rx_delay, fd, dim, interp)
    return ch.new_empty(ops.out_shape(dim, P, N, W, F))


@torch.library.custom_op("ustbx::das_python", mutates_args=(), device_types="cuda")
def das_python(ch: torch.Tensor, t0: float, dt: float, tx_apo: torch.Tensor, rx_apo: torch.Tensor,
               tx_delay: torch.Tensor, rx_delay: torch.Tensor, w0: float, dim: int) -> torch.Tensor:
    """DAS._beamform (python das.py:203-261) on the device; see the module
    docstring."""
    _check(ch, 1.0 / dt, t0, tx_apo, rx_apo, tx_delay, rx_delay, w0, dim, L.INTERP_LINEAR)
    return ops.das(ch.contiguous(), float(dt), float(t0), tx_apo.contiguous(), rx_apo.contiguous(),
                   tx_delay.contiguous(), rx_delay.contiguous(), float(w0), int(dim), edge=L.EDGE_PYTHON,
                   python_arithmetic=True, dasc_order=True)


@das_python.register_fake
def _das_python_fake(ch, t0, dt, tx_apo, rx_apo, tx_delay, rx_delay, w0, dim):
    P, N, W, F = _check(ch, 1.0 / dt, t0, tx_apo, rx_apo, tx_delay, rx_delay, w0, dim, L.INTERP_LINEAR)
    return ch.new_empty(ops.out_shape(dim, P, N, W, F))
