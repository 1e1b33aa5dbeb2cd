"""Enumerations of the reference (dimension.m, code.m, uff.window,
uff.wavefront, spherical_transmit_delay_model.m; python enums.py)."""
from enum import IntEnum


class Dimension(IntEnum):
    """Which axes the beamformer sums (dimension.m:27-32)."""
    none = 0
    receive = 1
    transmit = 2
    both = 3


class Code(IntEnum):
    """Implementation selector (code.m:23-28; python enums.py:36-42), plus the
    MI355X HIP backend as the first free value."""
    matlab = 0
    mex = 1
    python = 2
    matlab_gpu = 3
    mex_gpu = 4
    mex_hip = 5


class Wavefront(IntEnum):
    """uff.wavefront (wavefront.m:16-20)."""
    plane = 0
    spherical = 1
    photoacoustic = 2


class Window(IntEnum):
    """uff.window (window.m:24-37)."""
    none = 0
    boxcar = 1
    rectangular = 1
    flat = 1
    hanning = 2
    hamming = 3
    tukey25 = 4
    tukey50 = 5
    tukey75 = 6
    sta = 7
    scanline = 8
    triangle = 9


class SphericalTransmitDelayModel(IntEnum):
    """spherical_transmit_delay_model.m:32-37."""
    spherical = 1
    unified = 2
    hybrid = 3
    blended = 4


class Interpolation(IntEnum):
    """Sample interpolation of the beamformer: das_c/das_cuda/interp1 are
    linear; cubic (Keys, a = -1/2) has no reference counterpart."""
    linear = 0
    cubic = 1
