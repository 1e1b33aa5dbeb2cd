"""ustb_amd — MI355X (gfx950) delay-and-sum beamformer, a drop-in for the
midprocess.das path of the UltraSound ToolBox.

The compute path is the native library libustbx.so (HIP kernels behind the
C-ABI of include/ustbx_das.h); Python only describes the acquisition and moves
device pointers.
"""
from .enums import Code, Dimension, Interpolation, SphericalTransmitDelayModel, Wavefront, Window

__all__ = ["Code", "Dimension", "Interpolation", "SphericalTransmitDelayModel", "Wavefront", "Window"]
