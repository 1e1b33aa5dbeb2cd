"""A UFF file (written per write_object.m's conventions) read back and
beamformed on the device gives the image of the same channel data held in
memory, bit for bit; its frames feed the serving loop's payloads."""
import numpy as np
import pytest
import torch

from ustb_amd import uff_io as U
from ustb_amd.das import DAS
from ustb_amd.enums import Window
from ustb_amd.uff import Apodization, ChannelData, LinearArray, LinearScan, Wave

pytestmark = pytest.mark.gpu


def test_uff_file_beamforms_like_memory(cuda, tmp_path):
    try:
        U.h5()
    except OSError:
        pytest.skip("no HDF5 library")
    rng = np.random.default_rng(9)
    F, W, N, L = 3, 4, 32, 400
    prb = LinearArray(N, 0.3e-3)
    seq = [Wave.plane(a) for a in np.linspace(-0.1, 0.1, W)]
    x = (rng.standard_normal((F, W, N, L)) + 1j * rng.standard_normal((F, W, N, L))).astype(np.complex64)
    cd = ChannelData(probe=prb, sequence=seq, sampling_frequency=10.4e6, initial_time=0.0,
                     modulation_frequency=5.2e6, data=x.transpose(3, 2, 1, 0))
    path = str(tmp_path / "acq.uff")
    U.write_object(path, cd, "channel_data")
    back = U.read_channel_data(path)

    def image(c):
        d = DAS()
        d.channel_data = c
        d.scan = LinearScan(np.linspace(-4e-3, 4e-3, 33), np.linspace(5e-3, 25e-3, 65))
        d.receive_apodization = Apodization(Window.tukey25, f_number=1.7)
        return d.go().data.cpu().numpy()

    # the stored scalars and pitch are single precision: beamform the
    # in-memory data with the same values
    ref_cd = ChannelData(probe=back.probe, sequence=back.sequence, sampling_frequency=back.sampling_frequency,
                         initial_time=back.initial_time, modulation_frequency=back.modulation_frequency,
                         data=cd.data)
    assert np.array_equal(image(back), image(ref_cd))
    frames = U.serving_frames(path)
    assert len(frames) == F and all(np.array_equal(fr[0], x[i]) for i, fr in enumerate(frames))
