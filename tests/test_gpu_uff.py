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


def test_streamed_planes_beamform_like_memory(cuda, tmp_path):
    """The streaming form: FrameReader.planes into a pinned staging buffer,
    one copy, the complex frame built on the GPU, a one-frame plan. Every
    frame's image equals the in-memory frame's bit for bit."""
    try:
        U.h5()
    except OSError:
        pytest.skip("no HDF5 library")
    from ustb_amd.geometry import make_plan

    rng = np.random.default_rng(10)
    F, W, N, L = 3, 3, 16, 300
    prb = LinearArray(N, 0.3e-3)
    seq = [Wave.plane(a) for a in np.linspace(-0.1, 0.1, W)]
    x = (rng.standard_normal((F, W, N, L)) + 1j * rng.standard_normal((F, W, N, L))).astype(np.complex64)
    cd = ChannelData(probe=prb, sequence=seq, sampling_frequency=10.4e6, initial_time=0.0,
                     modulation_frequency=5.2e6, data=x.transpose(3, 2, 1, 0))
    path = str(tmp_path / "stream.uff")
    U.write_object(path, cd, "channel_data")
    scan = LinearScan(np.linspace(-2e-3, 2e-3, 17), np.linspace(5e-3, 15e-3, 64))
    plan = make_plan(prb, scan, seq, n_samples=L, n_frames=1, sampling_frequency=10.4e6, initial_time=0.0,
                     modulation_frequency=5.2e6, sound_speed=1540.0)
    reader = U.FrameReader(path, "/channel_data")
    host = torch.empty((2, W, N, L), dtype=torch.float32).pin_memory()
    dpl = torch.empty_like(host, device=cuda)
    dch = torch.empty((1, W, N, L), dtype=torch.complex64, device=cuda)
    try:
        for f in range(F):
            reader.planes(f, out=host.numpy())
            dpl.copy_(host, non_blocking=True)
            torch.complex(dpl[0], dpl[1], out=dch[0])
            got = plan.run(dch).cpu().numpy()
            ref = plan.run(torch.from_numpy(x[f:f + 1].copy()).to(cuda)).cpu().numpy()
            assert np.array_equal(got, ref), f
    finally:
        plan.close()
