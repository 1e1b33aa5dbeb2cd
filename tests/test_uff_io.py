"""UFF file I/O (CPU): uff.write_object / uff.read_object conventions
(+uff/write_object.m, +uff/read_object.m) through the HDF5 C library.

- round trip of channel data (probe, sequence, scalars, complex and real
  data, one and several frames) through write_object / read_object;
- the stored form as write_object.m writes it: the version attribute, class /
  name / size / array attributes, real and imag datasets, dimensions reversed
  with trailing singletons of the MATLAB size dropped, arrays of objects as
  <name>_0001 ..., a one-element array as the object itself;
- wave.fix_origin_from_source applied on reading (wave.m:89-114);
- frame-by-frame reading (a hyperslab per frame).
No UFF file from the reference is available offline, so the format is checked
against the conventions the reference's writer and reader follow.
"""
import numpy as np
import pytest

from ustb_amd import uff_io as U
from ustb_amd.enums import Wavefront
from ustb_amd.uff import ChannelData, LinearArray, MatrixArray, Point, Wave

try:
    U.h5()
    HAVE_HDF5 = True
except OSError:
    HAVE_HDF5 = False

pytestmark = pytest.mark.skipif(not HAVE_HDF5, reason="no HDF5 library")


def _cd(F=2, W=3, N=8, L=50, complex_data=True, probe=None):
    rng = np.random.default_rng(F * 100 + W)
    probe = probe or LinearArray(N, 0.3e-3)
    N = probe.N_elements
    seq = [Wave.plane(0.1 * w) for w in range(W - 1)] + [Wave.spherical(1e-3, 0.0, -10e-3)]
    x = rng.standard_normal((F, W, N, L)).astype(np.float32)
    if complex_data:
        x = (x + 1j * rng.standard_normal((F, W, N, L))).astype(np.complex64)
    return ChannelData(probe=probe, sequence=seq, sampling_frequency=20.8e6, initial_time=1.25e-6,
                       sound_speed=1500.0, modulation_frequency=5.2e6 if complex_data else 0.0,
                       data=x.transpose(3, 2, 1, 0))


@pytest.mark.parametrize("F,W", [(2, 3), (1, 3), (3, 1), (1, 1)])
@pytest.mark.parametrize("complex_data", [True, False])
def test_channel_data_round_trip(tmp_path, F, W, complex_data):
    cd = _cd(F=F, W=W, complex_data=complex_data) if W > 1 else _cd(F=F, W=1, complex_data=complex_data)
    if W == 1:
        cd.sequence = [Wave.plane(0.05)]
    path = str(tmp_path / "a.uff")
    assert U.write_object(path, cd, "channel_data") > 0
    got = U.read_channel_data(path, "/channel_data")
    assert got.data.shape == cd.data.shape
    assert np.array_equal(np.asarray(got.data), np.asarray(cd.data))
    for k in ("sampling_frequency", "initial_time", "sound_speed", "modulation_frequency"):
        assert getattr(got, k) == pytest.approx(float(np.float32(getattr(cd, k))), rel=0, abs=0)
    # numbers are stored in single precision (write_object.m:55): the pitch,
    # hence the element positions, come back to FP32 accuracy
    np.testing.assert_allclose(got.probe.x, cd.probe.x, rtol=1e-6, atol=0)
    assert len(got.sequence) == len(cd.sequence)
    for a, b in zip(got.sequence, cd.sequence):
        assert a.wavefront == b.wavefront
        np.testing.assert_allclose(a.source.xyz, b.source.xyz, rtol=1e-6, atol=1e-12)
        assert a.delay == pytest.approx(b.delay) and a.sound_speed == pytest.approx(b.sound_speed)
    # frames one at a time
    fr = U.FrameReader(path, "/channel_data")
    assert fr.frames == F and fr.complex == complex_data
    for f in range(F):
        assert np.array_equal(fr.frame(f), np.asarray(cd.data).transpose(3, 2, 1, 0)[f])


def test_stored_form_follows_write_object(tmp_path):
    cd = _cd(F=2, W=3, N=8, L=50)
    path = str(tmp_path / "b.uff")
    U.write_object(path, cd, "channel_data")
    with U._File(path, "r") as f:
        L = U.h5()
        assert U._read_attr(f.id, "version") == "v1.2.0"
        g = L.H5Gopen2(f.id, b"channel_data", 0)
        assert U._read_attr(g, "class") == "uff.channel_data"
        assert U._read_attr(g, "name") == "channel_data"
        np.testing.assert_array_equal(U._read_attr(g, "size"), [1, 1])
        np.testing.assert_array_equal(U._read_attr(g, "array"), [0])
        names = dict(U._children(g))
        assert names["data"] == U.H5I_GROUP and names["sampling_frequency"] == U.H5I_DATASET
        d = L.H5Gopen2(g, b"data", 0)
        assert U._read_attr(d, "class") == "single" and U._read_attr(d, "complex")[0] == 1
        re = U._read_dataset(d, "real")
        assert re.shape == (2, 3, 8, 50)  # [L N W F] reversed
        L.H5Gclose(d)
        s = L.H5Gopen2(g, b"sequence", 0)
        assert U._read_attr(s, "class") == "uff.wave" and U._read_attr(s, "array")[0] == 1
        np.testing.assert_array_equal(U._read_attr(s, "size"), [1, 3])
        assert [n for n, _ in U._children(s)] == ["sequence_0001", "sequence_0002", "sequence_0003"]
        L.H5Gclose(s)
        p = L.H5Gopen2(g, b"probe", 0)
        assert U._read_attr(p, "class") == "uff.linear_array"
        assert U._read_dataset(p, "geometry").shape == (7, 8)  # [N 7] reversed
        L.H5Gclose(p)
        assert U._read_dataset(g, "sampling_frequency").shape == (1, 1)
        L.H5Gclose(g)
    # one frame: the trailing singleton of [L N W 1] is dropped
    path1 = str(tmp_path / "c.uff")
    U.write_object(path1, _cd(F=1, W=3, N=8, L=50), "cd")
    with U._File(path1, "r") as f:
        L = U.h5()
        d = L.H5Gopen2(f.id, b"cd/data", 0)
        assert U._read_dataset(d, "real").shape == (3, 8, 50)
        L.H5Gclose(d)


def test_single_wave_is_written_as_the_object(tmp_path):
    cd = _cd(F=1, W=1)
    cd.sequence = [Wave.plane(0.0)]
    path = str(tmp_path / "d.uff")
    U.write_object(path, cd, "channel_data")
    node = U.read_object(path, "/channel_data")
    assert isinstance(node["sequence"], U.Node) and node["sequence"].cls == "uff.wave"
    assert node["sequence"]["wavefront"] == Wavefront.plane


def test_fix_origin_from_source_on_read(tmp_path):
    """A spherical wave with a finite source and a zero origin gets the
    source's x, y and z (z only when behind the probe) on reading."""
    cd = _cd(F=1, W=2)
    cd.sequence = [Wave(Wavefront.spherical, Point.from_xyz(2e-3, 0.0, -10e-3), Point()),
                   Wave(Wavefront.spherical, Point.from_xyz(1e-3, 0.0, 20e-3), Point())]
    path = str(tmp_path / "e.uff")
    U.write_object(path, cd, "channel_data")
    got = U.read_channel_data(path)
    np.testing.assert_allclose(got.sequence[0].origin.xyz, (2e-3, 0.0, -10e-3), rtol=1e-6, atol=1e-12)
    np.testing.assert_allclose(got.sequence[1].origin.xyz, (1e-3, 0.0, 0.0), rtol=1e-6, atol=1e-12)


def test_matrix_array_and_errors(tmp_path):
    cd = _cd(F=1, W=2, probe=MatrixArray(4, 3, 0.3e-3, 0.25e-3))
    path = str(tmp_path / "f.uff")
    U.write_object(path, cd, "channel_data")
    got = U.read_channel_data(path)
    np.testing.assert_allclose(got.probe.x, cd.probe.x, rtol=1e-6, atol=0)
    np.testing.assert_allclose(got.probe.y, cd.probe.y, rtol=1e-6, atol=0)
    with pytest.raises(FileExistsError):
        U.write_object(path, cd, "channel_data")
    with pytest.raises(KeyError, match="does not exist"):
        U.read_object(path, "/nothing")
    with pytest.raises(FileNotFoundError):
        U.read_object(str(tmp_path / "missing.uff"), "/channel_data")


@pytest.mark.parametrize("F", [1, 3])
def test_frame_reader_into_buffers(tmp_path, F):
    """frame(f, out=...) fills a caller's buffer (IQ and RF) and planes(f)
    returns the stored real and imag planes of the same frame."""
    for complex_data in (True, False):
        cd = _cd(F=F, W=3, complex_data=complex_data)
        path = str(tmp_path / f"g{int(complex_data)}.uff")
        U.write_object(path, cd, "channel_data")
        fr = U.FrameReader(path, "/channel_data")
        ref = np.asarray(cd.data).transpose(3, 2, 1, 0)
        out = np.empty(ref.shape[1:], ref.dtype)
        for f in reversed(range(F)):
            got = fr.frame(f, out=out)
            assert got is out and np.array_equal(out, ref[f])
        with pytest.raises(ValueError, match="C-contiguous"):
            fr.frame(0, out=np.empty(ref.shape[1:], np.float64))
        if complex_data:
            buf = np.empty((2,) + ref.shape[1:], np.float32)
            for f in range(F):
                assert fr.planes(f, out=buf) is buf
                assert np.array_equal(buf[0], ref[f].real) and np.array_equal(buf[1], ref[f].imag)
            with pytest.raises(IndexError):
                fr.planes(F)
        else:
            with pytest.raises(ValueError, match="IQ"):
                fr.planes(0)
