"""Host-side tests (CPU): the C-ABI library loads and exports every symbol the
headers declare; das_c's argument contract (+mex/source/das_c.cpp:53-252) as
enforced by ustbx_das_check_mex_args; loud failure without the library;
enumerations; the synthetic generator's determinism."""
import ctypes
import os
import re

import numpy as np
import pytest

from ustb_amd import _lib as L
from ustb_amd import synth
from ustb_amd.enums import Code, Dimension, SphericalTransmitDelayModel, Wavefront, Window
from ustb_amd.uff import LinearArray, Wave

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_functions():
    names = set()
    for h in os.listdir(os.path.join(ROOT, "include")):
        if h.endswith(".h"):
            text = open(os.path.join(ROOT, "include", h)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names |= set(re.findall(r"\b(ustbx_\w+)\s*\(", text))
    return names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(L.LIB_PATH)
    declared = _declared_functions()
    assert len(declared) >= 14
    missing = [n for n in declared if not hasattr(lib, n)]
    assert not missing, missing
    assert set(L.EXPORTS) == declared


def test_abi_version_and_device_count():
    assert L.lib().ustbx_abi_version() == 3
    assert L.lib().ustbx_device_count() >= 0


def _check(nrhs=9, nlhs=1, ch_ndim=4, ch_dims=(64, 8, 3, 2), ch_double=0, mat_ndim=(2, 2, 2, 2),
           mat_dims=((100, 3), (100, 8), (100, 3), (100, 8)), mat_double=(0, 0, 0, 0), numel=(1, 1, 1),
           sc_double=(0, 0, 0), mode=3, cplx=1, fd=5e6):
    d = L.DasDesc()
    st = L.Status()
    arr = lambda t, v: (t * len(v))(*v)
    rc = L.lib().ustbx_das_check_mex_args(
        nrhs, nlhs, ch_ndim, arr(ctypes.c_int64, list(ch_dims) + [0] * (4 - len(ch_dims))), ch_double,
        arr(ctypes.c_int32, mat_ndim), arr(ctypes.c_int64, [x for p in mat_dims for x in p]),
        arr(ctypes.c_int32, mat_double), arr(ctypes.c_int64, numel), arr(ctypes.c_int32, sc_double), mode, cplx,
        fd, ctypes.byref(d), ctypes.byref(st))
    return rc, L.lib().ustbx_status_id(rc).decode(), st.message.decode(), d


def test_mex_args_accept_valid_call():
    rc, ident, msg, d = _check()
    assert rc == 0 and (d.L, d.N, d.W, d.F, d.P, d.dim) == (64, 8, 3, 2, 100, 3)
    rc, _, _, d = _check(ch_ndim=2, ch_dims=(64, 8), mat_dims=((100, 1), (100, 8), (100, 1), (100, 8)),
                         mat_ndim=(2, 2, 2, 2))
    assert rc == 0 and (d.W, d.F) == (1, 1)


@pytest.mark.parametrize("kw,ident,text", [
    (dict(nrhs=8), "Toolbox:SRP_SRC:nrhs", "Too few input arguments"),
    (dict(nrhs=11), "Toolbox:SRP_SRC:nrhs", "Too many input arguments"),
    (dict(nlhs=2), "Toolbox:SRP_SRC:nlhs", "Too many output arguments"),
    (dict(mode=4), "", "Unknown sum mode"),
    (dict(ch_ndim=5), "Toolbox:SRP_SRC:Dimensions", "Unknown channel data format"),
    (dict(ch_double=1), "", "channel data should be single"),
    (dict(mat_ndim=(2, 2, 3, 2)), "Toolbox:SRP_SRC:Dimensions", "Unknown transmit delay format"),
    (dict(mat_dims=((100, 3), (100, 8), (100, 4), (100, 8))), "Toolbox:SRP_SRC:Dimensions",
     "Number of waves in channel data & transmit delays"),
    (dict(mat_dims=((100, 3), (100, 8), (100, 3), (99, 8))), "Toolbox:SRP_SRC:Dimensions",
     "Number of pixels in transmit and receive delays"),
    (dict(mat_dims=((100, 3), (100, 8), (100, 3), (100, 9))), "Toolbox:SRP_SRC:Dimensions",
     "Number of channels in channel data & receive delays"),
    (dict(mat_dims=((101, 3), (100, 8), (100, 3), (100, 8))), "Toolbox:SRP_SRC:Dimensions",
     "Number of pixels in transmit apodization"),
    (dict(mat_dims=((100, 3), (100, 7), (100, 3), (100, 8))), "Toolbox:SRP_SRC:Dimensions",
     "Number of channels in channel data & receive apodization"),
    (dict(mat_double=(0, 0, 1, 0)), "", "transmiy delays"),
    (dict(numel=(2, 1, 1)), "", "sampling frequency should be an escalar"),
    (dict(sc_double=(0, 1, 0)), "", "initial time should be single"),
    (dict(numel=(1, 1, 3)), "", "Modulation frequency 'fd' should be an escalar"),
    (dict(cplx=0), "", "input data is real"),
])
def test_mex_arg_errors(kw, ident, text):
    rc, got_ident, msg, _ = _check(**kw)
    assert rc < 0 and got_ident == ident and text in msg


def test_real_data_without_modulation_is_accepted():
    rc, *_ = _check(cplx=0, fd=0.0)
    assert rc == 0


def test_missing_library_fails_loudly(monkeypatch):
    monkeypatch.setattr(L, "_lib", None)
    monkeypatch.setattr(L, "LIB_PATH", "/nonexistent/libustbx.so")
    with pytest.raises(RuntimeError, match="not built"):
        L.lib()


def test_ops_reject_host_tensors():
    import torch

    from ustb_amd import ops

    x = torch.zeros((1, 1, 2, 8), dtype=torch.complex64)
    m = torch.zeros((1, 4))
    with pytest.raises(TypeError, match="device tensors"):
        ops.das(x, 1.0, 0.0, m, torch.zeros((2, 4)), m, torch.zeros((2, 4)), 0.0, 3)


def test_enums_match_reference():
    assert [int(d) for d in Dimension] == [0, 1, 2, 3]           # dimension.m:27-32
    assert (Code.matlab, Code.mex, Code.matlab_gpu, Code.mex_gpu) == (0, 1, 3, 4)  # code.m:23-28
    assert Code.python == 2 and Code.mex_hip == 5
    assert (Window.none, Window.boxcar, Window.hanning, Window.tukey25, Window.sta, Window.scanline,
            Window.triangle) == (0, 1, 2, 4, 7, 8, 9)              # window.m:24-37
    assert (Wavefront.plane, Wavefront.spherical, Wavefront.photoacoustic) == (0, 1, 2)
    assert SphericalTransmitDelayModel.hybrid == 3


def test_synth_deterministic_across_thread_counts():
    prb = LinearArray(8, 0.3e-3)
    seq = [Wave.plane(0.1), Wave.spherical(1e-3, 0.0, -5e-3)]
    rng = np.random.default_rng(0)
    pts = np.stack([rng.uniform(-3e-3, 3e-3, 50), np.zeros(50), rng.uniform(5e-3, 20e-3, 50)], 1)
    amps = rng.standard_normal(50)
    for iq in (True, False):
        kw = dict(fs=20e6, n_samples=512, f0=5e6, bandwidth=0.6, iq=iq)
        a = synth.echoes(prb, seq, pts, amps, threads=1, **kw)
        b = synth.echoes(prb, seq, pts, amps, threads=7, **kw)
        assert np.array_equal(a, b)
        assert a.shape == (1, 2, 8, 512) and np.abs(a).max() > 0


def test_bench_refuses_more_than_one_gpu():
    import json
    import subprocess
    import sys

    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2
    assert "one GPU" in json.loads(r.stdout.strip().splitlines()[-1])["error"]


def test_rtb_workload_follows_the_reference_example():
    """examples/fresnel/linear_array/RTB_linear_array.m: 128 elements at
    0.3 mm, 31 focused waves with sources at x = -15..15 mm, z = 20 mm and
    origins on the array below them, fs 50 MHz, 512 x 512 scan over x
    +-19.2 mm and z 0-45 mm, Hamming receive f# 2.5 and transmit f# 1.5 with
    3 mm minimum apertures, dimension both."""
    from ustb_amd import configs
    from ustb_amd.uff import LinearScan

    acq = configs.rtb(with_data=False)
    assert acq.probe.N_elements == 128 and np.isclose(acq.probe.x[1] - acq.probe.x[0], 0.3e-3)
    assert len(acq.sequence) == 31 and acq.fs == 50e6 and int(acq.dimension) == int(Dimension.both)
    src = np.array([w.source.xyz for w in acq.sequence])
    org = np.array([w.origin.xyz for w in acq.sequence])
    np.testing.assert_allclose(src[:, 0], np.linspace(-15e-3, 15e-3, 31), atol=1e-12)
    np.testing.assert_allclose(src[:, 2], 20e-3, atol=1e-12)
    np.testing.assert_allclose(org, np.stack([src[:, 0], np.zeros(31), np.zeros(31)], 1), atol=1e-12)
    assert all(w.wavefront == Wavefront.spherical for w in acq.sequence)
    assert isinstance(acq.scan, LinearScan) and acq.scan.shape[:2] == (512, 512)
    assert np.isclose(acq.scan.z_axis[0], 0.0) and np.isclose(acq.scan.z_axis[-1], 45e-3)
    assert np.isclose(acq.scan.x_axis[0], -19.2e-3) and np.isclose(acq.scan.x_axis[-1], 19.2e-3)
    for ap, f in ((acq.rx_apod, 2.5), (acq.tx_apod, 1.5)):
        assert ap.window == Window.hamming and ap.f_number == (f, f) and ap.minimum_aperture == (3e-3, 3e-3)
    assert acq.tx_model == SphericalTransmitDelayModel.hybrid
    assert configs.rtb(with_data=False, tx_model=SphericalTransmitDelayModel.unified).tx_model == 2


def test_python_arithmetic_flag_argument_checks():
    """USTBX_FLAG_PYTHON_ARITHMETIC needs the Python edge rule, linear
    interpolation and a positive sample period; the descriptor is checked
    before any device pointer is used."""
    import ctypes

    base = dict(L=16, N=2, W=1, F=1, P=4, fs=1e-7, t0=0.0, fd=0.0, dim=3, interp=L.INTERP_LINEAR,
                edge=L.EDGE_PYTHON, flags=L.FLAG_PYTHON_ARITHMETIC, split=0)
    for over in (dict(edge=L.EDGE_DASC), dict(interp=L.INTERP_CUBIC), dict(fs=0.0)):
        d = L.DasDesc(**dict(base, **over))
        st = L.Status()
        rc = L.lib().ustbx_das_run(ctypes.byref(d), None, None, None, None, None, None, None, ctypes.byref(st))
        assert rc == L.E_ARG, over
        assert b"Python arithmetic" in st.message


def test_plan_create_rejects_oversized_and_empty_geometry():
    """ustbx_plan_create checks the geometry on the host before any device
    work: one wave of channel data (N x L samples) of 2 GiB or more (the
    gathers address a wave with 32-bit byte offsets; larger frames are
    fine) is E_CAPACITY; an empty scan, F < 1 or L < 2 are E_DIMENSIONS."""
    import ctypes

    from ustb_amd.geometry import build_geometry
    from ustb_amd.uff import LinearArray, LinearScan, Wave

    probe = LinearArray(128, 0.3e-3)
    scan = LinearScan(np.linspace(-5e-3, 5e-3, 8), np.linspace(5e-3, 10e-3, 8))
    base = dict(n_samples=1024, n_frames=1, sampling_frequency=20e6, initial_time=0.0,
                modulation_frequency=0.0, sound_speed=1540.0)

    def create(seq=None, sc=None, **kw):
        g, keep = build_geometry(probe, sc or scan, seq or [Wave.plane(0.0)], **dict(base, **kw))
        h, st = ctypes.c_void_p(), L.Status()
        rc = L.lib().ustbx_plan_create(ctypes.byref(g), ctypes.byref(h), ctypes.byref(st))
        assert not h.value  # nothing created
        return rc, st.message.decode()

    # 128 channels x 2^21 samples x 8 B = 2 GiB in one wave
    rc, msg = create(n_samples=1 << 21)
    assert rc == L.E_CAPACITY and "2 GiB" in msg
    for kw in (dict(n_samples=1), dict(n_frames=0)):
        rc, _ = create(**kw)
        assert rc == L.E_DIMENSIONS, kw
    rc, _ = create(sc=LinearScan(np.zeros(0), np.zeros(4)))
    assert rc == L.E_DIMENSIONS


def test_host_entry_rejects_null_inputs_before_device_work():
    """ustbx_das_run_host checks its host arrays before it stages anything:
    a NULL input with a nonzero size is E_ARG, here even without a GPU."""
    import ctypes

    d = L.DasDesc(L=16, N=2, W=1, F=1, P=4, fs=1.0, t0=0.0, fd=0.0, dim=3, interp=L.INTERP_LINEAR,
                  edge=L.EDGE_DASC, flags=0, split=0)
    out = np.zeros(4, np.complex64)
    st = L.Status()
    rc = L.lib().ustbx_das_run_host(ctypes.byref(d), None, None, None, None, None, out.ctypes.data, 0,
                                    ctypes.byref(st))
    assert rc == L.E_ARG and b"NULL input" in st.message


def test_plan_key_sees_every_apodization_element():
    """Two receive apodization vectors of 1024 elements that differ in one
    middle element (NumPy's repr elides it) give different plan keys."""
    import numpy as np

    from ustb_amd.das import DAS
    from ustb_amd.uff import Apodization, ChannelData, LinearArray, LinearScan, Wave

    d = DAS()
    prb = LinearArray(4, 0.3e-3)
    d.channel_data = ChannelData(probe=prb, sequence=[Wave.plane(0.0)], sampling_frequency=1e7, initial_time=0.0,
                                 modulation_frequency=0.0, sound_speed=1540.0, data=np.zeros((8, 4, 1, 1)))
    d.scan = LinearScan(np.zeros(1), np.linspace(1e-3, 2e-3, 4))
    v = np.ones(1024, np.float32)
    d.receive_apodization = Apodization(apodization_vector=v.copy())
    k1 = d._plan_key(8, 1)
    v[511] = 0.5
    d.receive_apodization = Apodization(apodization_vector=v.copy())
    assert d._plan_key(8, 1) != k1
    d.receive_apodization = Apodization(apodization_vector=v.copy())
    assert d._plan_key(8, 1) == d._plan_key(8, 1)


def test_hilbert_rule_follows_w0():
    """RF is made analytic whenever |2*pi*single(fd)| < eps (das.m:212,
    das.py:143), for host and device data alike."""
    from ustb_amd.das import DAS
    from ustb_amd.uff import ChannelData

    d = DAS()
    d.channel_data = ChannelData(probe=None, sequence=[], sampling_frequency=1e7, modulation_frequency=0.0)
    assert d._needs_hilbert()
    d.channel_data = ChannelData(probe=None, sequence=[], sampling_frequency=1e7, modulation_frequency=5.2e6)
    assert not d._needs_hilbert()


def test_torch_op_fake_kernel_shapes_and_errors():
    """torch.ops.ustbx.das: the fake (meta) kernel gives das_c's output shape
    per dimension and das_c's errors; CPU tensors have no kernel (no
    fallback)."""
    import torch

    from ustb_amd import torch_ops  # noqa: F401  (registers the operator)

    ch = torch.empty((2, 3, 4, 64), dtype=torch.complex64, device="meta")
    ta, ra = torch.empty((3, 100), device="meta"), torch.empty((4, 100), device="meta")
    shapes = {0: (2, 3, 4, 100), 1: (2, 3, 1, 100), 2: (2, 1, 4, 100), 3: (2, 1, 1, 100)}
    for dim, shape in shapes.items():
        out = torch.ops.ustbx.das(ch, 1.0, 0.0, ta, ra, ta, ra, 0.0, dim)
        assert tuple(out.shape) == shape and out.dtype == torch.complex64
    with pytest.raises(RuntimeError, match="Toolbox:SRP_SRC:Dimensions: Number of channels"):
        torch.ops.ustbx.das(ch, 1.0, 0.0, ta, ra, ta, torch.empty((5, 100), device="meta"), 0.0, 3)
    with pytest.raises(RuntimeError, match="^Unknown sum mode"):
        torch.ops.ustbx.das(ch, 1.0, 0.0, ta, ra, ta, ra, 0.0, 4)
    with pytest.raises(RuntimeError, match="single precision"):
        torch.ops.ustbx.das(ch.to(torch.complex128), 1.0, 0.0, ta, ra, ta, ra, 0.0, 3)
    for dim, shape in shapes.items():  # the Python reference's _beamform contract
        out = torch.ops.ustbx.das_python(ch, 0.0, 1e-7, ta, ra, ta, ra, 0.0, dim)
        assert tuple(out.shape) == shape and out.dtype == torch.complex64
    c = torch.zeros((1, 1, 2, 8), dtype=torch.complex64)
    a, b = torch.zeros((1, 5)), torch.zeros((2, 5))
    with pytest.raises(NotImplementedError):
        torch.ops.ustbx.das(c, 1.0, 0.0, a, b, a, b, 0.0, 3)
