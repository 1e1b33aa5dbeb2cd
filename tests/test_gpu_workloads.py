"""The batched (C4) and served (C5) workloads and the full-size C1 image
against the das_c restatement of the oracle, on the oracle's own FP32
matrices (its restatement of das.m / apodization.m):

- C1 at full size (65,536 pixels x 128 channels, RF): the geometry plan
  within 1e-5 of the peak, and the matrix contract (ustbx_das_run, the MEX
  path) in das_c's summation order bit-identical to das_c;
- C4 reduced (a ragged batch of STA RF, CPWC IQ and 3-D matrix-array
  acquisitions, scale 0.25; the 3-D acquisitions with the receive state
  computed on the fly, the others from receive tables): every image of the
  persistent batch launch and of its hipGraph replay within 1e-5 of das_c
  (IQ: of the FP64 restatement, and of das_c up to das_c's own phase
  rounding);
- C5 reduced (a Poisson stream over the three serving configurations, scale
  0.25, the 3-D one on the fly): every served image within 1e-5 of das_c.
"""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from tests.cases import oracle_matrices
from ustb_amd import configs, ops
from ustb_amd.batch import Batch
from ustb_amd.geometry import make_plan
from ustb_amd.serve import Server, poisson_trace
from ustb_amd.uff import SectorScan

pytestmark = pytest.mark.gpu
TOL = 1e-5


def _case(acq):
    return dict(probe=acq.probe, scan=acq.scan, seq=acq.sequence, rx=acq.rx_apod, tx=acq.tx_apod,
                model=int(acq.tx_model), scan_kind="sector" if isinstance(acq.scan, SectorScan) else "linear")


def _rel(img, ref):
    ok = ~np.isnan(ref)
    assert np.array_equal(np.isnan(img), ~ok), "NaN positions differ"
    return float(np.max(np.abs(img[ok] - ref[ok])) / np.max(np.abs(ref[ok])))


def _dasc(acq, data, mats=None):
    ta, ra, td, rd = mats if mats is not None else oracle_matrices(_case(acq))
    return O.das_c(data, acq.fs, acq.t0, ta, ra, td, rd, acq.fd, int(acq.dimension))


def _check(img, acq, data, mats, what):
    """img against das_c on the oracle's matrices within 1e-5 of the peak.
    IQ data under the separable phase: das_c's own FP32 phase rounding is the
    larger error, so the FP64 restatement (das_c's w0 and index arithmetic)
    is held to 1e-5 and das_c to 1e-5 plus its own distance from it."""
    ref = _dasc(acq, data, mats)
    err = _rel(img, ref)
    if acq.fd == 0.0:
        assert err <= TOL, (what, err)
        return
    ta, ra, td, rd = mats
    exact = O.beamform_fp64(data, acq.fs, acq.t0, ta, ra, td, rd, acq.fd, int(acq.dimension),
                            w0=O.dasc_wd(acq.fd), index="fp32")
    err_exact, dasc_exact = _rel(img, exact), _rel(ref, exact)
    print(f"{what}: |gpu-das_c| {err:.2e}, |gpu-fp64| {err_exact:.2e}, |das_c-fp64| {dasc_exact:.2e}", flush=True)
    assert err_exact <= TOL, (what, err_exact)
    assert err <= TOL + dasc_exact, (what, err, dasc_exact)


def test_c1_full_size(cuda):
    acq = configs.c1()
    assert (acq.P, acq.probe.N_elements, acq.n_samples) == (65536, 128, 2048)
    print(f"case C1 full size P={acq.P}", flush=True)
    mats = oracle_matrices(_case(acq))
    ref = _dasc(acq, acq.data, mats)
    ch = torch.from_numpy(acq.data).to(cuda)
    plan = make_plan(acq.probe, acq.scan, acq.sequence, **acq.plan_kwargs())
    img = plan.run(ch).cpu().numpy()
    plan.close()
    assert _rel(img, ref) <= TOL
    ta, ra, td, rd = [torch.from_numpy(np.ascontiguousarray(m)).to(cuda) for m in mats]
    exact = ops.das(ch, acq.fs, acq.t0, ta, ra, td, rd, acq.fd, 3, dasc_order=True).cpu().numpy()
    assert np.array_equal(exact, ref)
    fast = ops.das(ch, acq.fs, acq.t0, ta, ra, td, rd, acq.fd, 3).cpu().numpy()
    assert _rel(fast, ref) <= TOL


def test_c4_reduced_batch(cuda):
    wl = configs.c4(scale=0.25, n_acq=10, n_speckle=800)
    kinds = {acq.name.split()[0] for acq in wl.configs}
    assert kinds == {"C4a", "C4b", "C4c"}, kinds
    print(f"case C4 reduced: {len(wl.acquisitions)} acquisitions, configs {[a.name for a in wl.configs]}",
          flush=True)
    plans = [make_plan(a.probe, a.scan, a.sequence, separable_phase=True, **a.plan_kwargs()) for a in wl.configs]
    for a, p in zip(wl.configs, plans):
        if a.name.startswith("C4c"):
            p.set_variant(4)  # the receive state on the fly: the 3-D class's small-memory plan
            assert p.variant[0] == 4
    batch = Batch([plans[k] for k, _ in wl.acquisitions])
    chs = [torch.from_numpy(wl.datasets[k][d]).to(cuda) for k, d in wl.acquisitions]
    outs = batch.set_io(chs)
    batch.launch()
    torch.cuda.synchronize()
    first = [o.cpu().numpy() for o in outs]
    for o in outs:
        o.zero_()
    batch.replay()
    torch.cuda.synchronize()
    mats = [oracle_matrices(_case(a)) for a in wl.configs]
    for i, (k, d) in enumerate(wl.acquisitions):
        acq = wl.configs[k]
        assert np.array_equal(outs[i].cpu().numpy(), first[i], equal_nan=True), (i, acq.name)
        _check(first[i], acq, wl.datasets[k][d], mats[k], f"{i} {acq.name}")
    batch.close()
    for p in plans:
        p.close()


def test_c5_reduced_stream(cuda):
    cfgs, datas = configs.c5(scale=0.25, n_speckle=800)
    print(f"case C5 reduced: {[a.name for a in cfgs]}", flush=True)
    plans = [make_plan(a.probe, a.scan, a.sequence, separable_phase=True, **a.plan_kwargs()) for a in cfgs]
    plans[2].set_variant(4)  # the 3-D configuration with the receive state on the fly
    srv = Server(plans, datas)
    trace = poisson_trace(40, rate=2000.0, n_configs=len(cfgs), n_data=len(datas[0]), seed=5)
    assert {r.config for r in trace} == {0, 1, 2}
    res = srv.run(trace, keep_outputs=True)
    mats = [oracle_matrices(_case(a)) for a in cfgs]
    seen = {}
    for r in trace:
        key = (r.config, r.data)
        if key not in seen:
            seen[key] = res.outputs[r.rid]
            _check(res.outputs[r.rid], cfgs[r.config], datas[r.config][r.data], mats[r.config], str(r))
        else:  # the same frame again: the same image
            assert np.array_equal(res.outputs[r.rid], seen[key], equal_nan=True), r
    assert np.all(res.latency_s > 0)
    srv.close()
    for p in plans:
        p.close()
