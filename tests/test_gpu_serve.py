"""The serving loop (continuous batching): every request's image is the same
alone and in the mix. Requests arrive faster than one at a time can be
served, large frames are sliced over several iterations, and a small HBM
budget forces admission to wait; each served image must equal its plan's own
run with the batcher's tile shape bit for bit, and the plain one-launch loop
must agree within 1e-5 of the peak magnitude."""
import numpy as np
import pytest
import torch

from tests.cases import channel_noise, geom_cases
from ustb_amd.geometry import make_plan
from ustb_amd import _lib as L
from ustb_amd.serve import Server, poisson_trace

pytestmark = pytest.mark.gpu
CASES = geom_cases()
NAMES = ["cpwc_tukey", "dw_sector", "matrix3d_dw", "cpwc_depth0"]  # the last: NaN weights on its first row


def _setup():
    plans, datas = [], []
    for i, name in enumerate(NAMES):
        case = CASES[name]
        plan = make_plan(case["probe"], case["scan"], case["seq"], n_samples=case["L"], n_frames=1,
                         sampling_frequency=case["fs"], initial_time=0.0, modulation_frequency=case["fd"],
                         sound_speed=1540.0, receive_apodization=case["rx"], transmit_apodization=case["tx"],
                         dimension=3, tx_model=case["model"], separable_phase=True)
        assert plan.has_variant(2), name
        plans.append(plan)
        datas.append([channel_noise(case, seed=300 + 10 * i + j) for j in range(2)])
    return plans, datas


def _alone(plans, datas, tiles):
    ref = {}
    for k, p in enumerate(plans):
        p.set_variant(2)
        p.set_tiles(*tiles)
        for d, x in enumerate(datas[k]):
            ref[(k, d)] = p.run(torch.from_numpy(x).cuda()).cpu().numpy()
    return ref


@pytest.mark.parametrize("budget_frames,slice_ms", [(None, 4.0), (None, 0.0), (2.5, 0.01)])
def test_served_images_equal_alone(cuda, budget_frames, slice_ms):
    plans, datas = _setup()
    kw = {}
    if budget_frames is not None:
        big = max(int(np.prod(p.out_shape)) * 8 + d[0].nbytes for p, d in zip(plans, datas))
        kw["hbm_budget"] = int(budget_frames * big)
    if slice_ms < 1.0:  # tiny iterations whether or not a backlog waits
        kw["idle_slice_ms"] = slice_ms
    srv = Server(plans, datas, slice_ms=slice_ms, **kw)
    trace = poisson_trace(60, rate=20000.0, n_configs=len(plans), n_data=2, seed=11)
    res = srv.run(trace, keep_outputs=True)
    assert len(res.outputs) == len(trace)
    assert res.iterations >= 2
    ref = _alone(plans, datas, srv.tiles)
    for r in trace:
        assert np.array_equal(res.outputs[r.rid], ref[(r.config, r.data)], equal_nan=True), r
    assert np.all(res.latency_s > 0)
    # the plain loop: one launch per request with the plans' default tiles
    for p in plans:
        p.set_variant(2)
    plain = srv.plain_loop(trace[:20], keep_outputs=True)
    for r in trace[:20]:
        a, b = plain.outputs[r.rid], ref[(r.config, r.data)]
        assert np.array_equal(np.isnan(a), np.isnan(b))
        assert np.nanmax(np.abs(a - b)) <= 1e-5 * np.nanmax(np.abs(b))
    srv.close()
    for p in plans:
        p.close()


def _trace_ok(res, trace, ref, want):
    """Requests with status "ok" equal their image alone, bit for bit; the
    statuses are the expected ones."""
    for r in trace:
        assert res.status[r.rid] == want(r), (r, res.status[r.rid], res.errors.get(r.rid))
        if want(r) == "ok":
            assert np.array_equal(res.outputs[r.rid], ref[(r.config, r.data)], equal_nan=True), r
            assert res.latency_s[r.rid] > 0
        else:
            assert r.rid not in res.outputs and np.isnan(res.latency_s[r.rid])


def test_over_budget_requests_are_rejected(cuda):
    """A request whose input and output alone exceed the HBM budget is
    rejected (never admitted); the others are served as alone."""
    plans, datas = _setup()
    need = [int(np.prod(p.out_shape)) * 8 + d[0].nbytes for p, d in zip(plans, datas)]
    big = {k for k, v in enumerate(need) if v == max(need)}
    assert len(big) < len(plans)
    srv = Server(plans, datas, hbm_budget=max(need) - 1)
    trace = poisson_trace(40, rate=20000.0, n_configs=len(plans), n_data=2, seed=12)
    assert any(r.config in big for r in trace)
    res = srv.run(trace, keep_outputs=True)
    ref = _alone(plans, datas, srv.tiles)
    _trace_ok(res, trace, ref, lambda r: "rejected" if r.config in big else "ok")
    assert "budget" in res.errors[next(r.rid for r in trace if r.config in big)]
    assert res.summary()["status"]["rejected"] == sum(r.config in big for r in trace)
    srv.close()
    for p in plans:
        p.close()


def test_a_failing_request_does_not_take_the_others_down(cuda):
    """The library refuses every submission that holds configuration 1 (forced
    on the host): the iterations that mix it with others are retried one
    request at a time, so only configuration 1's requests fail and every
    other request is served as alone."""
    plans, datas = _setup()
    srv = Server(plans, datas, slice_ms=4.0)
    bad = plans[1]
    real = srv._submit

    def refusing(ps, chs, outs, begin, end):
        if any(p is bad for p in ps):
            raise L.USTBXError(L.E_ARG, "", "forced failure")
        real(ps, chs, outs, begin, end)

    srv._submit = refusing
    trace = poisson_trace(50, rate=20000.0, n_configs=len(plans), n_data=2, seed=13)
    res = srv.run(trace, keep_outputs=True)
    ref = _alone(plans, datas, srv.tiles)
    _trace_ok(res, trace, ref, lambda r: "failed" if r.config == 1 else "ok")
    assert all("forced failure" in res.errors[r.rid] for r in trace if r.config == 1)
    srv.close()
    for p in plans:
        p.close()


def test_cancelled_requests_release_and_others_complete(cuda):
    """Every third request is cancelled once the loop has asked about it 30
    times (so some go before admission, some while copied in or between the
    slices of a sliced 3-D frame; one that finished first stays "ok"), and
    one request has a deadline of 0 s; every request not cancelled is served
    as alone."""
    plans, datas = _setup()
    srv = Server(plans, datas, slice_ms=0.05, idle_slice_ms=0.05)
    trace = poisson_trace(45, rate=20000.0, n_configs=len(plans), n_data=2, seed=14)
    trace[4].deadline_s = 0.0
    asked = {}

    def cancel(rid):
        asked[rid] = asked.get(rid, 0) + 1
        return rid % 3 == 0 and asked[rid] > 30

    res = srv.run(trace, keep_outputs=True, cancel=cancel)
    ref = _alone(plans, datas, srv.tiles)
    assert res.status[trace[4].rid] == "cancelled"
    n_cancelled = 0
    for r in trace:
        st = res.status[r.rid]
        if r.rid % 3 == 0 or r is trace[4]:
            assert st in ("cancelled", "ok"), (r, st)
        else:
            assert st == "ok", (r, st)
        if st == "ok":
            assert np.array_equal(res.outputs[r.rid], ref[(r.config, r.data)], equal_nan=True), r
        else:
            n_cancelled += 1
            assert r.rid not in res.outputs
    assert n_cancelled >= 2
    srv.close()
    for p in plans:
        p.close()
