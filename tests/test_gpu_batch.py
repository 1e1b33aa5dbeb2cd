"""The persistent batch launch (many acquisitions, dimension both).

Every image of a batch must equal its plan's own run with the same tile shape
bit for bit (both evaluate each tile with the same body in the same order);
the plans' own runs are checked against the oracle in test_gpu_geom.py. The
hipGraph replay must give the same images on every replay, with eager work
and synchronizations between replays, and must follow set_io to new buffers.
"""
import numpy as np
import pytest
import torch

from tests.cases import channel_noise, geom_cases
from ustb_amd import _lib as L
from ustb_amd.batch import Batch
from ustb_amd.enums import Interpolation
from ustb_amd.geometry import make_plan

pytestmark = pytest.mark.gpu
CASES = geom_cases()


def _plan(case, F=1, interp=Interpolation.linear, dim=3):
    return make_plan(case["probe"], case["scan"], case["seq"], n_samples=case["L"], n_frames=F,
                     sampling_frequency=case["fs"], initial_time=0.0, modulation_frequency=case["fd"],
                     sound_speed=1540.0, receive_apodization=case["rx"], transmit_apodization=case["tx"],
                     dimension=dim, tx_model=case["model"], separable_phase=True, interpolation=int(interp))


def _mix(interp, frames=(1, 2, 1)):
    """Plans of the cases that have variant-2 state, with mixed frame counts,
    one plan listed twice (two acquisitions of one configuration). The focused
    cases and cpwc_depth0 hold NaN weights: their tiles take the SAFE body."""
    plans, datas, used = [], [], []
    k = 0
    for name in sorted(CASES):
        case = CASES[name]
        F = frames[k % len(frames)]
        plan = _plan(case, F=F, interp=interp)
        if not plan.has_variant(2) or not plan.batchable:
            plan.close()
            continue
        plans.append(plan)
        used.append(case)
        datas.append(channel_noise(case, F=F, seed=100 + k))
        k += 1
    plans.append(plans[0])
    datas.append(channel_noise(used[0], F=plans[0].F, seed=999))
    return plans, datas


def _reference(plans, chs, wc, ws):
    outs = []
    for p, c in zip(plans, chs):
        p.set_variant(2)
        p.set_tiles(wc, ws)
        outs.append(p.run(c).cpu().numpy())
    return outs


def _close(plans):
    for p in {id(p): p for p in plans}.values():
        p.close()


@pytest.mark.parametrize("interp", [Interpolation.linear, Interpolation.cubic])
def test_batch_equals_plan_runs(cuda, interp):
    plans, datas = _mix(interp)
    assert len(plans) >= 8
    chs = [torch.from_numpy(d).to(cuda) for d in datas]
    b = Batch(plans)
    shape = b.shape
    assert shape["items"] == sum(((p.P + 63) // 64) * p.F for p in plans)
    outs = b.set_io(chs)
    b.launch()
    torch.cuda.synchronize()
    ref = _reference(plans, chs, shape["wave_chunk"], shape["split"])
    for o, r in zip(outs, ref):
        assert np.array_equal(o.cpu().numpy(), r, equal_nan=True)
    b.close()
    _close(plans)


@pytest.mark.parametrize("wc,ws", [(4, 4), (5, 5), (6, 8), (10, 8), (16, 4)])
def test_batch_tile_shapes(cuda, wc, ws):
    plans, datas = _mix(Interpolation.linear)
    chs = [torch.from_numpy(d).to(cuda) for d in datas]
    b = Batch(plans)
    b.set_tiles(wc, ws)
    outs = b.set_io(chs)
    b.launch()
    torch.cuda.synchronize()
    for o, r in zip(outs, _reference(plans, chs, wc, ws)):
        assert np.array_equal(o.cpu().numpy(), r, equal_nan=True)
    b.close()
    _close(plans)


def test_batch_graph_replays(cuda):
    plans, datas = _mix(Interpolation.linear)
    chs = [torch.from_numpy(d).to(cuda) for d in datas]
    b = Batch(plans)
    outs = b.set_io(chs)
    b.launch()
    torch.cuda.synchronize()
    eager = [o.cpu().numpy() for o in outs]
    junk = torch.randn(1 << 20, device=cuda)
    for rep in range(4):
        for o in outs:
            o.zero_()
        junk.mul_(1.0001).add_(0.5)  # unrelated eager work between replays
        b.replay()
        torch.cuda.synchronize()
        for o, e in zip(outs, eager):
            assert np.array_equal(o.cpu().numpy(), e, equal_nan=True), f"replay {rep}"
    # new buffers under the captured graph
    chs2 = [torch.flip(c, dims=[-1]).contiguous() for c in chs]
    outs2 = b.set_io(chs2)
    b.replay()
    torch.cuda.synchronize()
    shape = b.shape
    for o, r in zip(outs2, _reference(plans, chs2, shape["wave_chunk"], shape["split"])):
        assert np.array_equal(o.cpu().numpy(), r, equal_nan=True)
    b.close()
    _close(plans)


def test_batch_rejections(cuda):
    case = CASES["cpwc_tukey"]
    lin = _plan(case)
    cub = _plan(case, interp=Interpolation.cubic)
    with pytest.raises(L.USTBXError, match="share one interpolation"):
        Batch([lin, cub])
    tx = _plan(case, dim=2)
    with pytest.raises(L.USTBXError):
        Batch([lin, tx])
    b = Batch([lin])
    with pytest.raises(L.USTBXError, match="buffers first"):
        b.launch()
    b.close()
    assert lin.batchable and not tx.batchable
    for p in (lin, cub, tx):
        p.close()
