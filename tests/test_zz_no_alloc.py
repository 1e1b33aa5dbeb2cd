"""The allocation-free entry ustbx_plan_run_cf_noalloc (the fused coherence
factor with its partial sums in the library's device scratch array) against
ustbx_plan_run_cf with a caller workspace, through the C-ABI: bit-identical
images and factors on two calls in a row, the second on a smaller
acquisition (fewer pixels and frames); a plan whose partial sums exceed the
scratch array is rejected with USTBX_E_CAPACITY."""
import ctypes

import numpy as np
import pytest
import torch

from tests.cases import channel_noise, geom_cases
from ustb_amd import _lib as L
from ustb_amd.geometry import make_plan
from ustb_amd.uff import SectorScan

pytestmark = pytest.mark.gpu


def _plan(case, scan, F):
    return make_plan(case["probe"], scan, case["seq"], n_samples=case["L"], n_frames=F,
                     sampling_frequency=case["fs"], initial_time=0.0, modulation_frequency=case["fd"],
                     sound_speed=1540.0, receive_apodization=case["rx"], transmit_apodization=case["tx"],
                     dimension=2, tx_model=case["model"], separable_phase=True)


def _both(plan, ch, M, cuda):
    """(workspace entry, allocation-free entry) images and factors."""
    F, P = plan.F, plan.P
    ws_bytes = int(L.lib().ustbx_plan_cf_workspace_bytes(plan.handle))
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device=cuda)
    res = []
    for noalloc in (False, True):
        img = torch.full((F, P), complex(7.0, 7.0), dtype=torch.complex64, device=cuda)
        cf = torch.full((F, P), 7.0, dtype=torch.float32, device=cuda)
        st = L.Status()
        sp = torch.cuda.current_stream(cuda).cuda_stream
        if noalloc:
            rc = L.lib().ustbx_plan_run_cf_noalloc(plan.handle, ch.data_ptr(), img.data_ptr(), cf.data_ptr(),
                                                   M.data_ptr(), L.CF_MATLAB, sp, ctypes.byref(st))
        else:
            rc = L.lib().ustbx_plan_run_cf(plan.handle, ch.data_ptr(), img.data_ptr(), cf.data_ptr(), M.data_ptr(),
                                           L.CF_MATLAB, ws.data_ptr(), ws_bytes, sp, ctypes.byref(st))
        L.check(rc, st)
        torch.cuda.synchronize(cuda)
        res.append((img.cpu().numpy(), cf.cpu().numpy()))
    return res


def test_noalloc_cf_equals_workspace_entry(cuda):
    case = geom_cases()["dw_sector"]
    big = case["scan"]
    small = SectorScan(big.azimuth_axis[::2], big.depth_axis[: big.depth_axis.size // 2])
    for k, (scan, F) in enumerate(((big, 9), (small, 3))):
        print(f"case noalloc call {k}: P={scan.N_pixels} F={F}", flush=True)
        plan = _plan(case, scan, F)
        assert plan.variant[0] == 2
        assert int(L.lib().ustbx_plan_cf_workspace_bytes(plan.handle)) <= int(L.lib().ustbx_scratch_bytes())
        ch = torch.from_numpy(channel_noise(case, F=F, seed=60 + k)).to(cuda)
        M = plan.active_elements(L.DIM_RX)
        (img_w, cf_w), (img_n, cf_n) = _both(plan, ch, M, cuda)
        assert np.isfinite(img_w).any() and np.abs(img_w).max() > 0
        np.testing.assert_array_equal(img_n, img_w)
        np.testing.assert_array_equal(cf_n, cf_w)
        plan.close()


def test_noalloc_cf_rejects_more_than_the_scratch(cuda):
    case = geom_cases()["dw_sector"]
    scan = case["scan"]
    cap = int(L.lib().ustbx_scratch_bytes())
    per_frame = int(np.ceil(case["probe"].N_elements / 8)) * scan.N_pixels * 16
    F = cap // per_frame + 1
    print(f"case noalloc capacity: F={F}", flush=True)
    plan = _plan(case, scan, F)
    assert int(L.lib().ustbx_plan_cf_workspace_bytes(plan.handle)) > cap
    ch = torch.zeros((F, len(case["seq"]), case["probe"].N_elements, case["L"]), dtype=torch.complex64, device=cuda)
    img = torch.empty((F, plan.P), dtype=torch.complex64, device=cuda)
    st = L.Status()
    rc = L.lib().ustbx_plan_run_cf_noalloc(plan.handle, ch.data_ptr(), img.data_ptr(), None, None, L.CF_MATLAB,
                                           torch.cuda.current_stream(cuda).cuda_stream, ctypes.byref(st))
    assert rc == L.E_CAPACITY and b"scratch" in st.message
    plan.close()
