"""The tile autotuner (ustb_amd/autotune.py) on the GPU: it times every
candidate tile shape of a plan, checks each image against the first, applies
the fastest, and caches the choice per shape class in memory and in a JSON
file."""
import json

import numpy as np
import pytest
import torch

from tests.cases import channel_noise, geom_cases
from ustb_amd import autotune
from ustb_amd.enums import Interpolation
from ustb_amd.geometry import make_plan

pytestmark = pytest.mark.gpu
CASES = geom_cases()


def _plan(case, dim, **kw):
    return make_plan(case["probe"], case["scan"], case["seq"], n_samples=case["L"], n_frames=1,
                     sampling_frequency=case["fs"], initial_time=0.0, modulation_frequency=case["fd"],
                     sound_speed=1540.0, receive_apodization=case["rx"], transmit_apodization=case["tx"],
                     dimension=dim, tx_model=case["model"], **kw)


@pytest.mark.parametrize("name,dim,interp", [("cpwc_tukey", 3, Interpolation.cubic), ("dw_sector", 2,
                                                                                     Interpolation.linear)])
def test_autotune_picks_applies_and_caches(cuda, tmp_path, name, dim, interp):
    case = CASES[name]
    print(f"case autotune {name} dim={dim}", flush=True)
    plan = _plan(case, dim, separable_phase=True, interpolation=interp)
    ch = torch.from_numpy(channel_noise(case, seed=3)).to(cuda)
    autotune._cache.clear()
    cache = str(tmp_path / "tiles.json")
    best, times = autotune.tune(plan, ch, reps=1, cache_file=cache)
    cands = autotune.candidates(plan)
    assert set(times) == set(cands) and best in cands
    assert best == min(times, key=times.get)
    assert plan.variant == best  # (variant, waves or frames per chunk, splits or channels per block)
    tuned = plan.run(ch).cpu().numpy()
    # the same candidate set by hand gives the same image, bit for bit
    autotune.apply(plan, cands[0])
    first = plan.run(ch).cpu().numpy()
    autotune.apply(plan, best)
    assert np.array_equal(plan.run(ch).cpu().numpy(), tuned, equal_nan=True)
    ok = ~np.isnan(first)
    assert np.max(np.abs(tuned[ok] - first[ok])) <= 1e-5 * np.max(np.abs(first[ok]))
    # the cache: a fresh plan of the same shape class takes the choice without timing
    key = autotune.shape_class(plan)
    assert json.load(open(cache))[key] == list(best)
    autotune._cache.clear()
    other = _plan(case, dim, separable_phase=True, interpolation=interp)
    again, t2 = autotune.tune(other, ch, cache_file=cache)
    assert again == best and t2 == {} and other.variant == best
    other.close()
    plan.close()
