"""Tile autotuner for geometry plans.

Times every applicable (kernel variant, waves per chunk, split) candidate of a
plan on its own channel data with HIP events, checks that each candidate's
image agrees with the first one (the candidates only regroup the same sums),
and keeps the fastest. Results are cached per shape class -- (dimension,
interpolation, phase, waves, channels, frames, pixels rounded to a power of
two) -- in memory and, optionally, in a JSON file.
"""
from __future__ import annotations

import json
import math
import os
import threading

import torch

from . import _lib as L

VARIANT0 = [(0, wc, sp) for wc in (8, 16) for sp in (1, 4)]
VARIANT0_EXACT = [(0, 8, 1), (0, 8, 4)]
VARIANT1 = [(1, wc, ws) for ws in (4, 8) for wc in (8, 10, 12, 16)]
VARIANT2 = [(2, wc, ws) for ws in (4, 5, 8) for wc in (4, 5, 6, 8, 10, 12, 16)]
VARIANT3 = [(3, wc, ws) for ws in (4, 8) for wc in (6, 10, 16)]
VARIANT2_TX = [(2, fc, cs) for fc in (4, 8) for cs in (4, 8)]   # frames per block, channels per block

_cache: dict = {}
_lock = threading.Lock()


def shape_class(plan) -> str:
    P2 = 1 << max(0, int(math.ceil(math.log2(max(plan.P, 1)))))
    return f"dim{plan.dim}-W{plan.W}-N{plan.N}-F{plan.F}-P{P2}-interp{plan.interp}-sep{int(plan.separable)}"


def candidates(plan):
    if plan.dim in (L.DIM_BOTH, L.DIM_TX):
        c = list(VARIANT0 if plan.separable else VARIANT0_EXACT)
    else:
        c = [(0, 8, 1)] if not plan.separable else [(0, 8, 1), (0, 16, 1)]
    if plan.dim == L.DIM_BOTH and plan.separable and plan.F == 1 and plan.fd != 0.0:
        c += VARIANT1
    if plan.dim == L.DIM_TX and plan.has_variant(2):
        c += VARIANT2_TX
    if plan.dim == L.DIM_BOTH and plan.has_variant(2):
        c += VARIANT2
        if plan.interp == L.INTERP_CUBIC and plan.has_variant(3):
            c += VARIANT3
    return c


def apply(plan, cand):
    variant, wc, sp = cand
    plan.set_variant(variant)
    plan.set_tiles(wc, sp)


def tune(plan, ch: torch.Tensor, reps: int = 3, cache_file: str | None = None, verbose: bool = False):
    """Pick and apply the fastest tiles; returns (candidate, {candidate: ms})."""
    key = shape_class(plan)
    with _lock:
        if key not in _cache and cache_file and os.path.exists(cache_file):
            try:
                _cache.update({k: tuple(v) for k, v in json.load(open(cache_file)).items()})
            except (OSError, ValueError):
                pass
        if key in _cache:
            apply(plan, _cache[key])
            return _cache[key], {}
    out = torch.empty(plan.out_shape, dtype=torch.complex64, device=ch.device)
    ref = None
    times = {}
    stream = torch.cuda.current_stream(ch.device)
    cands = candidates(plan)
    for cand in cands:
        apply(plan, cand)
        plan.run(ch, out)  # warm-up (and first-touch of code)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            plan.run(ch, out)
        e1.record(stream)
        torch.cuda.synchronize(ch.device)
        ms = e0.elapsed_time(e1) / reps
        img = out.clone()
        if ref is None:
            ref = img
        else:
            peak = torch.nan_to_num(ref).abs().max().item()
            err = torch.nan_to_num(img - ref).abs().max().item()
            if not err <= 1e-5 * max(peak, 1e-30):
                raise RuntimeError(f"tile candidate {cand} disagrees with {cands[0]}: {err / peak:.2e}")
        times[cand] = ms
        if verbose:
            print(f"[autotune] {key} {cand}: {ms:.3f} ms", flush=True)
    best = min(times, key=times.get)
    apply(plan, best)
    with _lock:
        _cache[key] = best
        if cache_file:
            try:
                json.dump({k: list(v) for k, v in _cache.items()}, open(cache_file, "w"), indent=1)
            except OSError:
                pass
    return best, times
