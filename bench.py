#!/usr/bin/env python3
"""Benchmark of the delay-and-sum hot path on one MI355X.

Workload (one step = one beamforming pass over one acquisition): C2, the
compounded CPWC IQ configuration -- 128-element linear array, 75 plane waves
(-16..16 deg), 1024 IQ samples at 10.4 MHz, 512 x 1024 pixels, cubic
interpolation, phase rotation, tx/rx Tukey-25 apodization at f# 1.7,
dimension both. Delays/apodization are computed on the device (geometry plan);
channel data is resident in HBM before the timed region.

Prints one JSON line: Gsamples/s (pixel x channel x wave per second), the
kernel's roofline position (HIP events on the launch stream), the error
against the das_c restatement and an FP64 restatement on a pixel sample, and
the das_c restatement timed on the host cores on a bounded pixel sample.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

VALU_PEAK_TFLOPS = 157.3     # MI355X published FP32 vector peak (packed FMA)
VALU_NONPACKED_TFLOPS = 78.6  # the same without packed FP32 (one FMA per lane per cycle)
HBM_PEAK_GBS = 8000.0        # MI355X published HBM3E peak


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="timed steps (default 20; C4: 5)")
    ap.add_argument("--warmup", type=int, default=None,
                    help="untimed steps (default 25, about 60 ms on C2: the clocks settle within about 40 "
                         "launches, and 3 warm-up steps measured C2 4%% slower; C4: 1)")
    ap.add_argument("--config", default="C2", choices=["C1", "C2", "C3", "C4", "C4C", "C5", "RTB", "RF2IQ", "UFF"],
                    help="C4C: C4's 3-D class alone (one frame of the 32 x 32 matrix array, 25 waves, "
                         "128 x 128 x 256 voxels); RTB: focused transmits (retrospective transmit beamforming), the "
                         "focused-wave delay models; RF2IQ: the RF front end (fast_demodulation and hilbert) on STA RF "
                         "of 128 waves x 128 channels x 2048 samples; UFF: C3's frames streamed from a UFF file through the reader "
                         "into a one-frame plan")
    ap.add_argument("--tx-model", default="hybrid", choices=["spherical", "unified", "hybrid", "blended"],
                    help="RTB: spherical_transmit_delay_model of the focused waves")
    ap.add_argument("--frames", type=int, default=600, help="C5: requests served (the trace's first ones)")
    ap.add_argument("--rate", type=float, default=2000.0, help="C5: Poisson arrival rate, frames/s")
    ap.add_argument("--slice-ms", type=float, default=2.0, help="C5: estimated work per serving iteration")
    ap.add_argument("--long-slices", type=int, default=8,
                    help="C5: iteration length, in slices, while no small frame is pending")
    ap.add_argument("--depth", type=int, default=2,
                    help="C5: serving iterations queued on the GPU at once")
    ap.add_argument("--idle", default="1,4,1",
                    help="C5: slice ms, long slices and depth while the arrived work not yet submitted is "
                         "below 250 ms ('none': the same as under a backlog)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-check", action="store_true")
    ap.add_argument("--phase", default="auto", choices=["auto", "separable", "dasc"],
                    help="auto: das_c's phase for dimension transmit (its sine and cosine are shared by the "
                         "frames of a block), the separable phase otherwise")
    ap.add_argument("--abi", default="geometry", choices=["geometry", "matrix"],
                    help="C1-C3: geometry plan (delays and apodization from resident tables) or the das_c "
                         "matrix contract (ustbx_das_run over das.m's FP32 matrices in HBM, the plain kernel)")
    ap.add_argument("--postprocess", default="none", choices=["none", "cf"],
                    help="C3: das with the coherence factor over the receive channels, the fused kernel "
                         "against the per-channel images followed by the coherence-factor kernel")
    ap.add_argument("--wave-chunk", type=int, default=0)
    ap.add_argument("--split", type=int, default=0)
    ap.add_argument("--variant", type=int, default=-1, help="kernel variant (default: the plan's choice)")
    ap.add_argument("--stage", type=int, default=-1, choices=[-1, 0, 1],
                    help="stage sample windows through LDS (1) or gather (0); default (-1): the plan's "
                         "choice (on for dimension both, off for dimension transmit)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only time a 1 GiB device-to-device copy (achievable HBM bandwidth; also the "
                         "known-size kernel that calibrates the FETCH_SIZE/WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.steps is None:
        args.steps = 5 if args.config == "C4" else 20
    if args.warmup is None:
        args.warmup = 1 if args.config == "C4" else 25
    return args


COPY_BYTES = 1 << 30


def copy_probe(torch, dev, reps=10):
    """Streaming copy of COPY_BYTES (read once, written once) timed with HIP
    events: the achievable HBM bandwidth, counting read + write bytes."""
    src = torch.ones(COPY_BYTES // 4, dtype=torch.float32, device=dev)
    dst = torch.empty_like(src)
    dst.copy_(src)
    stream = torch.cuda.current_stream(dev)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record(stream)
    for _ in range(reps):
        dst.copy_(src)
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    del src, dst
    return {"bytes_per_copy": 2 * COPY_BYTES, "ms": round(ms, 4), "GBs": round(2 * COPY_BYTES / (ms * 1e-3) / 1e9, 1)}


def flops_per_unit(interp: int, iq: bool) -> int:
    """Algorithmic FP32 flops of one active (pixel, channel, wave) term:
    fractional index (c = tx + rx, (c - t0) * fs: 3), interpolation of a
    complex sample (linear: 2 complex-by-real FMAs = 8; cubic: Keys weights 16
    + 4 complex-by-real FMAs 16 = 32), phase rotation and accumulation
    (one complex multiply-add: 8; RF: a real weight, 4)."""
    return 3 + (32 if interp == 1 else 8) + (8 if iq else 4)


def active_units(plan, torch) -> int:
    """Number of (pixel, channel, wave, frame) terms with nonzero apodization:
    the terms das_c evaluates (das_c.cpp:366,371), counted on the device."""
    return plan.active_terms()


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU; --gpus %d is not supported (a frame/"
                                   "acquisition-parallel multi-GPU layout is described in DESIGN.md)" % args.gpus}))
        sys.exit(2)
    import torch

    from ustb_amd import configs
    from ustb_amd.geometry import make_plan

    dev = torch.device("cuda", 0)
    if args.copy_probe:
        print(json.dumps({"copy_probe": copy_probe(torch, dev)}))
        return
    if args.config == "C4":
        run_batch(args, torch, dev)
        return
    if args.config == "C5":
        run_serve(args, torch, dev)
        return
    if args.config == "RF2IQ":
        run_frontend(args, torch, dev)
        return
    if args.config == "UFF":
        run_uff(args, torch, dev)
        return
    if args.abi == "matrix":
        run_matrix(args, torch, dev)
        return
    if args.postprocess == "cf":
        run_cf(args, torch, dev)
        return
    t_gen = time.time()
    acq = build_acquisition(args)
    print(f"[bench] {acq.name}: data {acq.data.shape} generated in {time.time() - t_gen:.1f}s", file=sys.stderr,
          flush=True)
    ch = torch.from_numpy(acq.data).to(dev)
    if args.phase == "auto":
        args.phase = "dasc" if int(acq.dimension) == 2 else "separable"
    torch.cuda.synchronize()
    t_plan = time.perf_counter()
    plan = make_plan(acq.probe, acq.scan, acq.sequence, separable_phase=(args.phase == "separable"),
                     **acq.plan_kwargs())
    torch.cuda.synchronize()
    plan_ms = (time.perf_counter() - t_plan) * 1e3
    if args.variant >= 0:
        plan.set_variant(args.variant)
    if args.wave_chunk or args.split:
        plan.set_tiles(args.wave_chunk, args.split)
    if args.stage >= 0:
        plan.set_staging(bool(args.stage))
    staging = bool(args.stage) if args.stage >= 0 else int(acq.dimension) == 3
    staged_pairs = plan.stage_stats if int(acq.dimension) == 3 and staging else (0, 0)
    variant, wave_chunk, split = plan.variant
    out = torch.empty(plan.out_shape, dtype=torch.complex64, device=dev)
    stream = torch.cuda.current_stream(dev)
    for _ in range(args.warmup):
        plan.run(ch, out)
    torch.cuda.synchronize()

    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev0.record(stream)
    for _ in range(args.steps):
        plan.run(ch, out)
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kernel_ms = ev0.elapsed_time(ev1) / args.steps
    ms_per_step = wall * 1e3 / args.steps
    units = acq.units
    value = units / (ms_per_step * 1e-3) / 1e9
    print(f"[bench] {args.steps} steps: {ms_per_step:.3f} ms/step (kernel {kernel_ms:.3f} ms)", file=sys.stderr,
          flush=True)

    iq = acq.fd != 0.0
    act = active_units(plan, torch)
    fl = flops_per_unit(int(acq.interpolation), iq) * act
    achieved_tflops = fl / (kernel_ms * 1e-3) / 1e12
    in_bytes = acq.data.nbytes
    out_bytes = int(np.prod(plan.out_shape)) * 8
    alg_bytes = in_bytes + out_bytes
    achieved_gbs = alg_bytes / (kernel_ms * 1e-3) / 1e9
    copy = copy_probe(torch, dev)
    kname = "das_tx_rxtab" if int(acq.dimension) == 2 else "das_both_rxtab"
    traffic = _traffic(f"{kname}_{args.config.lower()}_traffic.txt")

    result = {
        "metric": "DAS Gsamples/s (pixel x channel x wave)",
        "value": round(value, 3),
        "unit": "Gsamples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": acq.name, "pixels": acq.P, "channels": acq.probe.N_elements,
                   "waves": len(acq.sequence), "frames": acq.n_frames, "samples": acq.n_samples,
                   "interp": "cubic" if acq.interpolation == 1 else "linear",
                   "dimension": int(acq.dimension), "phase": args.phase,
                   "geometry": "on-device, resident per-configuration tables",
                   "plan_build_ms": round(plan_ms, 2),
                   "kernel_variant": variant, "waves_per_split": wave_chunk, "splits": split,
                   "lds_staging": staging,
                   "staged_tile_channels": round(staged_pairs[0] / staged_pairs[1], 4) if staged_pairs[1] else None},
        "roofline": {"bound": "valu", "achieved": round(achieved_tflops, 3), "peak": VALU_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": round(achieved_tflops / VALU_PEAK_TFLOPS, 4),
                     "traffic": traffic,
                     "flops_per_active_unit": flops_per_unit(int(acq.interpolation), iq),
                     "active_units": act, "kernel_ms": round(kernel_ms, 4),
                     "peak_nonpacked": VALU_NONPACKED_TFLOPS,
                     "frac_nonpacked": round(achieved_tflops / VALU_NONPACKED_TFLOPS, 4)},
        "hbm": {"algorithmic_bytes": alg_bytes, "achieved_GBs": round(achieved_gbs, 1),
                "peak_GBs": HBM_PEAK_GBS, "frac": round(achieved_gbs / HBM_PEAK_GBS, 5),
                "achievable_copy_GBs": copy["GBs"],
                "frac_of_achievable": round(achieved_gbs / copy["GBs"], 5)},
    }

    if not args.no_check:
        result["errors"] = check_errors(acq, plan, out, ch, torch)
    if not args.no_cpu_baseline:
        lin = gpu_linear_rate(acq, ch, args, torch, dev) if int(acq.interpolation) != 0 else None
        result["cpu_baseline"] = cpu_baseline(acq, plan, torch, gpu_linear=lin)
    plan.close()
    print(json.dumps(result))


def _traffic(name):
    """HBM bytes per launch of the dominant kernel from its committed
    rocprofv3 counter summary (FETCH_SIZE and WRITE_SIZE, calibrated on the
    copy probe), or None."""
    path = os.path.join(ROOT, "profiles", name)
    if os.path.exists(path):
        for line in open(path):
            if line.startswith("traffic_bytes_per_launch"):
                return float(line.split("=")[1])
    return None


def build_acquisition(args):
    from ustb_amd import configs
    from ustb_amd.enums import SphericalTransmitDelayModel

    if args.config == "RTB":
        return configs.rtb(tx_model=int(SphericalTransmitDelayModel[args.tx_model]))
    return configs.BUILDERS[args.config]()


def _timed(torch, stream, fn, warmup, steps):
    """(kernel ms from HIP events on the launch stream, wall ms) per step.
    The kernel time is taken with the steps queued ahead: the stream first runs
    a spin kernel that lasts longer than the host takes to enqueue the steps,
    so the events see the kernels back to back and not the host's launch rate
    (a Python call costs about 0.1 ms on some hosts, more than the shortest
    kernels here)."""
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(stream)
    for _ in range(steps):
        fn()
    e1.record(stream)
    torch.cuda.synchronize()
    w_ms = (time.perf_counter() - t0) * 1e3 / steps
    host_paced = e0.elapsed_time(e1) / steps
    if not hasattr(torch.cuda, "_sleep"):  # no spin kernel in this torch build: the host-paced events
        log("torch.cuda._sleep is missing: kernel times are host-paced")
        return host_paced, w_ms
    # spin long enough for the host to enqueue every step (twice the wall time)
    cal = 1 << 20
    with torch.cuda.stream(stream):
        e0.record(stream)
        torch.cuda._sleep(cal)
        e1.record(stream)
    torch.cuda.synchronize()
    spin = int(cal * max(1.0, 2.0 * w_ms * steps / max(e0.elapsed_time(e1), 1e-3)))
    with torch.cuda.stream(stream):
        torch.cuda._sleep(spin)
    e0.record(stream)
    for _ in range(steps):
        fn()
    e1.record(stream)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / steps, w_ms


def run_frontend(args, torch, dev):
    """The RF front end on the device: preprocess.fast_demodulation
    (fast_demodulation.m:58-183: down-mix, Kaiser low-pass, decimation by
    floor(fs / (2 f0)) = 2), the Python reference's zero-phase FastDemodulation
    (decimation floor(fs / (4 f0)) = 1) and the analytic signal of das.m's RF
    branch (das.m:211-214, hipFFT), on STA RF of 128 waves x 128 channels x
    2048 samples at 20.8 MHz (f0 5.2 MHz): seeded noise plus tones at f0.
    One step = one pass of fast_demodulation over the whole acquisition;
    units are RF samples. Each is HBM-bound: algorithmic bytes = RF read once
    (4 B per sample) + output written once (8 B per output sample)."""
    from oracle import oracle as O
    from ustb_amd import _lib as L
    from ustb_amd import ops
    from ustb_amd.preprocess import matlab_lowpass, python_lowpass

    W, N, Ls, fs, f0, t0 = 128, 128, 2048, 20.8e6, 5.2e6, 0.0
    rng = np.random.default_rng(7)
    t = np.arange(Ls) / fs
    rf = (0.3 * rng.standard_normal((1, W, N, Ls))
          + np.cos(2 * np.pi * f0 * t + rng.uniform(0, 6, (1, W, N, 1)))).astype(np.float32)
    x = torch.from_numpy(rf).to(dev)
    S = rf.size
    stream = torch.cuda.current_stream(dev)
    b_m = matlab_lowpass(0.5 * f0, f0, fs)
    down_m = int(math.floor(fs / (2 * f0)))
    b_p = python_lowpass(Ls, 0.5 * f0, fs)
    down_p = int(math.floor(fs / (4 * f0)))
    iq_m = torch.empty((1, W, N, (Ls + down_m - 1) // down_m), dtype=torch.complex64, device=dev)
    an = torch.empty((1, W, N, Ls), dtype=torch.complex64, device=dev)
    copy = copy_probe(torch, dev)

    def line(name, fn, out_elems, taps, samples=S):
        k_ms, w_ms = _timed(torch, stream, fn, args.warmup, args.steps)
        alg = 4 * samples + 8 * out_elems
        gbs = alg / (k_ms * 1e-3) / 1e9
        log(f"{name}: {k_ms:.4f} ms/step (kernel), {w_ms:.4f} ms wall, {gbs:.0f} GB/s")
        return {"kernel_ms": round(k_ms, 4), "wall_ms": round(w_ms, 4), "taps": taps,
                "Gsamples_s": round(samples / (w_ms * 1e-3) / 1e9, 3), "algorithmic_bytes": alg,
                "achieved_GBs": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4),
                "frac_of_achievable": round(gbs / copy["GBs"], 4)}

    demod = line("fast_demodulation (causal Kaiser, /2)",
                 lambda: ops.demodulate(x, fs, t0, f0, b_m, down_m, L.FIR_CAUSAL), iq_m.numel(), len(b_m))
    zp = line("FastDemodulation (zero-phase, /1)",
              lambda: ops.demodulate(x, fs, t0, f0, b_p, down_p, L.FIR_ZERO_PHASE), S, len(b_p))
    hil = line("hilbert", lambda: ops.hilbert(x, an), S, 0)
    # a length of the form 2^a 3^b 5^c 7^d (the mixed-radix LDS path), as RTB's 3500
    x35 = x.reshape(-1, Ls)[:, :1750].contiguous().reshape(-1, 3500)
    an35 = torch.empty(x35.shape, dtype=torch.complex64, device=dev)
    hil35 = line("hilbert L=3500", lambda: ops.hilbert(x35, an35), x35.numel(), 0, samples=x35.numel())

    errors = {}
    if not args.no_check:
        k = 256                                           # traces checked against the oracle
        sub = rf.reshape(-1, Ls)[:k]
        got = ops.demodulate(x, fs, t0, f0, b_m, down_m, L.FIR_CAUSAL).reshape(-1, iq_m.shape[-1])[:k].cpu().numpy()
        ref = O.fast_demodulation(sub, fs, t0, f0)[0]
        errors["fast_demodulation_vs_oracle"] = float(np.max(np.abs(got - ref)) / np.max(np.abs(ref)))
        got = ops.demodulate(x, fs, t0, f0, b_p, down_p, L.FIR_ZERO_PHASE).reshape(-1, Ls)[:k].cpu().numpy()
        ref = O.python_fast_demodulation(sub, fs, t0, f0)[0]
        errors["python_fast_demodulation_vs_oracle"] = float(np.max(np.abs(got - ref)) / np.max(np.abs(ref)))
        got = ops.hilbert(x, an).reshape(-1, Ls)[:k].cpu().numpy()
        ref = O.hilbert(sub.astype(np.float64))
        errors["hilbert_vs_scipy"] = float(np.max(np.abs(got - ref)) / np.max(np.abs(ref)))
        got = ops.hilbert(x35, an35)[:k].cpu().numpy()
        ref = O.hilbert(x35[:k].cpu().numpy().astype(np.float64))
        errors["hilbert_L3500_vs_scipy"] = float(np.max(np.abs(got - ref)) / np.max(np.abs(ref)))
        log(f"errors {errors}")
    result = {
        "metric": "RF front end Gsamples/s (RF time samples)",
        "value": demod["Gsamples_s"],
        "unit": "Gsamples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": demod["wall_ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "RF2IQ fast_demodulation of STA RF, 128 waves x 128 channels x 2048 samples, "
                               "20.8 MHz, f0 5.2 MHz", "rf_samples": S},
        "roofline": {"bound": "hbm", "achieved": demod["achieved_GBs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": demod["frac"], "traffic": _traffic("rf_frontend_rf2iq_traffic.txt"),
                     "kernel_ms": demod["kernel_ms"],
                     "achievable_copy_GBs": copy["GBs"], "frac_of_achievable": demod["frac_of_achievable"]},
        "fast_demodulation": demod,
        "python_fast_demodulation": zp,
        "hilbert": hil,
        "hilbert_L3500": hil35,
        "errors": errors,
    }
    if not args.no_cpu_baseline:
        sub = rf.reshape(-1, Ls)
        O.fast_demodulation(sub[:64], fs, t0, f0)
        t1 = time.perf_counter()
        O.fast_demodulation(sub, fs, t0, f0)
        dt = time.perf_counter() - t1
        result["cpu_baseline"] = {"value": round(sub.size / dt / 1e9, 4), "unit": "Gsamples/s", "cores": 1,
                                  "kind": "port", "sample": f"the oracle's fast_demodulation (NumPy/SciPy lfilter, "
                                                            f"FP64) on {sub.shape[0]} of {W * N} traces, {dt:.2f} s"}
    print(json.dumps(result))


def run_uff(args, torch, dev):
    """UFF channel-data reader in the serving path: C3's acquisition (phased
    64 elements, 31 diverging waves, 2048 IQ samples, sector 256 x 1024,
    dimension transmit) with 8 frames written to a UFF file
    (uff.write_object conventions), then streamed frame by frame: the
    reader's hyperslabs of one frame's real and imag datasets straight into
    a pinned staging buffer, one host to device copy, the complex frame
    built on the GPU from the two planes, one beamforming pass of a
    one-frame plan. Two staging buffers let the host read frame f+1 while
    the GPU copies and beamforms frame f. Units: frames per second end to
    end over the pipelined pass; the reader's own rate in MB/s of channel
    data and the copy + beamform time come from a pass with a
    synchronisation after every frame. The first frame's image equals the
    in-memory frame's bit for bit."""
    import shutil
    import tempfile

    from ustb_amd import configs, uff_io
    from ustb_amd.geometry import make_plan
    from ustb_amd.uff import ChannelData

    frames = 8
    acq = configs.c3(frames=frames)
    cd = ChannelData(probe=acq.probe, sequence=acq.sequence, sampling_frequency=acq.fs, initial_time=acq.t0,
                     sound_speed=acq.c, modulation_frequency=acq.fd,
                     data=np.ascontiguousarray(acq.data.transpose(3, 2, 1, 0)))
    tmp = tempfile.mkdtemp(prefix="ustbx_uff_")
    path = os.path.join(tmp, "c3.uff")
    try:
        t = time.perf_counter()
        uff_io.write_object(path, cd, "channel_data")
        write_s = time.perf_counter() - t
        fbytes = acq.data[0].nbytes
        log(f"wrote {frames} frames ({os.path.getsize(path) / 1e6:.0f} MB) in {write_s:.1f} s")
        plan = make_plan(acq.probe, acq.scan, acq.sequence, separable_phase=False,
                         **acq.plan_kwargs(n_frames=1))
        reader = uff_io.FrameReader(path, "/channel_data")
        pshape = (2,) + acq.data.shape[1:]
        host = [torch.empty(pshape, dtype=torch.float32).pin_memory() for _ in range(2)]
        hview = [h.numpy() for h in host]
        dpl = [torch.empty(pshape, dtype=torch.float32, device=dev) for _ in range(2)]
        dch = [torch.empty((1,) + acq.data.shape[1:], dtype=torch.complex64, device=dev) for _ in range(2)]
        outs = [torch.empty(plan.out_shape, dtype=torch.complex64, device=dev) for _ in range(2)]
        copied = [None, None]

        def step(f, k):
            if copied[k] is not None:
                copied[k].synchronize()  # staging buffer k's previous copy has left the host
            reader.planes(f, out=hview[k])
            dpl[k].copy_(host[k], non_blocking=True)
            ev = torch.cuda.Event()
            ev.record()
            copied[k] = ev
            torch.complex(dpl[k][0], dpl[k][1], out=dch[k][0])
            plan.run(dch[k], outs[k])

        read_ms, gpu_ms = [], []
        first = None
        for f in range(frames):  # serial pass: warms the page cache and the plan, times the parts
            t0 = time.perf_counter()
            reader.planes(f, out=hview[0])
            t1 = time.perf_counter()
            dpl[0].copy_(host[0], non_blocking=True)
            torch.complex(dpl[0][0], dpl[0][1], out=dch[0][0])
            plan.run(dch[0], outs[0])
            torch.cuda.synchronize()
            t2 = time.perf_counter()
            if f == 0:
                first = outs[0].cpu().numpy()
            else:  # frame 0 carries the plan's first launch and a cold page cache
                read_ms.append((t1 - t0) * 1e3)
                gpu_ms.append((t2 - t1) * 1e3)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for f in range(frames):  # pipelined pass: the timed one
            step(f, f % 2)
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / frames
        ref_plan = make_plan(acq.probe, acq.scan, acq.sequence, separable_phase=False,
                             **acq.plan_kwargs(n_frames=1))
        ref = ref_plan.run(torch.from_numpy(np.ascontiguousarray(acq.data[:1])).to(dev)).cpu().numpy()
        same = bool(np.array_equal(first, ref))
        last = bool(np.array_equal(outs[(frames - 1) % 2].cpu().numpy(),
                                   ref_plan.run(torch.from_numpy(np.ascontiguousarray(acq.data[frames - 1:]))
                                                .to(dev)).cpu().numpy()))
        ref_plan.close()
        plan.close()
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    rd = float(np.mean(read_ms))
    result = {
        "metric": "UFF frames/s read and beamformed (C3 frames, one at a time)",
        "value": round(1e3 / ms, 2),
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": frames,
        "warmup": frames,
        "ms_per_step": round(ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "UFF: C3 phased DW 31 waves, sector 256x1024, transmit, frames from a UFF file",
                   "frames": frames, "frame_bytes": fbytes},
        "reader": {"ms_per_frame": round(rd, 3), "MB_s": round(fbytes / (rd * 1e-3) / 1e6, 1),
                   "note": "HDF5 hyperslabs of the real and imag datasets into pinned staging; "
                           "the complex frame is built on the GPU"},
        "copy_and_beamform_ms": round(float(np.mean(gpu_ms)), 3),
        "errors": {"first_frame_equals_memory_frame_bitwise": same,
                   "last_pipelined_frame_equals_memory_frame_bitwise": last},
    }
    log(f"UFF: {ms:.2f} ms per frame pipelined (reader {rd:.2f} ms, copy + beamform {np.mean(gpu_ms):.2f} ms), "
        f"same={same} last={last}")
    print(json.dumps(result))


def run_cf(args, torch, dev):
    """das (dimension transmit) + postprocess.coherence_factor over the
    receive channels: the fused kernel (ustbx_plan_run_cf, no per-channel
    images in HBM) against the plan's transmit kernel writing the per-channel
    images followed by the coherence-factor kernel. Same units as the das
    line (pixel x channel x wave x frame)."""
    from ustb_amd import _lib as L
    from ustb_amd import configs, ops
    from ustb_amd.geometry import make_plan

    acq = build_acquisition(args)
    if int(acq.dimension) != 2:
        raise SystemExit("--postprocess cf needs a dimension-transmit workload (C3)")
    ch = torch.from_numpy(acq.data).to(dev)
    if args.phase == "auto":
        args.phase = "dasc"
    plan = make_plan(acq.probe, acq.scan, acq.sequence, separable_phase=(args.phase == "separable"),
                     **acq.plan_kwargs())
    if args.wave_chunk or args.split:
        plan.set_tiles(args.wave_chunk, args.split)
    if args.stage >= 0:
        plan.set_staging(bool(args.stage))
    M = plan.active_elements(L.DIM_RX)
    per_channel = torch.empty(plan.out_shape, dtype=torch.complex64, device=dev)
    image = torch.empty((plan.F, plan.P), dtype=torch.complex64, device=dev)
    cf = torch.empty((plan.F, plan.P), dtype=torch.float32, device=dev)
    stream = torch.cuda.current_stream(dev)

    def fused():
        plan.run_cf(ch, M, L.CF_MATLAB, image, cf)

    def unfused():
        plan.run(ch, per_channel)
        return ops.coherence_factor(per_channel, L.DIM_RX, M, L.CF_MATLAB)

    def timed(fn):
        for _ in range(args.warmup):
            fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t0 = time.perf_counter()
        e0.record(stream)
        for _ in range(args.steps):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / args.steps, (time.perf_counter() - t0) * 1e3 / args.steps

    fused_ms, fused_wall = timed(fused)
    unfused_ms, unfused_wall = timed(unfused)
    img_u, cf_u = unfused()
    fused()
    torch.cuda.synchronize()
    peak = float(img_u.abs().max())
    err = float((image - img_u[:, 0, 0]).abs().max()) / max(peak, 1e-30)
    cf_err = float((cf - cf_u[:, 0, 0]).abs().max())
    log(f"das + coherence factor: fused {fused_ms:.3f} ms, unfused {unfused_ms:.3f} ms, image {err:.2e}, "
        f"cf {cf_err:.2e}")
    result = {
        "metric": "DAS Gsamples/s (pixel x channel x wave)",
        "value": round(acq.units / (fused_wall * 1e-3) / 1e9, 3),
        "unit": "Gsamples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(fused_wall, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": acq.name + " + coherence factor over the receive channels",
                   "phase": args.phase, "postprocess": "coherence_factor (MATLAB flavour, M from the receive "
                                                       "apodization above 0.16)"},
        "fused": {"kernel_ms": round(fused_ms, 4), "output_bytes": image.numel() * 8 + cf.numel() * 4,
                  "workspace_bytes": int(plan._cf_workspace.numel())},
        "unfused": {"kernel_ms": round(unfused_ms, 4), "per_channel_bytes": per_channel.numel() * 8,
                    "Gsamples_s": round(acq.units / (unfused_wall * 1e-3) / 1e9, 3)},
        "fused_speedup": round(unfused_ms / fused_ms, 3),
        "errors": {"fused_vs_unfused_image": err, "fused_vs_unfused_cf_abs": cf_err},
    }
    plan.close()
    print(json.dumps(result))


def run_matrix(args, torch, dev):
    """The das_c / das_cuda contract on device buffers (ustbx_das_run): das.m's
    FP32 delay and apodization matrices, computed once and resident in HBM,
    read by the plain kernel on every call. Algorithmic bytes per call: channel
    data + the four matrices + the image, each once."""
    from oracle import oracle as O
    from ustb_amd import configs, ops
    from ustb_amd.geometry import make_plan

    acq = build_acquisition(args)
    ch = torch.from_numpy(acq.data).to(dev)
    plan = make_plan(acq.probe, acq.scan, acq.sequence, separable_phase=False, **acq.plan_kwargs())
    ta, ra, td, rd = plan.materialize()
    plan.close()
    dim = int(acq.dimension)
    interp = int(acq.interpolation)
    out = torch.empty(ops.out_shape(dim, acq.P, acq.probe.N_elements, len(acq.sequence), acq.n_frames),
                      dtype=torch.complex64, device=dev)

    def call():
        ops.das(ch, acq.fs, acq.t0, ta, ra, td, rd, acq.fd, dim, interp=interp, split=args.split, out=out)

    for _ in range(args.warmup):
        call()
    stream = torch.cuda.current_stream(dev)
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev0.record(stream)
    for _ in range(args.steps):
        call()
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kernel_ms = ev0.elapsed_time(ev1) / args.steps
    ms_per_step = wall * 1e3 / args.steps
    value = acq.units / (ms_per_step * 1e-3) / 1e9
    mat_bytes = 4 * (ta.numel() + ra.numel() + td.numel() + rd.numel())
    alg_bytes = acq.data.nbytes + mat_bytes + out.numel() * 8
    achieved = alg_bytes / (kernel_ms * 1e-3) / 1e9
    copy = copy_probe(torch, dev)
    log(f"matrix ABI {args.steps} steps: {ms_per_step:.4f} ms/step (kernel {kernel_ms:.4f} ms), "
        f"{achieved:.0f} GB/s of {copy['GBs']:.0f} achievable")
    traffic = _traffic(f"das_matrix_{args.config.lower()}_traffic.txt")
    result = {
        "metric": "DAS Gsamples/s (pixel x channel x wave)",
        "value": round(value, 3),
        "unit": "Gsamples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": acq.name, "abi": "matrix (ustbx_das_run, das_c contract)", "pixels": acq.P,
                   "channels": acq.probe.N_elements, "waves": len(acq.sequence), "frames": acq.n_frames,
                   "samples": acq.n_samples, "interp": "cubic" if interp == 1 else "linear", "dimension": dim,
                   "split": args.split or "auto"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel_ms": round(kernel_ms, 4),
                     "algorithmic_bytes": alg_bytes, "achievable_copy_GBs": copy["GBs"],
                     "frac_of_achievable": round(achieved / copy["GBs"], 4)},
    }
    result["host_entry"] = host_entry(acq, [t.cpu().numpy() for t in (ta, ra, td, rd)], dim, interp)
    if not args.no_check:
        pix = _sample_pixels(acq.P, 2048, seed=5)
        sel = lambda t: np.ascontiguousarray(t.cpu().numpy()[:, pix])
        ref = O.das_c(acq.data, acq.fs, acq.t0, sel(ta), sel(ra), sel(td), sel(rd), acq.fd, dim)
        if interp == 1:
            # das_c is linear: its parity from a linear call on the same
            # matrices; the cubic image against the FP64 beamformer with das_c's
            # FP32 index
            got_c = out.cpu().numpy()[..., pix]
            ref64 = O.beamform_fp64(acq.data, acq.fs, acq.t0, sel(ta), sel(ra), sel(td), sel(rd), acq.fd, dim,
                                    interp="cubic", w0=O.dasc_wd(acq.fd), index="fp32")
            ops.das(ch, acq.fs, acq.t0, ta, ra, td, rd, acq.fd, dim, interp=0, split=args.split, out=out)
            torch.cuda.synchronize()
        got = out.cpu().numpy()[..., pix]
        peak = float(np.nanmax(np.abs(ref)))
        result["errors"] = {"pixels_sampled": int(pix.size),
                            "vs_dasc" if interp == 0 else "linear_vs_dasc": float(np.nanmax(np.abs(got - ref)) / peak),
                            "bitwise_vs_dasc" if interp == 0 else "linear_bitwise_vs_dasc":
                                bool(np.array_equal(got, ref, equal_nan=True))}
        if interp == 1:
            result["errors"]["cubic_vs_fp64_dasc_index"] = float(np.nanmax(np.abs(got_c - ref64)) /
                                                                 float(np.nanmax(np.abs(ref64))))
        log(f"errors {result['errors']}")
    print(json.dumps(result))


def host_entry(acq, mats, dim, interp, runs=5):
    """The same call through ustbx_das_run_host, the MEX gateway's entry: host
    (pageable) channel data and matrices, as MATLAB hands them over, copied in,
    beamformed and copied back per call (through the entry's pinned staging
    and its device buffer kept across calls). Median of `runs` calls after
    one warm-up: the PCIe-inclusive rate of the drop-in boundary, with the
    pinned host-to-device copy of the same bytes as its floor."""
    import ctypes

    from ustb_amd import _lib as L

    ta, ra, td, rd = (np.ascontiguousarray(m) for m in mats)
    F, W, N, Ls = acq.data.shape
    from ustb_amd import ops

    out = np.zeros(ops.out_shape(dim, acq.P, N, W, F), np.complex64)
    d = L.DasDesc(L=Ls, N=N, W=W, F=F, P=acq.P, fs=acq.fs, t0=acq.t0, fd=acq.fd, dim=dim, interp=interp, edge=0,
                  flags=0, split=0)
    ch = np.ascontiguousarray(acq.data)
    times = []
    for k in range(runs + 1):
        st = L.Status()
        t = time.perf_counter()
        rc = L.lib().ustbx_das_run_host(ctypes.byref(d), ch.ctypes.data, ta.ctypes.data, ra.ctypes.data,
                                        td.ctypes.data, rd.ctypes.data, out.ctypes.data, 0, ctypes.byref(st))
        L.check(rc, st)
        if k:
            times.append(time.perf_counter() - t)
    ms = float(np.median(times)) * 1e3
    host_bytes = ch.nbytes + sum(m.nbytes for m in (ta, ra, td, rd)) + out.nbytes
    # the floor: one pinned host-to-device copy of the input bytes
    import torch

    in_bytes = host_bytes - out.nbytes
    pin = torch.empty(in_bytes, dtype=torch.uint8, pin_memory=True)
    dst = torch.empty(in_bytes, dtype=torch.uint8, device="cuda")
    dma = []
    for k in range(runs + 1):
        torch.cuda.synchronize()
        t = time.perf_counter()
        dst.copy_(pin, non_blocking=True)
        torch.cuda.synchronize()
        if k:
            dma.append(time.perf_counter() - t)
    dma_ms = float(np.median(dma)) * 1e3
    log(f"host entry: {ms:.3f} ms per call ({host_bytes / 1e6:.1f} MB over the host link); "
        f"pinned DMA of the inputs alone {dma_ms:.3f} ms")
    return {"ms": round(ms, 3), "Gsamples_s": round(acq.units / (ms * 1e-3) / 1e9, 3), "host_bytes": host_bytes,
            "pinned_dma_ms": round(dma_ms, 3),
            "note": "pageable host buffers copied in and out per call through two pinned 8 MiB staging buffers "
                    "(threaded CPU copy overlapping the DMA), device buffer kept across calls"}


def run_batch(args, torch, dev):
    """C4: the ragged batch in one persistent launch (hipGraph replay) against
    the plain loop of one launch per acquisition on the same buffers."""
    from ustb_amd import configs
    from ustb_amd.batch import Batch
    from ustb_amd.geometry import make_plan

    t = time.time()
    wl = configs.c4()
    log(f"{wl.name}: {len(wl.acquisitions)} acquisitions of {len(wl.configs)} configurations, "
        f"data in {time.time() - t:.1f}s")
    plans = []
    for cfg in wl.configs:
        t = time.time()
        p = make_plan(cfg.probe, cfg.scan, cfg.sequence, separable_phase=True, **cfg.plan_kwargs())
        torch.cuda.synchronize()
        log(f"plan {cfg.name}: variant {p.variant}, {p.device_bytes / 2**30:.2f} GiB, {time.time() - t:.1f}s")
        plans.append(p)
    dsets = [[torch.from_numpy(d).to(dev) for d in ds] for ds in wl.datasets]
    chs = [dsets[k][d] for k, d in wl.acquisitions]
    bplans = [plans[k] for k, _ in wl.acquisitions]
    b = Batch(bplans)
    outs = b.set_io(chs)
    stream = torch.cuda.current_stream(dev)

    def timed(fn, reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        e0.record(stream)
        for _ in range(reps):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps, (time.perf_counter() - t0) * 1e3 / reps

    # tile shape: the default (work-weighted) and neighbours, timed on the batch
    shapes = {}
    default = (b.shape["wave_chunk"], b.shape["split"])
    for wc, ws in dict.fromkeys([default, (8, 4), (16, 4), (16, 8), (8, 8), (4, 8), (5, 5), (6, 4), (10, 4)]):
        b.set_tiles(wc, ws)
        b.launch()
        shapes[(wc, ws)] = timed(b.launch, 1)[0]
        log(f"batch tiles ({wc},{ws}): {shapes[(wc, ws)]:.2f} ms")
    best = min(shapes, key=shapes.get)
    b.set_tiles(*best)
    for _ in range(args.warmup):
        b.replay()
    batch_ms, batch_wall = timed(b.replay, args.steps)
    plain_outs = [torch.empty_like(o) for o in outs]
    # the plain loop runs every plan with its own autotuned tiles
    from ustb_amd import autotune
    for k in sorted(set(k for k, _ in wl.acquisitions)):
        i = next(i for i, (kk, _) in enumerate(wl.acquisitions) if kk == k)
        best_k, _ = autotune.tune(plans[k], chs[i], reps=1)
        log(f"plain loop tiles for {wl.configs[k].name}: {best_k}")

    def plain():
        for p, c, o in zip(bplans, chs, plain_outs):
            p.run(c, o)
    for _ in range(args.warmup):
        plain()
    plain_ms, plain_wall = timed(plain, args.steps)
    log(f"batch {batch_ms:.2f} ms (wall {batch_wall:.2f}), plain loop {plain_ms:.2f} ms (wall {plain_wall:.2f})")

    # parity: the batch against the plain loop (tile shapes may differ:
    # FP32 regrouping), and bit for bit against plan runs with its own tiles
    err = 0.0
    for o, q in zip(outs, plain_outs):
        peak = float(q.abs().max().item())
        err = max(err, float((o - q).abs().max().item()) / max(peak, 1e-30))
    bitwise = True
    seen = set()
    for i, (k, _) in enumerate(wl.acquisitions):
        if k in seen:
            continue
        seen.add(k)
        p = bplans[i]
        v = p.variant
        p.set_tiles(*best)
        ref = p.run(chs[i])
        bitwise = bitwise and bool(torch.equal(ref, outs[i]))
        p.set_tiles(v[1], v[2])
    units = wl.units
    act_k = {k: plans[k].active_terms() for k in set(k for k, _ in wl.acquisitions)}
    act = sum(act_k[k] for k, _ in wl.acquisitions)
    fl = sum(flops_per_unit(int(wl.configs[k].interpolation), wl.configs[k].fd != 0.0) * act_k[k]
             for k, _ in wl.acquisitions)
    achieved = fl / (batch_ms * 1e-3) / 1e12
    in_bytes = sum(int(c.numel()) * 8 for c in chs)
    out_bytes = sum(int(o.numel()) * 8 for o in outs)
    copy = copy_probe(torch, dev)
    gbs = (in_bytes + out_bytes) / (batch_ms * 1e-3) / 1e9
    kinds = {}
    for k, _ in wl.acquisitions:
        kinds[wl.configs[k].name] = kinds.get(wl.configs[k].name, 0) + 1
    result = {
        "metric": "DAS Gsamples/s (pixel x channel x wave)",
        "value": round(units / (batch_wall * 1e-3) / 1e9, 3),
        "unit": "Gsamples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(batch_wall, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": wl.name, "acquisitions": len(wl.acquisitions), "mix": kinds,
                   "units_per_batch": units, "kernel_variant": "batch (persistent, hipGraph)",
                   "waves_per_split": best[0], "splits": best[1], "blocks": b.shape["blocks"],
                   "work_items": b.shape["items"]},
        "roofline": {"bound": "valu", "achieved": round(achieved, 3), "peak": VALU_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / VALU_PEAK_TFLOPS, 4), "traffic": None, "active_units": act,
                     "kernel_ms": round(batch_ms, 3), "peak_nonpacked": VALU_NONPACKED_TFLOPS,
                     "frac_nonpacked": round(achieved / VALU_NONPACKED_TFLOPS, 4)},
        "hbm": {"algorithmic_bytes": in_bytes + out_bytes, "achieved_GBs": round(gbs, 1), "peak_GBs": HBM_PEAK_GBS,
                "frac": round(gbs / HBM_PEAK_GBS, 5), "achievable_copy_GBs": copy["GBs"]},
        "plain_loop": {"ms": round(plain_wall, 3), "kernel_ms": round(plain_ms, 3),
                       "Gsamples_s": round(units / (plain_wall * 1e-3) / 1e9, 3),
                       "batch_speedup": round(plain_wall / batch_wall, 3)},
        "tile_shapes_ms": {f"{k[0]}x{k[1]}": round(v, 3) for k, v in shapes.items()},
        "errors": {"batch_vs_plain": err, "batch_bitwise_vs_plan_same_tiles": bitwise},
    }
    if not args.no_cpu_baseline:
        cpwc = [k for k, c in enumerate(wl.configs) if "CPWC" in c.name]
        cfg = wl.configs[cpwc[0]] if cpwc else wl.configs[0]
        cfg.data = wl.datasets[wl.configs.index(cfg)][0]
        result["cpu_baseline"] = cpu_baseline(cfg, plans[wl.configs.index(cfg)], torch)
        result["cpu_baseline"]["sample"] = f"{cfg.name}: " + result["cpu_baseline"]["sample"]
    b.close()
    for p in plans:
        p.close()
    print(json.dumps(result))


def idle_settings(spec: str) -> dict:
    """--idle 'slice_ms,long_slices,depth' (or 'none') as Server keywords."""
    if spec == "none":
        return dict(idle_slice_ms=None, idle_long_slices=None, idle_depth=None)
    sl, lg, dp = spec.split(",")
    return dict(idle_slice_ms=float(sl), idle_long_slices=int(lg), idle_depth=int(dp))


def run_serve(args, torch, dev):
    """C5: the continuous-batching server against the plain loop (one request
    at a time, one launch per frame) on the same Poisson trace."""
    from ustb_amd import configs
    from ustb_amd.geometry import make_plan
    from ustb_amd.serve import Server, poisson_trace

    t = time.time()
    cfgs, datasets = configs.c5()
    plans = [make_plan(c.probe, c.scan, c.sequence, separable_phase=True, **c.plan_kwargs()) for c in cfgs]
    if args.stage >= 0:
        for p in plans:
            p.set_staging(bool(args.stage))
    torch.cuda.synchronize()
    log(f"C5: {len(cfgs)} configurations, data and plans in {time.time() - t:.1f}s")
    srv = Server(plans, datasets, hbm_budget=64 << 30, slice_ms=args.slice_ms,
                 tiles=(args.wave_chunk or 16, args.split or 4),
                 long_slices=args.long_slices, depth=args.depth, **idle_settings(args.idle))
    log(f"calibrated ms per position {[round(x, 4) for x in srv.ms_per_pos]}, positions {srv.positions}")
    names = [c.name for c in cfgs]
    trace = poisson_trace(args.frames, rate=args.rate, n_configs=len(cfgs), n_data=len(datasets[0]), seed=5)
    warm = poisson_trace(min(30, args.frames), rate=args.rate, n_configs=len(cfgs), n_data=len(datasets[0]),
                         seed=6)
    srv.run(warm)
    srv.plain_loop(warm)
    res = srv.run(trace)
    log(f"served {res.summary(names)}")
    plain = srv.plain_loop(trace)
    log(f"plain {plain.summary(names)}")
    # parity: a few requests alone (plan runs with the server's tiles) and in the mix
    chk = poisson_trace(12, rate=args.rate, n_configs=len(cfgs), n_data=len(datasets[0]), seed=7)
    mixed = srv.run(chk, keep_outputs=True)
    plain_chk = srv.plain_loop(chk, keep_outputs=True)
    same, err = True, 0.0
    for r in chk:
        p = plans[r.config]
        v = p.variant
        p.set_tiles(*srv.tiles)
        alone = p.run(torch.from_numpy(datasets[r.config][r.data]).to(dev)).cpu().numpy()
        p.set_tiles(v[1], v[2])
        same = same and bool(np.array_equal(alone, mixed.outputs[r.rid]))
        q = plain_chk.outputs[r.rid]
        err = max(err, float(np.max(np.abs(q - alone)) / max(np.max(np.abs(alone)), 1e-30)))
    units = sum(cfgs[r.config].units for r in trace)
    result = {
        "metric": "DAS Gsamples/s (pixel x channel x wave)",
        "value": round(units / res.elapsed_s / 1e9, 3),
        "unit": "Gsamples/s",
        "n_gpus": 1,
        "steps": len(trace),
        "warmup": len(warm),
        "ms_per_step": round(res.elapsed_s * 1e3 / len(trace), 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"C5 serving, Poisson {args.rate:g} frames/s, first {len(trace)} requests of the "
                               f"trace, uniform over {names}", "hbm_budget_GiB": 64, "slice_ms": args.slice_ms, "long_slices": args.long_slices,
                   "depth": args.depth, "idle_slice_long_depth": args.idle,
                   "waves_per_split": srv.tiles[0], "splits": srv.tiles[1],
                   "units_per_frame": {n: c.units for n, c in zip(names, cfgs)}},
        "serving": res.summary(names),
        "plain_loop": plain.summary(names),
        "serving_speedup_frames_per_s": round(res.frames_per_s / plain.frames_per_s, 3),
        "errors": {"served_equals_alone_bitwise": same, "plain_vs_alone": err},
    }
    srv.close()
    for p in plans:
        p.close()
    print(json.dumps(result))


def _sample_pixels(P, k, seed=0):
    return np.sort(np.random.default_rng(seed).choice(P, size=min(k, P), replace=False))


def check_errors(acq, plan, out, ch, torch):
    """das_c parity (linear mode, das_c has no cubic) and FP64 error of the
    benchmarked configuration, on a pixel sample, relative to the peak image
    magnitude."""
    from oracle import oracle as O
    from ustb_amd.geometry import make_plan

    t = time.time()
    ta, ra, td, rd = [x.cpu().numpy() for x in plan.materialize()]
    pix = _sample_pixels(acq.P, 1024, seed=5)
    img = out.cpu().numpy()
    peak = float(np.nanmax(np.abs(img)))
    res = {"pixels_sampled": int(pix.size)}
    wd = O.dasc_wd(acq.fd)
    ref64 = O.beamform_fp64(acq.data, acq.fs, acq.t0, ta, ra, td, rd, acq.fd, int(acq.dimension),
                            interp="cubic" if acq.interpolation == 1 else "linear", pixels=pix, w0=wd)
    res["vs_fp64"] = float(np.nanmax(np.abs(img[..., pix] - ref64)) / peak)
    # linear mode of the same plan (same phase arithmetic) vs das_c
    lin = make_plan(acq.probe, acq.scan, acq.sequence, separable_phase=plan.separable,
                    **acq.plan_kwargs(interpolation=0))
    img_lin = lin.run(ch).cpu().numpy()
    lin.close()
    peak_lin = float(np.nanmax(np.abs(img_lin)))
    ref_c = O.das_c(acq.data, acq.fs, acq.t0, ta[:, pix], ra[:, pix], td[:, pix], rd[:, pix], acq.fd,
                    int(acq.dimension))
    res["vs_dasc_linear"] = float(np.nanmax(np.abs(img_lin[..., pix] - ref_c)) / peak_lin)
    ref64l = O.beamform_fp64(acq.data, acq.fs, acq.t0, ta, ra, td, rd, acq.fd, int(acq.dimension),
                             pixels=pix)
    res["dasc_linear_vs_fp64_exact_w0"] = float(np.nanmax(np.abs(ref_c - ref64l)) / peak_lin)
    res["check_s"] = round(time.time() - t, 1)
    print(f"[bench] errors {res}", file=sys.stderr, flush=True)
    return res


def host_cpus() -> dict:
    """The host CPUs this process may use: the affinity mask, capped by the
    cgroup's CPU quota where one is set (a share of a larger machine), with
    nproc and the CPU model for the record."""
    aff = len(os.sched_getaffinity(0))
    quota = None
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            q, per = f.read().split()[:2]
        if q != "max":
            quota = max(1, int(int(q) // int(per)))
    except (OSError, ValueError):
        pass
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "affinity": aff, "cgroup_quota": quota,
            "threads": min(aff, quota) if quota else aff, "model": model}


def cpu_baseline(acq, plan, torch, gpu_linear=None):
    """The das_c restatement (C, pthreads over pixels like das_c's
    tbb::parallel_for) timed on the host on a pixel sample of the same
    workload, linear interpolation (das_c has no cubic), with every host CPU
    this process may use; median of 5 runs after 1 warm-up. gpu_linear: the
    GPU's rate on the same workload with linear interpolation, when the
    headline runs another interpolation, so that the two can be compared."""
    from oracle import oracle as O

    cpus = host_cpus()
    threads = cpus["threads"]
    ta, ra, td, rd = [x.cpu().numpy() for x in plan.materialize()]
    # bounded sample: every k-th pixel column so that one run is ~2-4 s
    per_unit_s = 1.0 / (60e6 * threads)                   # rough prior
    target_s = 2.5
    frac = min(1.0, target_s / max(acq.units * per_unit_s, 1e-9))
    k = max(1, int(round(acq.P * frac)))
    pix = _sample_pixels(acq.P, k, seed=7)
    args = (acq.data, acq.fs, acq.t0, np.ascontiguousarray(ta[:, pix]), np.ascontiguousarray(ra[:, pix]),
            np.ascontiguousarray(td[:, pix]), np.ascontiguousarray(rd[:, pix]), acq.fd, int(acq.dimension))
    O.das_c(*args, nthreads=threads)                      # warm-up
    times = []
    for _ in range(5):
        t = time.perf_counter()
        O.das_c(*args, nthreads=threads)
        times.append(time.perf_counter() - t)
    med = float(np.median(times))
    units = pix.size * acq.probe.N_elements * len(acq.sequence) * acq.n_frames
    res = {"value": round(units / med / 1e9, 4), "unit": "Gsamples/s", "cores": threads, "kind": "port",
           "sample": f"{pix.size} of {acq.P} pixels (all channels, waves, frames), linear interpolation, "
                     f"median of 5 runs after 1 warm-up, {med:.2f} s per run",
           "host": cpus}
    if gpu_linear is not None:
        res["gpu_linear_value"] = gpu_linear
        res["gpu_linear_over_cpu"] = round(gpu_linear / res["value"], 1)
    return res


def gpu_linear_rate(acq, ch, args, torch, dev):
    """Gsamples/s of the geometry plan with linear interpolation on the same
    workload (kernel time by HIP events over the steps)."""
    from ustb_amd.geometry import make_plan

    lin = make_plan(acq.probe, acq.scan, acq.sequence, separable_phase=(args.phase == "separable"),
                    **acq.plan_kwargs(interpolation=0))
    out = torch.empty(lin.out_shape, dtype=torch.complex64, device=dev)
    for _ in range(max(1, args.warmup)):
        lin.run(ch, out)
    stream = torch.cuda.current_stream(dev)
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    ev0.record(stream)
    for _ in range(args.steps):
        lin.run(ch, out)
    ev1.record(stream)
    torch.cuda.synchronize()
    ms = ev0.elapsed_time(ev1) / args.steps
    lin.close()
    return round(acq.units / (ms * 1e-3) / 1e9, 3)


if __name__ == "__main__":
    main()
