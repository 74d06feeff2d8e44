#!/usr/bin/env python3
"""Benchmark of the reconstruction hot path on one MI355X.

Workload (C1): 2 detector planes x 2000 frames; fan sinogram 768 detectors x
1000 source angles (uint16), rebinned to 512 x 1000, Shepp-Logan filter,
512^2 FP32 image with the circular mask.  One step = the whole job (4000
frame-planes = "slices") through Attenuation -> Fan2Para -> Filter ->
Backprojection(+mask), inputs already in HBM.

value       : plain FP32 chain (frame by frame, FFT filter, FP32 BP), slices/s.
mixed       : mixed-precision batched chain (DFT-GEMM filter on MFMA with
              FP16 output, FP16-sinogram / FP32-accumulate BP), slices/s and
              its relative RMSE / max error against the FP64 oracle.
roofline    : the dominant kernel (backprojection), FP32-ALU bound.
cpu_baseline: the faithful-FP32 CPU restatement (the oracle) on a bounded
              sample, on this host's cores.

python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "512² slices/sec on 1 GPU (mixed vs plain FP32) at stated RMSE; BP HBM GB/s"
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
FLOP_PER_UPDATE = 10          # t: mul+fma; two weights; two accumulate FMAs
BP_SLICES_IN_FLIGHT = 32      # the plain chain keeps 2 lanes x 16 frames in flight


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_frames(g, n_unique_per_plane, seed_base, dead):
    from risa_amd import synth

    planes = g.numberOfPlanes
    raws, pl = [], []
    for f in range(n_unique_per_plane):
        for k in range(planes):
            raws.append(synth.fan_frame(g, k, seed_base + f, 77 + f * planes + k, dead=dead))
            pl.append(k)
    return np.stack(raws), np.array(pl, np.int32)


def disk(N):
    c = (N - 1) * 0.5
    y, x = np.mgrid[0:N, 0:N].astype(np.float32)
    return ((x - np.float32(c)) ** 2 + (y - np.float32(c)) ** 2).astype(np.float64) <= N * N * 0.25


def rel_rmse(a, ref, m):
    a = np.asarray(a, np.float64)[m]
    ref = np.asarray(ref, np.float64)[m]
    return float(np.sqrt(np.mean((a - ref) ** 2)) / np.sqrt(np.mean(ref ** 2)))


def kernel_only(args, g, dev):
    """One kernel launched back to back on one stream, timed with HIP events on
    that stream: the figure a rocprofv3 kernel trace of the same command must
    agree with.  `copy` is a plain device copy of known size, the scale for
    the FETCH_SIZE / WRITE_SIZE counters."""
    import torch
    import risa_amd
    from risa_amd import device

    N, P, D = g.numberOfPixels, g.numberOfParallelProjections, g.numberOfParallelDetectors
    stream = torch.cuda.Stream()
    rng = np.random.default_rng(1)
    with torch.cuda.stream(stream):
        if args.kernel in ("bp", "bp16"):
            # every dispatch of the run has the timed shape (the autotuner's
            # included), so per-kernel counter averages are per such launch
            bp = device.Backprojection(g)
            if args.kernel == "bp":
                S = args.slices or BP_SLICES_IN_FLIGHT
                bp.autotune_batch(S)
                if args.variant is not None:
                    bp.variant = (bp.variant & 0xff00) | args.variant
                sino = torch.from_numpy((rng.standard_normal((S, P, D)) * 0.01).astype(np.float32)).to(dev)
                img = torch.empty((S, N, N), dtype=torch.float32, device=dev)
                fn, per = (lambda: bp.batch(sino, S, img)), S
                what = f"bp FP32, {S} slices {N}^2 x {P} angles per launch, variant {bp.variant & 0xff}"
            else:
                B = args.slices or 32
                bp.autotune_f16(B)
                if args.variant is not None:
                    bp.variant = (bp.variant & 0xff) | (args.variant << 8)
                sino = torch.from_numpy((rng.standard_normal((B // 4, P, D, 4)) * 2.0).astype(np.float16)).to(dev)
                img = torch.empty((B, N, N), dtype=torch.float32, device=dev)
                fn, per = (lambda: bp.f16x4(sino, 256.0, B, img)), B
                what = f"bp FP16 x4 layout, {B} slices {N}^2 x {P} per launch, variant {bp.variant >> 8}"
        elif args.kernel == "adj":
            # the differentiable FBP's backward projector at its C4 shape:
            # 16 slices of 512^2, 720 angles, 512 detectors per batched call
            ga = g.replace(numberOfParallelDetectors=512, numberOfParallelProjections=720, numberOfPixels=512,
                           filterType=0, fusedMasking=1)
            S4, P4, D4, N4 = 16, 720, 512, 512
            bpa = device.Backprojection(ga)
            if args.adjoint_gather:
                bpa.adjoint_method(1)
            im4 = torch.from_numpy(rng.standard_normal((S4, N4, N4)).astype(np.float32)).to(dev)
            gs4 = torch.empty((S4, P4, D4), dtype=torch.float32, device=dev)
            fn, per = (lambda: bpa.adjoint_batch(im4, S4, gs4)), S4
            what = (f"adjoint projector ({'gather' if args.adjoint_gather else 'sheared paths'}), {S4} slices of "
                    f"{N4}^2 -> {P4} x {D4} per call")
        elif args.kernel in ("att", "att_mm", "reorder"):
            # 32 C1 frames per launch: attenuation of detector-major frames,
            # the same reading ROFEX module-major frames through the fused
            # reordering, and the standalone reordering
            from risa_amd import synth
            F, Pf, Df = 32, g.numberOfFanProjections, g.numberOfFanDetectors
            avg_ref, avg_dark = synth.prepare_reference(g, synth.reference_frames(g, 4, seed=901))
            raw = torch.from_numpy(rng.integers(100, 60000, size=(F, Pf, Df), dtype=np.uint16)).to(dev)
            planes = np.arange(F, dtype=np.int32) % g.numberOfPlanes
            per = F
            if args.kernel == "reorder":
                out16 = torch.empty_like(raw)
                fn = lambda: device.reorder(raw, F, Pf, Df, out16)
                what = f"reorder module-major -> detector-major, {F} frames of {Pf} x {Df} uint16 per launch"
            else:
                att = device.Attenuation(g, avg_ref, avg_dark)
                if args.kernel == "att_mm":
                    att.input_layout = risa_amd.LAYOUT_MODULE_MAJOR
                out = torch.empty((F, Pf, Df), dtype=torch.float32, device=dev)
                fn = lambda: att.batch(raw, planes, out)
                layout = "module-major (reordering fused)" if args.kernel == "att_mm" else "detector-major"
                what = f"attenuation, {F} frames of {Pf} x {Df}, {layout} input"
        elif args.kernel in ("rebin", "dft"):
            # the mixed chain's front stages alone, 64 frames per launch (C1,
            # both planes; the chain's lane steps run 32): rebinning of
            # attenuated FP32 frames, and the DFT-GEMM filter to the FP16 x4
            # layout
            from risa_amd import synth
            F, Pf, Df = 64, g.numberOfFanProjections, g.numberOfFanDetectors
            P, D = g.numberOfParallelProjections, g.numberOfParallelDetectors
            planes = np.arange(F, dtype=np.int32) % g.numberOfPlanes
            per = F
            if args.kernel == "dft":
                filt = device.Filter(g, mode=risa_amd.FILTER_MODE_GEMM_SPLIT)
                par = torch.from_numpy((rng.standard_normal((F, P, D)) * 0.05).astype(np.float32)).to(dev)
                x4 = torch.empty((F // 4, P, D, 4), dtype=torch.float16, device=dev)
                fn = lambda: filt.f16x4(par, x4, F, 256.0)
                what = f"DFT-GEMM filter (hi+lo FP16 MFMA, FP32 in, FP16 x4 out), {F} slices of {P} x {D}"
            else:
                avg_ref, avg_dark = synth.prepare_reference(g, synth.reference_frames(g, 4, seed=901))
                raw = torch.from_numpy(rng.integers(100, 60000, size=(F, Pf, Df), dtype=np.uint16)).to(dev)
                att = device.Attenuation(g, avg_ref, avg_dark)
                f2p = device.Fan2Para(g)
                par = torch.empty((F, P, D), dtype=torch.float32, device=dev)
                fan = torch.empty((F, Pf, Df), dtype=torch.float32, device=dev)
                att.batch(raw, planes, fan)
                fn = lambda: f2p.batch(fan, planes, par)
                what = f"rebinning, {F} attenuated FP32 frames of {Pf} x {Df} -> {P} x {D}"
        elif args.kernel == "prep":
            # flat-field preparation of 8 reference frames per plane (1 % dead
            # columns): defect detection and repair of each, then the average
            from risa_amd import synth
            R = 8
            dead = synth.dead_columns(g.numberOfFanDetectors, 0.01, 7001)
            ref_np = synth.reference_frames(g, R, seed=901, dead=dead)
            t0 = time.perf_counter()
            synth.prepare_reference(g, ref_np)
            host_ms = (time.perf_counter() - t0) * 1e3
            frames = torch.from_numpy(ref_np).to(dev)
            avg_ref = torch.empty((g.numberOfPlanes,) + g.fan_shape, dtype=torch.float32, device=dev)
            avg_dark = torch.empty((g.numberOfPlanes, g.numberOfFanDetectors), dtype=torch.float32, device=dev)
            fn, per = (lambda: device.prepare_reference(g, frames, R, avg_ref, avg_dark)), 1
            what = (f"flat-field preparation, {R} x {g.numberOfPlanes} frames of {g.numberOfFanProjections} x "
                    f"{g.numberOfFanDetectors} (host path: {host_ms:.1f} ms on 1 core)")
        else:
            n = 64 << 20
            src = torch.ones(n, dtype=torch.float32, device=dev)
            dst = torch.empty_like(src)
            fn, per, what = (lambda: dst.copy_(src)), 1, f"copy {4 * n} bytes read + {4 * n} written"
        for _ in range(5):
            fn()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(args.launches):
            fn()
        e1.record(stream)
    stream.synchronize()
    us = e0.elapsed_time(e1) / args.launches * 1e3
    print(json.dumps({"kernel": args.kernel, "what": what, "launches": args.launches,
                      "us_per_launch": round(us, 2), "us_per_slice": round(us / per, 2)}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--frames", type=int, default=2000, help="frames per plane in one step")
    ap.add_argument("--unique", type=int, default=16, help="distinct seeded frames per plane")
    ap.add_argument("--chunk", type=int, default=512, help="frames per chain call (graph)")
    ap.add_argument("--cpu-frames", type=int, default=256,
                    help="frame-planes in the CPU baseline sample (about 10 s on 16 cores)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-rmse", action="store_true")
    ap.add_argument("--no-stream", action="store_true", help="skip the streamed C++ stage pipeline leg")
    ap.add_argument("--no-uneven", action="store_true",
                    help="skip the uneven-batch (96 geometries) and differentiable-FBP legs")
    ap.add_argument("--stream-frames", type=int, default=4000,
                    help="frame-planes through the stage pipeline (C1: 2 planes x 2000 frames)")
    ap.add_argument("--no-c0", action="store_true", help="skip the C0 GPU / CPU leg")
    ap.add_argument("--kernel", choices=("bp", "bp16", "adj", "att", "att_mm", "reorder", "rebin", "dft", "prep",
                                         "copy"),
                    help="time one kernel alone (for rocprofv3 runs) and exit")
    ap.add_argument("--launches", type=int, default=50, help="launches in --kernel mode")
    ap.add_argument("--variant", type=int, default=None, help="--kernel bp/bp16: this tile variant, not the autotuned one")
    ap.add_argument("--slices", type=int, default=None, help="--kernel bp / bp16: slices per launch (default 32 / 32)")
    ap.add_argument("--adjoint-gather", action="store_true", help="--kernel adj: the gather adjoint kernel")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus > 1 or world > 1:
        print("bench.py measures one GPU: the library has no multi-GPU path "
              "(frames would shard round-robin with no exchange). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import torch
    import risa_amd
    from risa_amd import device, synth

    dev = torch.device("cuda", 0)
    g = risa_amd.geom_from_config(risa_amd.config_path("c1"))
    planes = g.numberOfPlanes
    N = g.numberOfPixels
    fan_n = g.numberOfFanDetectors * g.numberOfFanProjections

    if args.kernel:
        kernel_only(args, g, dev)
        return

    t0 = time.time()
    dead = synth.dead_columns(g.numberOfFanDetectors, 0.01, 7001)
    ref_frames = synth.reference_frames(g, 8, seed=901, dead=dead)
    avg_ref, avg_dark = synth.prepare_reference(g, ref_frames)
    uniq_raw, uniq_planes = make_frames(g, args.unique, 1000, dead)
    log(f"generated {len(uniq_planes)} distinct frames + {len(ref_frames)} reference frames in {time.time() - t0:.1f}s")

    # the job: frames alternate planes; tiled from the distinct set, in HBM
    total = args.frames * planes
    raw_u = torch.from_numpy(uniq_raw).to(dev)
    reps = (total + len(uniq_planes) - 1) // len(uniq_planes)
    raw = raw_u.repeat(reps, 1, 1)[:total].contiguous()
    planes_all = np.tile(uniq_planes, reps)[:total]
    img = torch.empty((total, N, N), dtype=torch.float32, device=dev)
    log(f"job: {total} slices, raw {raw.numel() * 2 / 1e9:.2f} GB + images {img.numel() * 4 / 1e9:.2f} GB in HBM")

    stream = torch.cuda.Stream()
    chunk = args.chunk

    def run_chain(ch):
        with torch.cuda.stream(stream):
            for c0 in range(0, total, chunk):
                c1 = min(total, c0 + chunk)
                ch(raw[c0:c1], planes_all[c0:c1], img[c0:c1], use_graph=True)

    def timed(ch, label):
        for _ in range(args.warmup):
            run_chain(ch)
        stream.synchronize()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t = time.perf_counter()
        e0.record(stream)
        for _ in range(args.steps):
            run_chain(ch)
        e1.record(stream)
        stream.synchronize()
        torch.cuda.synchronize()
        wall = time.perf_counter() - t
        ms = e0.elapsed_time(e1) / args.steps
        log(f"{label}: {ms:.2f} ms/step ({total / (ms / 1e3):.0f} slices/s; wall {wall:.2f}s for {args.steps} steps)")
        return ms

    plain = device.Chain(g, avg_ref, avg_dark, mode=risa_amd.CHAIN_PLAIN, max_frames=chunk)
    log(f"plain chain: backprojector variant {plain.variant & 0xFF} (autotuned)")
    ms_plain = timed(plain, "plain FP32")
    img_plain_sample = img[:planes].cpu().numpy()

    mixed = device.Chain(g, avg_ref, avg_dark, mode=risa_amd.CHAIN_MIXED, max_frames=chunk)
    log(f"mixed chain: FP16 backprojector variant {mixed.variant >> 8} (autotuned)")
    ms_mixed = timed(mixed, "mixed FP16/FP32")
    img_mixed_sample = img[:planes].cpu().numpy()

    # Dominant kernel: backprojection, timed with HIP events on its stream,
    # at the concurrency the plain chain runs it (2 lanes x 16 frames in
    # flight = one launch of 32 slices here, tile shape autotuned for it).
    P, D = g.numberOfParallelProjections, g.numberOfParallelDetectors
    S = BP_SLICES_IN_FLIGHT
    bp = device.Backprojection(g)
    bp.autotune_batch(S)
    rng = np.random.default_rng(1)
    filt_sino = torch.from_numpy((rng.standard_normal((S, P, D)) * 0.01).astype(np.float32)).to(dev)
    bimg = torch.empty((S, N, N), dtype=torch.float32, device=dev)
    reps_bp = 100
    with torch.cuda.stream(stream):
        for _ in range(5):
            bp.batch(filt_sino, S, bimg)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps_bp):
            bp.batch(filt_sino, S, bimg)
        e1.record(stream)
    stream.synchronize()
    bp_ms = e0.elapsed_time(e1) / reps_bp
    bp_flops = FLOP_PER_UPDATE * N * N * P * S
    bp_bytes = (4 * P * D + 4 * N * N) * S
    bp_tflops = bp_flops / (bp_ms * 1e-3) / 1e12
    bp_gbs = bp_bytes / (bp_ms * 1e-3) / 1e9
    bp_variant = bp.variant & 0xFF
    log(f"backprojection FP32: {bp_ms * 1e3:.1f} us per {S}-slice launch ({bp_ms * 1e3 / S:.1f} us/slice), "
        f"{bp_tflops:.1f} TFLOP/s (10 flop/update), {bp_gbs:.1f} GB/s algorithmic, variant {bp_variant}")

    # BP in the mixed chain: FP16 sinogram in the 4-slice layout, 32 slices per launch
    bp16 = None
    try:
        B = 32
        bp.autotune_f16(B)
        s16 = torch.from_numpy((rng.standard_normal((B // 4, P, D, 4)) * 2.0).astype(np.float16)).to(dev)
        im16 = torch.empty((B, N, N), dtype=torch.float32, device=dev)
        with torch.cuda.stream(stream):
            for _ in range(3):
                bp.f16x4(s16, 256.0, B, im16)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            for _ in range(20):
                bp.f16x4(s16, 256.0, B, im16)
            e1.record(stream)
        stream.synchronize()
        us16 = e0.elapsed_time(e1) / 20 / B * 1e3
        bp16 = {"kernel": "bp_linear_f16_kernel", "us_per_slice": round(us16, 2), "slices_per_launch": B,
                "variant": bp.variant >> 8,
                "achieved_tflops": round(FLOP_PER_UPDATE * N * N * P / (us16 * 1e-6) / 1e12, 2)}
        log(f"backprojection FP16 (x4 layout): {us16:.1f} us/slice")
    except Exception as exc:  # reported, not fatal
        log(f"FP16 BP timing skipped: {exc}")

    # Accuracy against the FP64 oracle on one frame per plane (checker only).
    rmse = {}
    if not args.no_rmse:
        from oracle import oracle as orc

        oc = orc.Chain(g, avg_ref, avg_dark)
        m = disk(N)
        pr, mr, pmax, mmax = [], [], [], []
        for k in range(planes):
            ref = oc.f64(uniq_raw[k], int(uniq_planes[k]))
            pr.append(rel_rmse(img_plain_sample[k], ref, m))
            mr.append(rel_rmse(img_mixed_sample[k], ref, m))
            pmax.append(float(np.max(np.abs(img_plain_sample[k][m] - ref[m]))))
            mmax.append(float(np.max(np.abs(img_mixed_sample[k][m] - ref[m]))))
        rmse = {"plain_rel_rmse": max(pr), "plain_max_abs": max(pmax),
                "mixed_rel_rmse": max(mr), "mixed_max_abs": max(mmax),
                "ref_rms": float(np.sqrt(np.mean(ref[m] ** 2)))}
        log(f"accuracy vs FP64 oracle: {rmse}")
        oc.close()

    # Uneven batch (the 96-reconstruction workload: 256^2..768^2 grids,
    # 360..1000 angles, 256..768 detectors, FP32 and FP16 sinograms) in one
    # persistent launch, timed with HIP events; updates counted as N^2 * P.
    uneven = None
    if not args.no_uneven:
        try:
            Ns, Ps = (256, 384, 512, 768), (360, 500, 720, 1000)
            rng_u = np.random.default_rng(96)
            ujobs, upd = [], 0
            for i in range(96):
                Nj, Pj, Dj = Ns[i % 4], Ps[(i // 4) % 4], int(256 + (i * 37) % 513)
                f16 = (i // 16) % 2
                s_dev = torch.randn((Pj, Dj), device=dev) * 0.01
                if f16:
                    s_dev = (s_dev * 256.0).half()
                ujobs.append(dict(sino=s_dev, img=torch.empty((Nj, Nj), device=dev), N=Nj, P=Pj, D=Dj, dtype=f16,
                                  in_scale=256.0 if f16 else 1.0, rotationOffset=float(rng_u.uniform(-45, 45)),
                                  mask=int(i % 3 != 0), maskingValue=0.0))
                upd += Nj * Nj * Pj
            with torch.cuda.stream(stream):
                bb = device.BatchedBackprojection(ujobs)
                bb.autotune()
                for _ in range(2):
                    bb()
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record(stream)
                for _ in range(10):
                    bb()
                e1.record(stream)
            stream.synchronize()
            ums = e0.elapsed_time(e1) / 10
            utf = FLOP_PER_UPDATE * upd / (ums * 1e-3) / 1e12
            uneven = {"jobs": 96, "ms_per_launch": round(ums, 3), "jobs_per_s": round(96 / (ums * 1e-3), 1),
                      "updates_per_launch": upd, "achieved_tflops": round(utf, 2),
                      "frac_of_fp32_peak": round(utf / FP32_PEAK_TFLOPS, 4),
                      "what": "bp_persistent_kernel, 96 uneven jobs (48 FP32 + 48 FP16 sinograms) in one launch"}
            log(f"uneven batch: {uneven}")
            bb.close()
            del ujobs
        except Exception as exc:  # reported, not fatal
            log(f"uneven batch skipped: {exc}")

    # Differentiable FBP (16 slices, 512 detectors x 720 angles -> 512^2):
    # forward = filter + backprojection; backward = ray-driven adjoint
    # projector, filter-tap gradient and conjugate filter, per slice as the
    # autograd function runs them.
    adjoint = None
    if not args.no_uneven:
        try:
            ga = g.replace(numberOfParallelDetectors=512, numberOfParallelProjections=720, numberOfPixels=512,
                           filterType=0, fusedMasking=1)
            S4, P4, D4, N4 = 16, 720, 512, 512
            filt4, bp4 = device.Filter(ga, 0), device.Backprojection(ga)
            x4 = torch.randn((S4, P4, D4), device=dev) * 0.01
            y4 = x4.clone()
            im4 = torch.empty((S4, N4, N4), device=dev)
            gs4 = torch.empty_like(x4)
            dt4 = torch.empty(D4 // 2 + 1, device=dev)

            def fwd():
                y4.copy_(x4)
                filt4(y4, S4)
                bp4.batch(y4, S4, im4)

            def bwd():
                bp4.adjoint_batch(im4, S4, gs4)
                filt4.grad_taps(x4, gs4, S4, dt4)
                filt4.adjoint(gs4, S4)

            times = {}
            with torch.cuda.stream(stream):
                for name, fn in (("forward", fwd), ("backward", bwd)):
                    fn()
                    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                    e0.record(stream)
                    for _ in range(5):
                        fn()
                    e1.record(stream)
                    stream.synchronize()
                    times[name] = e0.elapsed_time(e1) / 5
            adjoint = {"slices": S4, "forward_ms": round(times["forward"], 3), "backward_ms": round(times["backward"], 3),
                       "forward_us_per_slice": round(times["forward"] * 1e3 / S4, 1),
                       "backward_us_per_slice": round(times["backward"] * 1e3 / S4, 1),
                       "what": "C4 geometry; forward = filter + batched backprojection; backward = "
                               "bp_adjoint_path_kernel (batched, 16 slices per pass, 2 per wave) + fft_grad_taps_kernel + "
                               "conjugate filter; one step of the autograd function"}
            log(f"differentiable FBP: {adjoint}")
        except Exception as exc:  # reported, not fatal
            log(f"differentiable FBP timing skipped: {exc}")

    # C1 as its definition reads: every frame streamed from pinned host memory
    # through the C++ stage queue (H2D -> ... -> D2H), all 2 x 2000 frames, on
    # device 0 only (the child sees one device); plain FP32 stages and the
    # mixed-precision stages (precision = "mixed").  PCIe-inclusive, reported
    # beside the device-resident value, never as it.
    streamed = None
    exe = ROOT / "risa_amd" / "bin" / "risa_pipeline"
    if exe.exists() and not args.no_stream:
        import subprocess
        import tempfile

        streamed = {}
        env = dict(os.environ, HIP_VISIBLE_DEVICES="0")
        base_cfg = risa_amd.config_path("c1")
        with tempfile.TemporaryDirectory() as tmp:
            # D2H's reconstruction rate over windows of 1000 images (the
            # reference's 10000 exceeds one run of 4000 frames)
            plain_cfg = Path(tmp) / "c1_plain.cfg"
            plain_cfg.write_text(base_cfg.read_text() + '\nrateInterval = 1000\n')
            mixed_cfg = Path(tmp) / "c1_mixed.cfg"
            mixed_cfg.write_text(base_cfg.read_text() + '\nrateInterval = 1000\nprecision = "mixed"\n')
            # fp32_fused: the Reconstruction stage (H2D .. D2H as one stage,
            # 3 lanes of 32-frame chain calls); its lanes hold up to 96 host
            # frames, so the source replays 64 distinct frames per plane
            legs = (("fp32", plain_cfg, ("--unique", "16")), ("mixed", mixed_cfg, ("--unique", "16")),
                    ("fp32_fused", plain_cfg, ("--unique", "64", "--fused")),
                    ("mixed_fused", mixed_cfg, ("--unique", "64", "--fused")))
            for label, cfg, extra in legs:
                try:
                    r = subprocess.run([str(exe), str(cfg), "--frames", str(args.stream_frames), *extra],
                                       capture_output=True, text=True, timeout=300, env=env)
                    if r.returncode == 0:
                        res = json.loads(r.stdout.strip().splitlines()[-1])
                        res["what"] = ("C++ glados-style stages on device 0, frames streamed from pinned host "
                                       "memory through H2D .. D2H, processorsPerDevice streams per stage"
                                       if res.get("stages") != "fused" else
                                       "the Reconstruction stage on device 0: H2D, the chain and D2H per batch "
                                       "of queued frames, reconstructionLanes streams")
                        streamed[label] = res
                    else:
                        log(f"stage pipeline ({label}) failed: {r.stderr[-500:]}")
                except Exception as exc:  # reported, not fatal
                    log(f"stage pipeline ({label}) skipped: {exc}")
        log(f"streamed stage pipeline: {streamed}")

    # C0: the reference's own CPU-runnable case (1 plane, 100 frames, 432 x 500
    # fan -> 256 x 500 -> 256^2, ramp): the GPU plain chain and the
    # faithful-FP32 CPU chain on all 100 frames.
    c0 = None
    if not args.no_c0:
        try:
            g0 = risa_amd.geom_from_config(risa_amd.config_path("c0"))
            F0 = 100
            ref0 = synth.reference_frames(g0, 8, seed=900)
            ar0, ad0 = synth.prepare_reference(g0, ref0)
            raw0_np = np.stack([synth.fan_frame(g0, 0, f, 55 + f) for f in range(F0)])
            pl0 = np.zeros(F0, np.int32)
            raw0 = torch.from_numpy(raw0_np).to(dev)
            img0 = torch.empty((F0,) + g0.img_shape, dtype=torch.float32, device=dev)
            ch0 = device.Chain(g0, ar0, ad0, mode=risa_amd.CHAIN_PLAIN, max_frames=F0)
            with torch.cuda.stream(stream):
                for _ in range(2):
                    ch0(raw0, pl0, img0, use_graph=True)
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record(stream)
                for _ in range(10):
                    ch0(raw0, pl0, img0, use_graph=True)
                e1.record(stream)
            stream.synchronize()
            ms0 = e0.elapsed_time(e1) / 10
            ch0.close()
            c0 = {"frames": F0, "gpu_ms": round(ms0, 3), "gpu_slices_per_s": round(F0 / (ms0 * 1e-3), 1),
                  "what": "C0 plain FP32 chain, 100 frames per call (one graph), inputs in HBM"}
            if not args.no_cpu:
                from oracle import oracle as orc

                oc0 = orc.Chain(g0, ar0, ad0)
                threads0 = min(16, os.cpu_count() or 1)
                oc0.f32(raw0_np[:1], pl0[:1], threads=1)
                t = time.perf_counter()
                oc0.f32(raw0_np, pl0, threads=threads0)
                dt0 = time.perf_counter() - t
                oc0.close()
                c0.update({"cpu_slices_per_s": round(F0 / dt0, 2), "cpu_cores": threads0, "cpu_kind": "port",
                           "cpu_seconds": round(dt0, 3),
                           "gpu_vs_cpu": round((F0 / (ms0 * 1e-3)) / (F0 / dt0), 1)})
            log(f"C0: {c0}")
        except Exception as exc:  # reported, not fatal
            log(f"C0 leg skipped: {exc}")

    cpu = None
    if not args.no_cpu:
        from oracle import oracle as orc

        oc = orc.Chain(g, avg_ref, avg_dark)
        n = args.cpu_frames
        reps_c = (n + len(uniq_planes) - 1) // len(uniq_planes)
        sample_raw = np.ascontiguousarray(np.tile(uniq_raw, (reps_c, 1, 1))[:n])
        sample_planes = np.ascontiguousarray(np.tile(uniq_planes, reps_c)[:n])
        threads = min(16, os.cpu_count() or 1)
        oc.f32(sample_raw[:1], sample_planes[:1], threads=1)
        t = time.perf_counter()
        oc.f32(sample_raw, sample_planes, threads=threads)
        dt = time.perf_counter() - t
        cpu_model = "unknown"
        try:
            for line in open("/proc/cpuinfo"):
                if line.startswith("model name"):
                    cpu_model = line.split(":", 1)[1].strip()
                    break
        except OSError:
            pass
        cpu = {"value": n / dt, "unit": "slices/s", "cores": threads, "kind": "port",
               "sample": f"{n} C1 frame-planes through the faithful-FP32 oracle chain "
                         f"(own mixed-radix CPU FFT), OpenMP over frames, {dt:.1f}s wall",
               "host": f"{cpu_model}, nproc {os.cpu_count()}"}
        log(f"cpu baseline: {cpu}")
        oc.close()

    # HBM bytes per launch of the backprojector from the committed rocprofv3
    # FETCH_SIZE / WRITE_SIZE runs (scaled by a copy of known size)
    # (per kernel: the profiled run dispatches every variant while autotuning)
    traffic, traffic_kernel = None, device.bp_kernel_name(False, bp_variant)
    tf = ROOT / "profiles" / "bp_c1_hbm_traffic.json"
    if tf.exists():
        tj = json.loads(tf.read_text())
        per = tj.get("bp_kernels", {})
        if traffic_kernel in per:
            traffic = per[traffic_kernel]

    plain_value = total / (ms_plain / 1e3)
    mixed_value = total / (ms_mixed / 1e3)
    out = {
        "metric": METRIC,
        "value": round(plain_value, 1),
        "unit": "slices/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_plain, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": (f"synthetic: {len(uniq_planes)} distinct seeded ellipse-phantom uint16 fan frames "
                 f"({args.unique} per plane, 1% dead detectors) tiled to {planes}x{args.frames} frames resident "
                 f"in HBM; every frame is reconstructed"),
        "config": {"workload": "C1: 2 planes x 2000 frames, fan 768x1000 uint16 -> parallel 512x1000 -> "
                               "512^2 FP32, Shepp-Logan, circular mask; plain FP32 chain frame by frame",
                   "slices_per_step": total, "frames_per_graph": chunk},
        "mixed": {"value": round(mixed_value, 1), "unit": "slices/s", "ms_per_step": round(ms_mixed, 3),
                  "vs_plain": round(mixed_value / plain_value, 3), "dtype": "f16 storage / f32 accumulate"},
        "accuracy_vs_fp64_oracle": rmse,
        "roofline": {"bound": "valu", "resource": "fp32 vector ALU (backprojection is compute-bound; "
                                                  "AI ~850 flop/B >> ridge ~20; its LDS reads run beside "
                                                  "it near their measured ceiling)",
                     "kernel": f"bp_linear_kernel variant {bp_variant}, {S} slices per launch",
                     "achieved": round(bp_tflops, 2), "peak": FP32_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": round(bp_tflops / FP32_PEAK_TFLOPS, 4), "traffic": traffic,
                     "traffic_unit": "bytes/launch (rocprofv3 FETCH_SIZE+WRITE_SIZE, profiles/)",
                     "traffic_kernel": traffic_kernel,
                     "launch_us": round(bp_ms * 1e3, 2), "flop_per_launch": bp_flops},
        "bp_hbm": {"achieved": round(bp_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                   "frac": round(bp_gbs / HBM_PEAK_GBS, 5), "bytes_per_launch": bp_bytes},
        "bp_f16": bp16,
        "uneven_batch": uneven,
        "differentiable_fbp": adjoint,
        "c1_streamed": streamed,
        "c0": c0,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
