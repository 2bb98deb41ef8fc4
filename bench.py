#!/usr/bin/env python3
"""Benchmark of the point-source model-map path on one MI355X.

A step is one pass of the whole path over the synthetic 100-source workload
(200x200 CAR image, 0.05 deg pixels, 37 energies, 8x8 Gauss-Legendre pixel
cubature): livetime gather + exposure contraction, mean-PSF build, and the
model map (cubature, solid angle, map integral, correction factor,
exposure).  All inputs are resident in HBM before the timed region; the
output is a float32 [S, E, H, W] map.

Prints one JSON line.  The CPU baseline is this project's FP64 C/OpenMP
restatement of the reference algorithm (the reference itself does not build
here), timed on the same workload on the host cores.
"""
from __future__ import annotations

import argparse
import gc
import json
import os
import sys
import time

import numpy as np

METRIC = "srcmap pix·E·src evals/s: surrogate vs cubature at ≤1e-4 max rel err, 1 GPU"
HBM_PEAK_GBS = 8000.0
FP32_VALU_PEAK_TFS = 157.3
STAGES = ("cubature", "near", "factor", "emit", "near_emit")


def _log(msg):
    print(msg, file=sys.stderr, flush=True)


def setup(cfg_name: str):
    import torch
    import xtsrcmaps_amd as xa
    from xtsrcmaps_amd import synthetic

    cfg = synthetic.CONFIGS[cfg_name]
    inp = synthetic.make_inputs(cfg)
    ctx = xa.Context(0)
    # one non-default stream for the library and for the timing events
    stream = torch.cuda.Stream()
    torch.cuda.set_stream(stream)
    ctx.set_stream(stream.cuda_stream)
    dev = torch.device("cuda:0")
    t = lambda a, dt=torch.float64: torch.as_tensor(np.ascontiguousarray(a), dtype=dt, device=dev)
    # IRF preparation (per IRF, not per source): normalisation and Aeff
    king = {}
    for side, (params, cos, loge) in inp.king_raw.items():
        king[side] = (t(ctx.normalize_rpsf(params, loge, inp.psf_scale)), t(cos), t(loge))
    aeff = {s: t(ctx.aeff_value(inp.costh, inp.logE, *g)) for s, g in inp.aeff_grid.items()}
    ltf = t(xa.livetime_efficiency_factors(inp.logE, *inp.effic))
    S, E, C = cfg.n_src, cfg.n_energy, len(inp.costh)
    H = W = cfg.npix
    d = dict(
        src=t(inp.src_sph), cube=t(inp.cube), wcube=t(inp.wcube), costh=t(inp.costh),
        logE=t(inp.logE), af=aeff["front"], ab=aeff["back"], ltf=ltf, king=king,
        ecb=torch.empty((S, C), dtype=torch.float64, device=dev),
        wecb=torch.empty((S, C), dtype=torch.float64, device=dev),
        expo=torch.empty((S, E), dtype=torch.float64, device=dev),
        upsf=torch.empty((S, 401, E), dtype=torch.float64, device=dev),
        parint=torch.empty((S, 401, E), dtype=torch.float64, device=dev),
        out=torch.empty((S, E, H, W), dtype=torch.float32, device=dev),
    )
    # the PSF stages on a second context and stream, beside the model map's
    # row-weight half (which needs no PSF); both streams at default priority
    d["stream"] = stream
    d["psf_stream"] = torch.cuda.Stream()
    d["psf_ctx"] = xa.Context(0)
    d["psf_ctx"].set_stream(d["psf_stream"].cuda_stream)
    d["serial"] = False
    torch.cuda.synchronize()
    return cfg, inp, ctx, d


def _profile(name: str):
    here = os.path.dirname(os.path.abspath(__file__))
    path = os.path.join(here, "profiles", name)
    return json.load(open(path)) if os.path.exists(path) else None


def measured_traffic(config: str):
    """HBM bytes per model map from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE passes (scaled by a copy kernel of known size), or None."""
    p = _profile(f"traffic_{config}_modelmap.json")
    return p["traffic_bytes_per_call"] if p else None


def psf_stages(pctx, inp, d, precision=0):
    """Cosine bins, exposure and mean PSF on context pctx (its stream)."""
    pctx.src_cosbins_dev(d["src"], d["cube"], inp.nside, d["ecb"])
    pctx.src_cosbins_dev(d["src"], d["wcube"], inp.nside, d["wecb"])
    pctx.exposure_dev(d["ecb"], d["wecb"], d["af"], d["ab"], d["ltf"], d["expo"], precision)
    pctx.psf_lookup_dev(d["king"]["front"], d["king"]["back"], d["costh"], d["logE"], d["af"],
                        d["ab"], d["ecb"], d["wecb"], d["ltf"], d["expo"], d["upsf"], d["parint"],
                        precision)


def step(ctx, inp, d, precision=0):
    """One pass of the hot path: exposure, mean PSF and the model map.  The
    model map's row-weight half runs on ctx's stream beside the PSF stages on
    the second context's stream; its second half waits for them.  With
    d["serial"], everything runs on ctx's stream, the model map in one call."""
    import xtsrcmaps_amd as xa
    flags = xa.LAYOUT_SEHW | xa.OUT_F32 | xa.COMPUTE_F32
    if d["serial"]:
        psf_stages(ctx, inp, d, precision)
        ctx.point_src_model_map_wcs_dev(inp.wcs, d["src"], d["expo"], d["upsf"], d["parint"],
                                        d["out"], inp.rule, flags=flags)
        return
    sa, sb = d["stream"], d["psf_stream"]
    # the previous step's second half has read expo, upsf and parint
    sb.wait_stream(sa)
    ctx.point_src_model_map_rows_dev(inp.wcs, d["src"], d["expo"].shape[1], inp.rule,
                                     flags=flags)
    psf_stages(d["psf_ctx"], inp, d, precision)
    sa.wait_stream(sb)
    ctx.point_src_model_map_finish_dev(inp.wcs, d["src"], d["expo"], d["upsf"], d["parint"],
                                       d["out"], inp.rule, flags=flags)


def cpu_baseline(cfg, inp, gpu_map):
    """FP64 oracle of the same workload on the host cores (the `port` kind)."""
    here = os.path.dirname(os.path.abspath(__file__))
    sys.path[:0] = [here, os.path.join(here, "tests")]
    from oracle import oracle as orc
    from pipelines import oracle_stages, oracle_rule, oracle_wcs
    orc.lib()
    t0 = time.perf_counter()
    st = oracle_stages(orc, inp)
    w = oracle_wcs(orc, inp.wcs)
    ref = orc.point_src_model_map_wcs(cfg.npix, cfg.npix, w, st["src"], st["expo"], st["upsf"],
                                      st["parint"], oracle_rule(inp))
    dt = time.perf_counter() - t0
    got = gpu_map.transpose(0, 2, 3, 1)
    rel = np.abs(got.astype(np.float64) - ref) / np.abs(ref)
    return dict(value=cfg.evals / dt, unit="evals/s", cores=orc.num_threads(), kind="port",
                sample=f"the full {cfg.name} workload ({cfg.n_src} sources), exposure + mean "
                       f"PSF + model map, FP64, OpenMP schedule(static,16) over sources, "
                       f"{dt:.1f} s"), float(rel.max()), float(np.median(rel))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="c1")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--precision", type=int, default=0, help="0 FP64, 1 BF16x3 contractions")
    ap.add_argument("--serial", action="store_true",
                    help="one stream, the model map in one call (no overlap with the PSF)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the path shards by source; see the design notes)", file=sys.stderr)
        sys.exit(2)

    import torch
    _log(f"bench: setup {args.config}")
    cfg, inp, ctx, d = setup(args.config)
    d["serial"] = args.serial
    _log("bench: warmup")
    for _ in range(args.warmup):
        step(ctx, inp, d, args.precision)
    torch.cuda.synchronize()
    _log(f"bench: timing {args.steps} steps")
    # a cyclic-GC pass of the interpreter (tens of ms with torch loaded) would
    # stall the host thread that enqueues the step; run it before, not inside
    gc.collect()
    gc.disable()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    ev0.record()
    for _ in range(args.steps):
        step(ctx, inp, d, args.precision)
    ev1.record()
    torch.cuda.synchronize()
    ms = ev0.elapsed_time(ev1) / args.steps
    # per-stage device time from HIP events on the library's stream (the
    # stream the model-map kernels are launched on), in a second pass so that
    # the events do not perturb the headline timing, and with the step on one
    # stream: beside the PSF stages the row-weight pass shares the CUs, and
    # its events would time the sharing rather than the kernel
    _log(f"bench: stage timing, {args.steps} steps (one stream)")
    serial = d["serial"]
    d["serial"] = True
    ctx.enable_timing(True)
    for name in STAGES:
        ctx.kernel_times(name)
    for _ in range(args.steps):
        step(ctx, inp, d, args.precision)
    d["serial"] = serial
    stage_ms = {}
    for name in STAGES:
        t, n = ctx.kernel_times(name)
        stage_ms[name] = t / args.steps
    ctx.enable_timing(False)
    gc.enable()
    evals = cfg.evals
    # graded roofline: the model map's compulsory output stream, 4 B per
    # FP32 map value, over the device time of the model-map passes
    mm_s = sum(stage_ms[n] for n in STAGES) / 1e3
    achieved = 4.0 * evals / mm_s / 1e9
    # VALU view of the row-weight pass.  Algorithmic: 4 flops per rule point
    # per eval, the work of a cubature that evaluates every point for every
    # energy; the pass evaluates each point once per pixel for all energies,
    # so this rate can exceed the FP32 peak.  The issued VALU lane operations
    # per second, as a fraction of the FP32 vector peak, come from the
    # committed rocprofv3 counter pass.
    n_points = len(inp.rule)
    valu_tf = 4.0 * n_points * evals / (stage_ms["cubature"] / 1e3) / 1e12
    valu_prof = _profile(f"valu_{cfg.name}_rowweight.json")
    result = {
        "metric": METRIC,
        "value": evals / (ms / 1e3),
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"{cfg.name}: {cfg.n_src} point sources, {cfg.npix}x{cfg.npix} "
                               f"CAR {cfg.cdelt} deg, {cfg.n_energy} energies, 8x8 Gauss-"
                               f"Legendre pixel cubature, FP32 [S,E,H,W] map",
                   "sources": cfg.n_src, "npix": cfg.npix, "energies": cfg.n_energy,
                   "rule": cfg.rule, "contraction": "bf16x3" if args.precision else "fp64",
                   "streams": 1 if args.serial else 2},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(cfg.name),
                     "kernel": "model-map passes (rowweight_kernel, near_rows_kernel, "
                               "hist_integral_kernel, emit_sehw_kernel, "
                               "near_emit_kernel); algorithmic bytes = "
                               "4 B per FP32 map value; time = their summed device time",
                     "stage_ms_per_step": stage_ms},
        "valu": {"kernel": "rowweight_kernel<float>",
                 "algorithmic_tflops": valu_tf, "peak": FP32_VALU_PEAK_TFS, "unit": "TFLOP/s",
                 "algorithmic_flops_per_eval": 4 * n_points,
                 "issued_frac_of_fp32_peak":
                     valu_prof["valu_frac_of_fp32_peak"] if valu_prof else None},
    }
    if not args.no_cpu_baseline:
        _log("bench: cpu baseline (oracle)")
        gpu_map = d["out"].cpu().numpy()
        cb, mx, md = cpu_baseline(cfg, inp, gpu_map)
        result["cpu_baseline"] = cb
        result["accuracy"] = {"max_rel_err_vs_fp64_oracle": mx, "median_rel_err": md}
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
