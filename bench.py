#!/usr/bin/env python3
"""Benchmark of the point-source model-map path on one MI355X.

A step is one pass of the whole path over the workload, with every input
resident in HBM before the timed region and nothing cached from the previous
step: effective area at the livetime cube's cos-theta bins, livetime gather,
exposure contraction, mean-PSF build, and the model map (pixel geometry and
solid angles, cubature, map integral, correction factor, exposure) into a
float32 [S, E, H, W] map.

The default workload is 2,000 point sources on a 400x400 CAR image at
0.05 deg with 37 energies and the 8x8 Gauss-Legendre pixel rule (1.18e10
map values, a 47 GB map).  It is run twice: with the plain FP32 cubature,
and with the learned surrogate for the far pixels whose rule points straddle
one PSF-table node (XTS_SURROGATE).  The surrogate's network is trained on
the device once per image and rule, before the timed region (its training
time is reported).  `value` is the surrogate path's evals/s; the plain
path's is reported beside it.

Prints one JSON line.  The CPU baseline is this project's FP64 C/OpenMP
restatement of the reference algorithm (the reference itself does not build
here), timed on a bounded sample of the same workload on the host cores; its
maps are also the accuracy check of both GPU paths.
"""
from __future__ import annotations

import argparse
import gc
import json
import os
import subprocess
import sys
import time

import numpy as np

METRIC = "srcmap pix·E·src evals/s: surrogate vs cubature at ≤1e-4 max rel err, 1 GPU"
HBM_PEAK_GBS = 8000.0
STAGES = ("cubature", "surrogate", "near", "factor", "emit", "wide", "near_emit")


def _log(msg):
    print(msg, file=sys.stderr, flush=True)


def setup_inputs(inp, with_out=True):
    """The workload's inputs resident on the device, two contexts and streams."""
    import torch
    import xtsrcmaps_amd as xa
    ctx = xa.Context(0)
    # one non-default stream for the library and for the timing events
    stream = torch.cuda.Stream()
    torch.cuda.set_stream(stream)
    ctx.set_stream(stream.cuda_stream)
    # the step recomputes the pixel geometry of the image every call
    ctx.set_tuning("geometry_cache", 0)
    dev = torch.device("cuda:0")
    t = lambda a, dt=torch.float64: torch.as_tensor(np.ascontiguousarray(a), dtype=dt, device=dev)
    # King normalisation happens once per IRF (the reference's IRF loader)
    king = {}
    for side, (params, cos, loge) in inp.king_raw.items():
        king[side] = (t(ctx.normalize_rpsf(params, loge, inp.psf_scale)), t(cos), t(loge))
    aeff_grid = {s: (t(g[0]), t(g[1]), t(np.asarray(g[2])), g[3]) for s, g in inp.aeff_grid.items()}
    ltf = t(xa.livetime_efficiency_factors(inp.logE, *inp.effic))
    S, E, C = inp.src_sph.shape[0], len(inp.logE), len(inp.costh)
    z = lambda *sh: torch.empty(sh, dtype=torch.float64, device=dev)
    d = dict(
        src=t(inp.src_sph), cube=t(inp.cube), wcube=t(inp.wcube), costh=t(inp.costh),
        logE=t(inp.logE), aeff_grid=aeff_grid, af=z(C, E), ab=z(C, E), ltf=ltf, king=king,
        ecb=z(S, C), wecb=z(S, C), expo=z(S, E), upsf=z(S, 401, E), parint=z(S, 401, E),
    )
    if with_out:
        H, W = inp.wcs.naxis1, inp.wcs.naxis2
        d["out"] = torch.empty((S, E, H, W), dtype=torch.float32, device=dev)
    # the PSF stages on a second context and stream, beside the model map's
    # row-weight half (which needs no PSF); both streams at default priority
    d["stream"] = stream
    d["psf_stream"] = torch.cuda.Stream()
    d["psf_ctx"] = xa.Context(0)
    d["psf_ctx"].set_stream(d["psf_stream"].cuda_stream)
    d["serial"] = False
    d["net"] = None
    torch.cuda.synchronize()
    return ctx, d


def setup(cfg_name: str, surrogate: bool = True):
    import torch
    from xtsrcmaps_amd import synthetic
    cfg = synthetic.CONFIGS[cfg_name]
    inp = synthetic.make_inputs(cfg)
    ctx, d = setup_inputs(inp)
    if surrogate:
        from xtsrcmaps_amd import surrogate as sg
        d["net"] = sg.fit(ctx, inp.wcs, inp.rule, seed=1)
    torch.cuda.synchronize()
    return cfg, inp, ctx, d


def _profile(name: str):
    here = os.path.dirname(os.path.abspath(__file__))
    path = os.path.join(here, "profiles", name)
    return json.load(open(path)) if os.path.exists(path) else None


def psf_stages(pctx, inp, d, precision=0):
    """Effective area, cosine bins, exposure and mean PSF on context pctx."""
    pctx.aeff_value_dev(d["costh"], d["logE"], d["aeff_grid"]["front"], d["af"])
    pctx.aeff_value_dev(d["costh"], d["logE"], d["aeff_grid"]["back"], d["ab"])
    pctx.src_cosbins_dev(d["src"], d["cube"], inp.nside, d["ecb"])
    pctx.src_cosbins_dev(d["src"], d["wcube"], inp.nside, d["wecb"])
    pctx.exposure_dev(d["ecb"], d["wecb"], d["af"], d["ab"], d["ltf"], d["expo"], precision)
    pctx.psf_lookup_dev(d["king"]["front"], d["king"]["back"], d["costh"], d["logE"], d["af"],
                        d["ab"], d["ecb"], d["wecb"], d["ltf"], d["expo"], d["upsf"], d["parint"],
                        precision)


def step(ctx, inp, d, flags=None, precision=0):
    """One pass of the hot path (flags: the model map's; default the plain
    FP32 cubature into the [S, E, H, W] FP32 map).  The model map's row-weight half runs on
    ctx's stream beside the PSF stages on the second context's stream; its
    second half waits for them.  With d["serial"], everything runs on ctx's
    stream, the model map in one call."""
    if flags is None:
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


def step_batched(ctx, inp, d, flags=None, precision=0):
    """One pass with the batched engine: the PSF stages, then every image's
    model map as a cost-ordered list of source batches on two streams into
    the footprint blocks of d["flat"]."""
    import xtsrcmaps_amd as xa
    if flags is None:
        flags = xa.LAYOUT_SEHW | xa.OUT_F32 | xa.COMPUTE_F32
    psf_stages(ctx, inp, d, precision)
    ctx.point_src_model_maps_batched_dev(d["wcs_list"], d["src_begin"], d["src"], d["expo"],
                                         d["upsf"], d["parint"], d["boxes"], d["flat"], inp.rule,
                                         flags=flags)


def setup_batched(ctx, inp, d, wcs_list, src_begin, containment=0.995):
    """Footprint boxes of the workload (from one pass of its PSF stages) and
    the flat output; the context autotunes and graph-captures the batched
    calls."""
    import torch
    psf_stages(ctx, inp, d)
    torch.cuda.synchronize()
    ctx.synchronize()
    boxes, radius, total = ctx.model_map_footprints(wcs_list, src_begin, d["src"], d["parint"],
                                                    containment)
    d["wcs_list"], d["src_begin"], d["boxes"] = wcs_list, src_begin, boxes
    if "out" in d and d["out"].numel() == total:
        d["flat"] = d["out"].view(-1)
    else:
        d["flat"] = torch.empty(total, dtype=torch.float32, device=d["src"].device)
    dense = sum((src_begin[i + 1] - src_begin[i]) * w.naxis1 * w.naxis2
                for i, w in enumerate(wcs_list)) * d["expo"].shape[1]
    d["footprint"] = {"containment": containment, "radius_deg_min": float(radius.min()),
                      "radius_deg_max": float(radius.max()), "evals": int(total),
                      "clipped_fraction_of_dense": 1.0 - total / dense}
    ctx.set_tuning("autotune", 1)
    ctx.set_tuning("graph", 1)
    torch.cuda.synchronize()


def timed(ctx, inp, d, flags, steps, warmup, precision, fn=None):
    """Warm-up steps, then exactly `steps` steps bracketed by device
    synchronisation, timed with events on the library's stream."""
    import torch
    fn = fn or step
    for _ in range(warmup):
        fn(ctx, inp, d, flags, precision)
    torch.cuda.synchronize()
    # a cyclic-GC pass of the interpreter (tens of ms with torch loaded) would
    # stall the host thread that enqueues the step; run it before, not inside
    gc.collect()
    gc.disable()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    ev0.record()
    for _ in range(steps):
        fn(ctx, inp, d, flags, precision)
    ev1.record()
    torch.cuda.synchronize()
    gc.enable()
    return ev0.elapsed_time(ev1) / steps


def stage_times(ctx, inp, d, flags, steps, precision):
    """Per-stage device time (ms per step) from HIP events on the library's
    stream around each model-map launch, with the step on one stream (beside
    the PSF stages the row-weight pass shares the CUs, and its events would
    time the sharing rather than the kernel)."""
    serial = d["serial"]
    d["serial"] = True
    ctx.enable_timing(True)
    for name in STAGES:
        ctx.kernel_times(name)
    for _ in range(steps):
        step(ctx, inp, d, flags, precision)
    out = {}
    for name in STAGES:
        tot, n = ctx.kernel_times(name)
        out[name] = {"ms": tot / steps, "launches_per_step": n / steps}
    ctx.enable_timing(False)
    d["serial"] = serial
    return out


# ---------------------------------------------------------- the large sky
def step_summed(ctx, inp, d, flags=0, precision=0):
    """One pass of the large-sky path: the PSF stages, then the flux-weighted
    summed model [E, H, W] of every source (xts_point_src_summed_model_dev)."""
    psf_stages(ctx, inp, d, precision)
    ctx.point_src_summed_model_dev(inp.wcs, d["src"], d["flux"], d["expo"], d["upsf"],
                                   d["parint"], d["sum"], inp.rule, flags=flags)


SUM_STAGES = ("cubature", "surrogate", "near", "factor", "sum")


def summed_stage_times(ctx, inp, d, flags, steps, precision):
    """Per-stage device time (ms per step, every batch) from HIP events on the
    library's stream."""
    ctx.enable_timing(True)
    for name in SUM_STAGES:
        ctx.kernel_times(name)
    for _ in range(steps):
        step_summed(ctx, inp, d, flags, precision)
    out = {}
    for name in SUM_STAGES:
        tot, n = ctx.kernel_times(name)
        out[name] = {"ms": tot / steps, "launches_per_step": n / steps}
    ctx.enable_timing(False)
    return out


def cpu_baseline_summed(cfg, inp, flux, gpu_sums: dict, n_src: int):
    """The FP64 oracle on the first n_src sources on the full image (exposure,
    mean PSF, per-source maps, flux-weighted sum); the GPU paths' summed
    models of the same sources against it."""
    here = os.path.dirname(os.path.abspath(__file__))
    sys.path[:0] = [here, os.path.join(here, "tests")]
    from oracle import oracle as orc
    from pipelines import oracle_stages, oracle_rule, oracle_wcs
    orc.lib()
    sub = type(inp)(**{**inp.__dict__, "src_sph": inp.src_sph[:n_src]})
    t0 = time.perf_counter()
    st = oracle_stages(orc, sub)
    maps = orc.point_src_model_map_wcs(cfg.npix, cfg.npix, oracle_wcs(orc, inp.wcs), st["src"],
                                       st["expo"], st["upsf"], st["parint"], oracle_rule(inp))
    ref = np.einsum("shwe,s->ehw", maps, flux[:n_src])
    dt = time.perf_counter() - t0
    del maps
    accuracy = {}
    for k, g in gpu_sums.items():
        rel = (np.abs(g.astype(np.float64) - ref) / np.abs(ref)).reshape(ref.shape[0], -1)
        pmax = rel.max(1)
        accuracy[k] = {"max_rel_err_vs_fp64_oracle": float(pmax.max()),
                       "median_rel_err": float(np.median(rel)),
                       "per_plane_max": [float(x) for x in pmax],
                       "per_plane_median": [float(x) for x in np.median(rel, 1)]}
    model, phys = host_cpu()
    evals = n_src * cfg.n_energy * cfg.npix * cfg.npix
    base = dict(value=evals / dt, unit="evals/s", cores=orc.num_threads(), kind="port",
                cpu_model=model, physical_cores=phys,
                sample=f"the first {n_src} of the {cfg.n_src} sources of {cfg.name} on the full "
                       f"{cfg.npix}x{cfg.npix} image, all {cfg.n_energy} energies ({evals:.3g} "
                       f"evals): exposure + mean PSF + model maps + flux-weighted sum, FP64, "
                       f"OpenMP over sources, {dt:.1f} s")
    return base, accuracy


def main_summed(args, cfg):
    """The large sky: the flux-weighted summed model of every source, plain
    FP32 cubature and the FP8 surrogate."""
    import torch
    import xtsrcmaps_amd as xa
    from xtsrcmaps_amd import synthetic
    inp = synthetic.make_inputs(cfg)
    ctx, d = setup_inputs(inp, with_out=False)
    flux = synthetic.fluxes(cfg)
    d["flux"] = torch.as_tensor(flux, dtype=torch.float64, device="cuda")
    E, H, W = cfg.n_energy, cfg.npix, cfg.npix
    d["sum"] = torch.empty((E, H, W), dtype=torch.float32, device="cuda")
    paths = {"cubature": 0}
    if not args.no_surrogate:
        from xtsrcmaps_amd import surrogate as sg
        d["net"] = sg.fit(ctx, inp.wcs, inp.rule, seed=1)
        paths["surrogate"] = xa.SURROGATE | xa.SURROGATE_FP8
    torch.cuda.synchronize()
    res, sums = {}, {}
    nref = min(args.cpu_sources, 4)
    for name, flags in paths.items():
        _log(f"bench: {name}: warmup {args.warmup}, timing {args.steps} steps")
        ms = timed(ctx, inp, d, flags, args.steps, args.warmup, args.precision, step_summed)
        res[name] = {"ms_per_step": ms, "value": cfg.evals / (ms / 1e3),
                     "stages": summed_stage_times(ctx, inp, d, flags, 2, args.precision)}
        _log(f"bench: {name}: {ms:.1f} ms per step")
        if not args.no_cpu_baseline:
            # the summed model of the CPU sample's sources, for the accuracy check
            out = torch.empty((E, H, W), dtype=torch.float32, device="cuda")
            ctx.point_src_summed_model_dev(inp.wcs, d["src"][:nref].contiguous(),
                                           d["flux"][:nref].contiguous(),
                                           d["expo"][:nref].contiguous(),
                                           d["upsf"][:nref].contiguous(),
                                           d["parint"][:nref].contiguous(), out, inp.rule,
                                           flags=flags)
            torch.cuda.synchronize()
            sums[name] = out.cpu().numpy()
    head = "surrogate" if "surrogate" in res else "cubature"
    ms = res[head]["ms_per_step"]
    st = res[head]["stages"]["sum"]
    # the dominant kernel: the sum pass on the FP32 matrix cores; algorithmic
    # flops = 8 per (source, energy, pixel) (four row weights x table values)
    flops = 8.0 * cfg.evals
    achieved = flops / (st["ms"] / 1e3) / 1e12
    result = {
        "metric": METRIC,
        "value": cfg.evals / (ms / 1e3),
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
        "config": {"workload": f"{cfg.name}: {cfg.n_src} point sources (flux log-uniform over 4 "
                               f"decades), {H}x{W} CAR {cfg.cdelt} deg, {E} energies, 8x8 Gauss-"
                               f"Legendre pixel cubature; output the flux-weighted summed model "
                               f"[{E}, {H}, {W}] FP32; value: the {head} path"
                               + (" (FP8 e4m3 network)" if head == "surrogate" else ""),
                   "sources": cfg.n_src, "energies": E, "rule": "gl8",
                   "per_step": "aeff, livetime gather, exposure, mean PSF, geometry, row "
                               "weights, factors, summed model"},
        "paths": {k: {"ms_per_step": round(v["ms_per_step"], 3), "value": v["value"],
                      "stage_ms": {n: round(x["ms"], 3) for n, x in v["stages"].items()}}
                  for k, v in res.items()},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": 157.3, "unit": "TFLOP/s",
                     "frac": achieved / 157.3, "traffic": None,
                     "kernel": "sum_far_kernel (the summed model's accumulation over sources on "
                               "v_mfma_f32_16x16x4_f32, with its near- and wide-pixel additions); "
                               "algorithmic flops = 8 per (source, energy, pixel); time = the sum "
                               "pass's HIP-event duration per step",
                     "launch_ms": st["ms"] / max(st["launches_per_step"], 1.0),
                     "launches_per_step": st["launches_per_step"]},
    }
    if "surrogate" in res:
        result["surrogate"] = {"speedup_vs_cubature":
                               res["cubature"]["ms_per_step"] / res["surrogate"]["ms_per_step"],
                               "target_speedup": 5.0,
                               "network": "4-64-64-64-1 SiLU; hidden layers FP8 e4m3 on MFMA "
                                          "with per-layer weight scales and per-wave activation "
                                          "scales, FP32 accumulation",
                               "training_s": d["net"].seconds}
    if not args.no_cpu_baseline:
        _log(f"bench: cpu baseline (oracle, {nref} sources)")
        cb, accuracy = cpu_baseline_summed(cfg, inp, flux, sums, nref)
        result["cpu_baseline"] = cb
        result["accuracy"] = accuracy
    print(json.dumps(result), flush=True)


def host_cpu():
    """CPU model and physical core count of this host (lscpu)."""
    model, cores, sockets = "unknown", None, 1
    try:
        txt = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        for line in txt.splitlines():
            k, _, v = line.partition(":")
            v = v.strip()
            if k.strip() == "Model name":
                model = v
            elif k.strip() == "Core(s) per socket":
                cores = int(v)
            elif k.strip() == "Socket(s)":
                sockets = int(v)
    except (OSError, ValueError, subprocess.SubprocessError):
        pass
    return model, (cores * sockets if cores else os.cpu_count())


def cpu_baseline(cfg, inp, gpu_maps: dict, n_src: int):
    """The FP64 oracle (the `port` kind) on the first n_src sources of the
    workload, on the host cores: exposure, mean PSF and model map.  Its maps
    are compared with the GPU paths' maps of the same sources."""
    here = os.path.dirname(os.path.abspath(__file__))
    sys.path[:0] = [here, os.path.join(here, "tests")]
    from oracle import oracle as orc
    from pipelines import oracle_stages, oracle_rule, oracle_wcs
    orc.lib()
    w = oracle_wcs(orc, inp.wcs)
    rule = oracle_rule(inp)
    acc = {k: [0.0, []] for k in gpu_maps}
    dt = 0.0
    chunk = 16
    for a in range(0, n_src, chunk):
        b = min(n_src, a + chunk)
        sub = type(inp)(**{**inp.__dict__, "src_sph": inp.src_sph[a:b]})
        t0 = time.perf_counter()
        st = oracle_stages(orc, sub)
        ref = orc.point_src_model_map_wcs(cfg.npix, cfg.npix, w, st["src"], st["expo"],
                                          st["upsf"], st["parint"], rule)
        dt += time.perf_counter() - t0
        for k, m in gpu_maps.items():
            got = m[a:b].transpose(0, 2, 3, 1).astype(np.float64)
            rel = np.abs(got - ref) / np.abs(ref)
            acc[k][0] = max(acc[k][0], float(rel.max()))
            acc[k][1].append(np.median(rel))
        del ref
    model, phys = host_cpu()
    evals = n_src * cfg.n_energy * cfg.npix * cfg.npix
    base = dict(value=evals / dt, unit="evals/s", cores=orc.num_threads(), kind="port",
                cpu_model=model, physical_cores=phys,
                sample=f"the first {n_src} of the {cfg.n_src} sources of {cfg.name} on the full "
                       f"{cfg.npix}x{cfg.npix} image, all {cfg.n_energy} energies ({evals:.3g} "
                       f"evals): exposure + mean PSF + model map, FP64, OpenMP "
                       f"schedule(static,16) over sources, {dt:.1f} s")
    accuracy = {k: {"max_rel_err_vs_fp64_oracle": v[0],
                    "median_rel_err": float(np.median(v[1]))} for k, v in acc.items()}
    return base, accuracy


DIFFUSE_METRIC = "diffuse srcmap pix·E values/s, 1 GPU"
DIFFUSE_STAGES = ("diffuse_kernel", "diffuse_fft", "diffuse_cols", "diffuse_sample")


def main_diffuse(args, cfg):
    """The diffuse-source model map: crop and resample of the cube, the PSF
    kernel's spectrum, the per-energy FFT convolution, sampling onto the
    model map (xts_diffuse_src_model_map_wcs_dev), FP32 output."""
    import torch
    import xtsrcmaps_amd as xa
    from xtsrcmaps_amd import synthetic
    roi, cube_wcs, cube, upsf, expo = synthetic.diffuse_inputs(cfg)
    E, H, W = cfg.n_energy, roi.naxis1, roi.naxis2
    ctx = xa.Context(0)
    stream = torch.cuda.Stream()
    torch.cuda.set_stream(stream)
    ctx.set_stream(stream.cuda_stream)
    # the step recomputes the model map's solid angles every call
    ctx.set_tuning("geometry_cache", 0)
    dev = torch.device("cuda:0")
    t_cube, t_u, t_e = (torch.as_tensor(a, device=dev) for a in (cube, upsf, expo))
    out = torch.empty((E, H, W), dtype=torch.float32, device=dev)
    torch.cuda.synchronize()

    def one(*_):
        ctx.diffuse_src_model_map_wcs_dev(roi, cube_wcs, t_cube, t_u, t_e, cfg.target, out,
                                          flags=xa.OUT_F32)
    _log(f"bench: diffuse: warmup {args.warmup}, timing {args.steps} steps")
    ms = timed(ctx, None, None, None, args.steps, args.warmup, 0, one)
    ctx.enable_timing(True)
    for name in DIFFUSE_STAGES:
        ctx.kernel_times(name)
    for _ in range(5):
        one()
    torch.cuda.synchronize()
    stages = {name: ctx.kernel_times(name)[0] / 5 for name in DIFFUSE_STAGES}
    ctx.enable_timing(False)
    g = xa.diffuse_geometry(roi, cube_wcs, cfg.target)
    values = E * H * W
    _log(f"bench: diffuse: {ms:.3f} ms per map; stages {stages}")
    # the dominant kernel: the convolution's column pass (forward transform,
    # product with the kernel's spectrum, inverse); algorithmic bytes per
    # launch: its row spectra read and written, the kernel's spectrum read
    qh = g["fft_p"] // 2 + 1
    col_bytes = E * qh * (2 * g["m2"] + g["fft_q"]) * 16
    achieved = col_bytes / (stages["diffuse_cols"] / 1e3) / 1e9
    result = {
        "metric": DIFFUSE_METRIC,
        "value": values / (ms / 1e3),
        "unit": "values/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"{cfg.name}: the diffuse-source model map of a {H}x{W} CAR "
                               f"{cfg.cdelt} deg celestial image at {E} energies over a "
                               f"{cube_wcs.naxis1}x{cube_wcs.naxis2} galactic CAR cube "
                               f"({cfg.cube_cdelt} deg, {cfg.planes} planes), resampled to "
                               f"{cfg.target} deg: grid {g['m1']}x{g['m2']}, kernel "
                               f"{g['nk']}x{g['nk']}, transforms {g['fft_p']}x{g['fft_q']}; "
                               f"FP64 arithmetic, FP32 output",
                   "per_step": "solid angles, crop and resample, PSF kernel and its spectrum, "
                               "convolution, sampling, exposure"},
        "stage_ms": {k: round(v, 4) for k, v in stages.items()},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "fft_cols_kernel<1> (the convolution's column pass); algorithmic "
                               "bytes per launch = E x (P/2+1) x (2 m2 + Q) x 16: the image's "
                               "row spectra read and written, the kernel's spectrum read",
                     "launch_ms": stages["diffuse_cols"], "bytes_per_launch": col_bytes},
    }
    if not args.no_cpu_baseline:
        result["cpu_baseline"], result["accuracy"] = cpu_baseline_diffuse(cfg, out.cpu().numpy())
    print(json.dumps(result), flush=True)


def cpu_baseline_diffuse(cfg, gpu_map, planes=None):
    """The oracle's resample, kernel and sampling around scipy's FFT
    convolution (the reference's FFTW algorithm: per-plane linear convolution
    by FFT) on `planes` of the workload (default all), on the host cores;
    those planes of the GPU map against it."""
    here = os.path.dirname(os.path.abspath(__file__))
    sys.path[:0] = [here, os.path.join(here, "tests")]
    from oracle import oracle as orc
    from scipy.signal import fftconvolve
    from xtsrcmaps_amd import synthetic
    orc.lib()
    roi, cube_wcs, cube, upsf, expo = synthetic.diffuse_inputs(cfg)
    E = cfg.n_energy
    planes = list(range(E)) if planes is None else [p for p in planes if p < E]

    def ow(w):
        return orc.make_wcs((w.naxis1, w.naxis2), (w.crpix1, w.crpix2), (w.crval1, w.crval2),
                            (w.cdelt1, w.cdelt2), w.crota2, "CAR" if w.proj == 0 else "AIT")
    t0 = time.perf_counter()
    g = orc.diffuse_geometry(ow(roi), roi.is_galactic, ow(cube_wcs), cube_wcs.is_galactic,
                             cfg.target)
    img = orc.diffuse_resample(cube, ow(cube_wcs), g, E)
    K = orc.psf_make_kernel(g.nk, g.grid, g.kc1, g.kc2, upsf[:, planes])
    c = (g.nk - 1) // 2
    conv = np.stack([fftconvolve(img[e], K[i], mode="full")[c:c + g.m1, c:c + g.m2]
                     for i, e in enumerate(planes)])
    om = orc.solid_angle(roi.naxis1, roi.naxis2, ow(roi))
    ref = orc.diffuse_sample(conv, g, ow(roi), om, expo[planes])
    dt = time.perf_counter() - t0
    rel = np.abs(gpu_map[planes].astype(np.float64) - ref) / np.abs(ref)
    model, phys = host_cpu()
    n = len(planes) * roi.naxis1 * roi.naxis2
    base = dict(value=n / dt, unit="values/s", cores=orc.num_threads(), kind="port",
                cpu_model=model, physical_cores=phys,
                sample=f"energy planes {planes} of {cfg.name} ({n:.3g} map values): the "
                       f"oracle's crop and resample of all {E} planes, kernel, sampling (C, "
                       f"OpenMP threads) around scipy FFT convolution (one thread), FP64, "
                       f"{dt:.2f} s")
    return base, {"max_rel_err_vs_fp64_port": float(rel.max()),
                  "median_rel_err": float(np.median(rel)), "planes": planes}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="c3")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sources", type=int, default=128,
                    help="sources of the bounded CPU-baseline sample")
    ap.add_argument("--no-surrogate", action="store_true", help="plain cubature only")
    ap.add_argument("--precision", type=int, default=0, help="0 FP64, 1 BF16x3 contractions")
    ap.add_argument("--serial", action="store_true",
                    help="one stream, the model map in one call (no overlap with the PSF)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the path shards by source; see the design notes)", file=sys.stderr)
        sys.exit(2)

    import torch
    import xtsrcmaps_amd as xa
    from xtsrcmaps_amd import synthetic
    base = xa.LAYOUT_SEHW | xa.OUT_F32 | xa.COMPUTE_F32
    sur = not args.no_surrogate
    _log(f"bench: setup {args.config}")
    if args.config in synthetic.DIFFUSE_CONFIGS:
        main_diffuse(args, synthetic.DIFFUSE_CONFIGS[args.config])
        return
    if args.config in synthetic.CONFIGS and synthetic.CONFIGS[args.config].extra.get("flux_decades"):
        main_summed(args, synthetic.CONFIGS[args.config])
        return
    if args.config in synthetic.BATCH_CONFIGS:
        bcfg = synthetic.BATCH_CONFIGS[args.config]
        B = synthetic.make_batch_inputs(bcfg)
        inp = B.phys
        ctx, d = setup_inputs(inp, with_out=False)
        if sur:
            from xtsrcmaps_amd import surrogate as sg
            d["net"] = sg.fit(ctx, B.wcs[-1], inp.rule, seed=1)   # the images share a pixel size
        setup_batched(ctx, inp, d, B.wcs, B.src_begin)
        paths = {"cubature": (step_batched, base)}
        if sur:
            paths["surrogate"] = (step_batched, base | xa.SURROGATE)
        evals = d["footprint"]["evals"]
        workload = (f"{bcfg.name}: {bcfg.n_src} point sources over "
                    + ", ".join(f"{p}x{p} ({n})" for p, n in bcfg.rois)
                    + f" CAR images at {bcfg.cdelt} deg, {bcfg.n_energy} energies, 8x8 Gauss-"
                      f"Legendre pixel cubature, FP32 footprint blocks at 99.5% containment; "
                      f"batched engine")
        cfg = None
    else:
        cfg, inp, ctx, d = setup(args.config, surrogate=sur)
        d["serial"] = args.serial
        setup_batched(ctx, inp, d, [inp.wcs], np.array([0, cfg.n_src], np.int64))
        # "cubature": the library's default for the map (at this size the
        # pipelined source batches); "cubature, one pass": every pass over all
        # sources at once, one after the other (its emit runs alone: the
        # roofline's measurement)
        paths = {"cubature": (step, base),
                 "cubature, one pass": (step, base, {"pipeline_batches": 1}),
                 "batched cubature": (step_batched, base)}
        if sur:
            paths["surrogate"] = (step, base | xa.SURROGATE)
            paths["surrogate, one pass"] = (step, base | xa.SURROGATE, {"pipeline_batches": 1})
            paths["fp8 surrogate"] = (step, base | xa.SURROGATE | xa.SURROGATE_FP8)
            paths["batched surrogate"] = (step_batched, base | xa.SURROGATE)
        evals = cfg.evals
        workload = (f"{cfg.name}: {cfg.n_src} point sources, {cfg.npix}x{cfg.npix} CAR "
                    f"{cfg.cdelt} deg, {cfg.n_energy} energies, 8x8 Gauss-Legendre pixel "
                    f"cubature, FP32 [S,E,H,W] map")
    res, maps = {}, {}
    S = inp.src_sph.shape[0]
    for name, spec in paths.items():
        fn, flags = spec[0], spec[1]
        tuning = spec[2] if len(spec) > 2 else {}
        for k, v in tuning.items():
            ctx.set_tuning(k, v)
        _log(f"bench: {name}: warmup {args.warmup}, timing {args.steps} steps")
        ms = timed(ctx, inp, d, flags, args.steps, args.warmup, args.precision, fn)
        entry = {"ms_per_step": ms, "value": evals / (ms / 1e3)}
        if fn is step_batched:
            entry["batched_shape"] = {k: ctx.get_tuning(k) for k in
                                      ("batch_sources", "rowweight_blocks", "batches")}
        else:
            _log(f"bench: {name}: {ms:.3f} ms per step; stage timing")
            entry["stages"] = stage_times(ctx, inp, d, flags, max(3, min(args.steps, 10)),
                                          args.precision)
        _log(f"bench: {name}: {ms:.3f} ms per step")
        res[name] = entry
        if not args.no_cpu_baseline and cfg is not None and not name.startswith("batched"):
            step(ctx, inp, d, flags, args.precision)
            torch.cuda.synchronize()
            maps[name] = d["out"][:min(args.cpu_sources, S)].cpu().numpy()
        for k in tuning:
            ctx.set_tuning(k, 0)
    sur_paths = [k for k in res if "surrogate" in k]
    cub_paths = [k for k in res if "cubature" in k]
    best = lambda ks: min(ks, key=lambda k: res[k]["ms_per_step"])
    head = best(sur_paths) if sur_paths else best(cub_paths)
    ms = res[head]["ms_per_step"]
    # graded roofline: the emit pass, the step's dominant kernel, against its
    # compulsory output stream (4 B per FP32 map value per launch), from the
    # one-call path's stage timing
    timed_paths = [k for k in res if "stages" in res[k]]
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
        "config": {"workload": workload + f"; value: the {head} path",
                   "sources": S, "energies": len(inp.logE), "rule": "gl8",
                   "contraction": "bf16x3" if args.precision else "fp64",
                   "per_step": "aeff, livetime gather, exposure, mean PSF, geometry, map"},
        "paths": {k: {kk: (round(vv, 6) if isinstance(vv, float) else vv)
                      for kk, vv in v.items() if kk != "stages"} |
                  ({"stage_ms": {st: round(x["ms"], 4) for st, x in v["stages"].items()}}
                   if "stages" in v else {}) for k, v in res.items()},
    }
    if "footprint" in d:
        result["footprint"] = d["footprint"]
    if timed_paths:
        # the emit's launches measured where they run alone (in the pipelined
        # map they share the CUs with the next batch's row weights)
        hp = "cubature, one pass" if "cubature, one pass" in timed_paths else (
            head if "stages" in res[head] else timed_paths[-1])
        emit = res[hp]["stages"]["emit"]
        achieved = 4.0 * evals / (emit["ms"] / 1e3) / 1e9
        traffic = _profile(f"traffic_{cfg.name}_emit.json") if cfg else None
        result["roofline"] = {
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic["traffic_bytes_per_call"] if traffic else None,
            "kernel": f"emit_sehw_kernel (the model map's store pass) of the {hp} path; "
                      "algorithmic bytes = 4 B per FP32 map value, time = its average launch "
                      "duration from HIP events on the library's stream",
            "launch_ms": emit["ms"],
            "step_frac": 4.0 * evals / (ms / 1e3) / 1e9 / HBM_PEAK_GBS,
            "best_cubature_step_frac": 4.0 * evals / (res[best(cub_paths)]["ms_per_step"] / 1e3)
                                       / 1e9 / HBM_PEAK_GBS}
    if sur:
        n = d["net"]
        result["surrogate"] = {
            "speedup_vs_cubature": res[best(cub_paths)]["ms_per_step"] / ms,
            "target_speedup": 5.0,
            "training_s": n.seconds, "training_samples": n.n_samples,
            "training_steps": n.steps, "final_loss": n.final_loss,
            "network": "4-64-64-64-1 SiLU, BF16 inference on MFMA (layer 1 with the inputs' and weights' high and low BF16 parts, hidden layers plain BF16)"}
    if not args.no_cpu_baseline and cfg is not None:
        ncpu = min(args.cpu_sources, S)
        _log(f"bench: cpu baseline (oracle, {ncpu} sources)")
        cb, accuracy = cpu_baseline(cfg, inp, maps, ncpu)
        result["cpu_baseline"] = cb
        result["accuracy"] = accuracy
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
