"""rocprofv3 helpers behind the committed profile summaries.

Workloads to profile (run under ``rocprofv3 ... -- python3 -m
xtsrcmaps_amd.profiling <workload>``):

    modelmap --config c1 --calls 3   the model-map passes on resident inputs
             [--pipeline-batches N]  (1: every pass over all sources at once)
    stages --config c3 --precision 1 the PSF stages (0 FP64, 1 BF16x3 contractions)
    copy --mib 592 --calls 3         an FP32 copy kernel of known size: the
                                     calibration of FETCH_SIZE / WRITE_SIZE
    diffuse --config diffuse --calls 5  the diffuse-source model map (FP32 out)

Summaries of rocprofv3 CSV output (plain Python, no GPU):

    stats <kernel_stats.csv> --cmd CMD           Markdown table of a --stats run
    traffic --mm-fetch D --mm-write D --copy-fetch D --copy-write D
            --copy-mib M --cmd CMD               JSON: HBM bytes per model map
    valu <pmc run dir> --kernel K --cmd CMD      JSON: VALU issue of one kernel
    mfma <pmc run dir> --kernel RX --cmd CMD     JSON: matrix-core use of kernels
"""
from __future__ import annotations

import argparse
import csv
import glob
import json
import os
import re
import sys

MM_KERNELS = (r"rowweight_kernel|rowweight_frames_kernel|sur_rows_kernel|near_rows_kernel|"
              r"hist_integral_kernel|factor_kernel|emit_s|wide_emit_kernel|near_emit_kernel")
COPY_KERNELS = r"MulFunctor"          # torch.mul(src, 1.0, out=dst): reads and writes N bytes


# ---------------------------------------------------------------- workloads
def _modelmap(config: str, calls: int, surrogate: bool = False, fp8: bool = False,
              pipeline_batches: int | None = None) -> None:
    import torch
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    sys.path.insert(0, root)
    import bench
    import xtsrcmaps_amd as xa
    cfg, inp, ctx, d = bench.setup(config, surrogate=surrogate)
    if pipeline_batches is not None:
        ctx.set_tuning("pipeline_batches", pipeline_batches)
    flags = xa.LAYOUT_SEHW | xa.OUT_F32 | xa.COMPUTE_F32 | (xa.SURROGATE if surrogate else 0)
    if surrogate and fp8:
        flags |= xa.SURROGATE_FP8
    bench.step(ctx, inp, d, flags)
    torch.cuda.synchronize()
    for _ in range(calls):
        ctx.point_src_model_map_wcs_dev(inp.wcs, d["src"], d["expo"], d["upsf"], d["parint"],
                                        d["out"], inp.rule, flags=flags)
    torch.cuda.synchronize()
    print(f"modelmap: {calls} calls of {config}" + (" with the surrogate" if surrogate else ""),
          flush=True)


def _stages(config: str, calls: int, precision: int) -> None:
    """The PSF stages (effective area, cosine bins, exposure, mean PSF and
    its integrals) of a workload, `calls` times at the given contraction
    precision (0 FP64, 1 BF16x3 on the matrix cores)."""
    import torch
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    sys.path.insert(0, root)
    import bench
    cfg, inp, ctx, d = bench.setup(config, surrogate=False)
    bench.psf_stages(ctx, inp, d, precision)
    torch.cuda.synchronize()
    for _ in range(calls):
        bench.psf_stages(ctx, inp, d, precision)
    torch.cuda.synchronize()
    print(f"stages: {calls} calls of {config} at precision {precision}", flush=True)


def _diffuse(config: str, calls: int) -> None:
    """The diffuse-source model map of a workload, `calls` times, FP32 output."""
    import torch
    import xtsrcmaps_amd as xa
    from . import synthetic
    cfg = synthetic.DIFFUSE_CONFIGS[config]
    roi, cube_wcs, cube, upsf, expo = synthetic.diffuse_inputs(cfg)
    ctx = xa.Context(0)
    ctx.set_tuning("geometry_cache", 0)
    dev = torch.device("cuda")
    t_cube, t_u, t_e = (torch.as_tensor(a, device=dev) for a in (cube, upsf, expo))
    out = torch.empty((cfg.n_energy, roi.naxis1, roi.naxis2), dtype=torch.float32, device=dev)
    torch.cuda.synchronize()
    for _ in range(calls + 1):
        ctx.diffuse_src_model_map_wcs_dev(roi, cube_wcs, t_cube, t_u, t_e, cfg.target, out,
                                          flags=xa.OUT_F32)
    ctx.synchronize()
    print(f"diffuse: {calls + 1} calls of {config}", flush=True)


def _copy(mib: int, calls: int) -> None:
    import torch
    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(calls):
        torch.mul(src, 1.0, out=dst)
    torch.cuda.synchronize()
    print(f"copy: {calls} x {mib} MiB", flush=True)


# ---------------------------------------------------------------- summaries
def _one(pattern: str) -> str:
    found = sorted(glob.glob(pattern, recursive=True))
    if not found:
        raise FileNotFoundError(pattern)
    return found[0]


def stats_markdown(path: str, cmd: str, top: int = 40) -> str:
    rows = list(csv.DictReader(open(path)))
    lines = [f"Command: `{cmd}`", "",
             "| kernel | calls | avg µs | total µs | % |", "|---|---|---|---|---|"]
    for r in rows[:top]:
        name = r["Name"].replace("|", "\\|")
        if len(name) > 90:
            name = name[:87] + "..."
        lines.append(f"| {name} | {r['Calls']} | {float(r['AverageNs']) / 1e3:.1f} | "
                     f"{float(r['TotalDurationNs']) / 1e3:.1f} | {float(r['Percentage']):.1f} |")
    return "\n".join(lines) + "\n"


def _counter_sum(run_dir: str, counter: str, kernel_rx: str) -> tuple[float, int]:
    """Sum of `counter` (KiB per dispatch) over dispatches whose kernel name
    matches kernel_rx; returns (total KiB, dispatch count)."""
    path = _one(os.path.join(run_dir, "**", "*counter_collection.csv"))
    per_dispatch = {}
    rx = re.compile(kernel_rx)
    for r in csv.DictReader(open(path)):
        if r["Counter_Name"] != counter or not rx.search(r["Kernel_Name"]):
            continue
        key = r["Dispatch_Id"]
        per_dispatch[key] = per_dispatch.get(key, 0.0) + float(r["Counter_Value"])
    return sum(per_dispatch.values()), len(per_dispatch)


def traffic_json(mm_fetch: str, mm_write: str, copy_fetch: str, copy_write: str,
                 copy_mib: int, cmd: str, maps: int = 0) -> dict:
    """HBM bytes per model map; maps: model maps in the profiled run (0: one
    per emit dispatch, as without source batches)."""
    copy_bytes = copy_mib * (1 << 20)
    cf, nf = _counter_sum(copy_fetch, "FETCH_SIZE", COPY_KERNELS)
    cw, nw = _counter_sum(copy_write, "WRITE_SIZE", COPY_KERNELS)
    # the copy kernel reads and writes copy_bytes per call
    f_fetch = cf * 1024 / (nf * copy_bytes)
    f_write = cw * 1024 / (nw * copy_bytes)
    mf, mnf = _counter_sum(mm_fetch, "FETCH_SIZE", MM_KERNELS)
    mw, mnw = _counter_sum(mm_write, "WRITE_SIZE", MM_KERNELS)
    # one emit dispatch per model map unless told (the pipelined map has one
    # per source batch)
    _, calls = _counter_sum(mm_write, "WRITE_SIZE", r"emit_s")
    if maps > 0:
        calls = maps
    fetch_b = mf * 1024 / calls / f_fetch
    write_b = mw * 1024 / calls / f_write
    per_kernel = {}
    for name, rx in (("rowweight", "rowweight_kernel|rowweight_frames_kernel"),
                     ("surrogate", "sur_rows_kernel"), ("wide", "wide_emit_kernel"),
                     ("near_rows", "near_rows_kernel"),
                     ("hist_factor", "hist_integral_kernel|factor_kernel"),
                     ("emit", "emit_s"), ("near_emit", "near_emit_kernel")):
        f, _ = _counter_sum(mm_fetch, "FETCH_SIZE", rx)
        w, _ = _counter_sum(mm_write, "WRITE_SIZE", rx)
        per_kernel[name] = {"fetch_bytes": f * 1024 / calls / f_fetch,
                            "write_bytes": w * 1024 / calls / f_write}
    return {"command": cmd, "copy_mib": copy_mib,
            "fetch_factor": f_fetch, "write_factor": f_write, "model_map_calls": calls,
            "dispatches_fetch": mnf, "dispatches_write": mnw,
            "fetch_bytes_per_call": fetch_b, "write_bytes_per_call": write_b,
            "traffic_bytes_per_call": fetch_b + write_b, "per_kernel": per_kernel}


# VALU counters of one kernel (one rocprofv3 --pmc pass with SQ_WAVES,
# SQ_INSTS_VALU, SQ_INSTS_LDS, SQ_WAIT_INST_ANY, SQ_WAVE_CYCLES and
# GRBM_GUI_ACTIVE).  The VALU rate is counted in lane operations: a wave64
# VALU instruction is 64 lane operations, and MI355X's FP32 vector peak of
# 157.3 TFLOP/s is 78.65e12 FMA lane operations per second.
VALU_LANE_OPS_PEAK = 157.3e12 / 2


# Matrix-core counters of the kernels matching kernel_rx (one --pmc pass
# with SQ_INSTS_VALU_MFMA_MOPS_BF16, SQ_INSTS_MFMA, SQ_VALU_MFMA_BUSY_CYCLES,
# SQ_BUSY_CYCLES and GRBM_GUI_ACTIVE).  One MOPS count is 512 floating-point
# operations of the matrix cores.
MFMA_PEAK_BF16 = 2.5e15


def mfma_json(run_dir: str, kernel_rx: str, cmd: str) -> dict:
    path = _one(os.path.join(run_dir, "**", "*counter_collection.csv"))
    rx = re.compile(kernel_rx)
    disp = {}
    for r in csv.DictReader(open(path)):
        if not rx.search(r["Kernel_Name"]):
            continue
        d = disp.setdefault(r["Dispatch_Id"], {"ns": int(r["End_Timestamp"]) - int(r["Start_Timestamp"]),
                                                "name": r["Kernel_Name"].split("(")[0]})
        d[r["Counter_Name"]] = d.get(r["Counter_Name"], 0.0) + float(r["Counter_Value"])
    if not disp:
        raise ValueError(f"no dispatch of {kernel_rx} in {path}")
    kernels = {}
    for d in disp.values():
        kernels.setdefault(d["name"], []).append(d)
    out = {"command": cmd, "kernels": {}}
    for name, rows in kernels.items():
        mean = lambda k: sum(x.get(k, 0.0) for x in rows) / len(rows)
        ns = mean("ns")
        flops = mean("SQ_INSTS_VALU_MFMA_MOPS_BF16") * 512
        out["kernels"][name] = {
            "dispatches": len(rows), "duration_us": ns / 1e3,
            "mfma_insts": mean("SQ_INSTS_MFMA"),
            "mfma_bf16_flops": flops,
            "mfma_bf16_tflops": flops / (ns * 1e-9) / 1e12 if ns else 0.0,
            "frac_of_bf16_dense_peak": flops / (ns * 1e-9) / MFMA_PEAK_BF16 if ns else 0.0,
            "mfma_busy_over_busy_cycles": (mean("SQ_VALU_MFMA_BUSY_CYCLES") /
                                           max(mean("SQ_BUSY_CYCLES"), 1.0))}
    return out


def valu_json(run_dir: str, kernel_rx: str, cmd: str) -> dict:
    path = _one(os.path.join(run_dir, "**", "*counter_collection.csv"))
    rx = re.compile(kernel_rx)
    disp = {}
    for r in csv.DictReader(open(path)):
        if not rx.search(r["Kernel_Name"]):
            continue
        d = disp.setdefault(r["Dispatch_Id"], {"ns": int(r["End_Timestamp"]) - int(r["Start_Timestamp"]),
                                                "name": r["Kernel_Name"]})
        d[r["Counter_Name"]] = d.get(r["Counter_Name"], 0.0) + float(r["Counter_Value"])
    if not disp:
        raise ValueError(f"no dispatch of {kernel_rx} in {path}")
    rows = list(disp.values())
    mean = lambda k: sum(d[k] for d in rows) / len(rows)
    ns = mean("ns")
    valu = mean("SQ_INSTS_VALU")
    lane_ops = valu * 64 / (ns * 1e-9)
    return {"command": cmd, "kernel": rows[0]["name"], "dispatches": len(rows),
            "duration_us": ns / 1e3, "waves": mean("SQ_WAVES"),
            "valu_insts": valu, "lds_insts": mean("SQ_INSTS_LDS"),
            "valu_insts_per_wave": valu / mean("SQ_WAVES"),
            "wait_any_per_wave_cycles": mean("SQ_WAIT_INST_ANY") / mean("SQ_WAVE_CYCLES"),
            "valu_lane_ops_per_s": lane_ops,
            "valu_frac_of_fp32_peak": lane_ops / VALU_LANE_OPS_PEAK}


def main(argv=None) -> None:
    ap = argparse.ArgumentParser(prog="python -m xtsrcmaps_amd.profiling")
    sub = ap.add_subparsers(dest="what", required=True)
    p = sub.add_parser("modelmap")
    p.add_argument("--config", default="c1")
    p.add_argument("--calls", type=int, default=3)
    p.add_argument("--surrogate", action="store_true")
    p.add_argument("--fp8", action="store_true", help="with --surrogate: the FP8 network")
    p.add_argument("--pipeline-batches", type=int, default=None,
                   help="tuning key pipeline_batches (1: every pass over all sources at once)")
    p = sub.add_parser("stages")
    p.add_argument("--config", default="c3")
    p.add_argument("--calls", type=int, default=5)
    p.add_argument("--precision", type=int, default=0)
    p = sub.add_parser("diffuse")
    p.add_argument("--config", default="diffuse")
    p.add_argument("--calls", type=int, default=5)
    p = sub.add_parser("copy")
    p.add_argument("--mib", type=int, default=592)
    p.add_argument("--calls", type=int, default=3)
    p = sub.add_parser("stats")
    p.add_argument("csv")
    p.add_argument("--cmd", required=True)
    p = sub.add_parser("traffic")
    for k in ("mm-fetch", "mm-write", "copy-fetch", "copy-write"):
        p.add_argument("--" + k, required=True)
    p.add_argument("--copy-mib", type=int, required=True)
    p.add_argument("--cmd", required=True)
    p.add_argument("--maps", type=int, default=0,
                   help="model maps in the profiled run (0: one per emit dispatch)")
    p.add_argument("--kernel-only", default="",
                   help="report this kernel alone (e.g. emit) as traffic_bytes_per_call")
    p = sub.add_parser("mfma")
    p.add_argument("run_dir")
    p.add_argument("--kernel", default="gemm|exposure|mean_psf|contract")
    p.add_argument("--cmd", required=True)
    p = sub.add_parser("valu")
    p.add_argument("run_dir")
    p.add_argument("--kernel", default="rowweight_kernel<float>")
    p.add_argument("--cmd", required=True)
    a = ap.parse_args(argv)
    if a.what == "modelmap":
        _modelmap(a.config, a.calls, a.surrogate, a.fp8, a.pipeline_batches)
    elif a.what == "stages":
        _stages(a.config, a.calls, a.precision)
    elif a.what == "copy":
        _copy(a.mib, a.calls)
    elif a.what == "diffuse":
        _diffuse(a.config, a.calls)
    elif a.what == "mfma":
        print(json.dumps(mfma_json(a.run_dir, a.kernel, a.cmd), indent=1))
    elif a.what == "stats":
        sys.stdout.write(stats_markdown(a.csv, a.cmd))
    elif a.what == "valu":
        print(json.dumps(valu_json(a.run_dir, re.escape(a.kernel), a.cmd), indent=1))
    else:
        t = traffic_json(a.mm_fetch, a.mm_write, a.copy_fetch, a.copy_write, a.copy_mib, a.cmd,
                         a.maps)
        if a.kernel_only:
            k = t["per_kernel"][a.kernel_only]
            t = {"command": a.cmd, "kernel": a.kernel_only, "fetch_factor": t["fetch_factor"],
                 "write_factor": t["write_factor"], "model_map_calls": t["model_map_calls"],
                 "fetch_bytes_per_call": k["fetch_bytes"], "write_bytes_per_call": k["write_bytes"],
                 "traffic_bytes_per_call": k["fetch_bytes"] + k["write_bytes"],
                 "model_map": {kk: t[kk] for kk in ("fetch_bytes_per_call", "write_bytes_per_call",
                                                    "traffic_bytes_per_call", "per_kernel")}}
        print(json.dumps(t, indent=1))


if __name__ == "__main__":
    main()
