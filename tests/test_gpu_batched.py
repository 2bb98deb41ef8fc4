"""GPU tests of the batched engine (xts_point_src_model_maps_batched_dev) on
the four-image workload (2,000 sources over 64^2, 128^2, 256^2 and 400^2
CAR images at 0.05 deg, 37 energies, 8x8 Gauss-Legendre):
- footprints at 99.5% containment, and the values inside them for two
  sources of every image against the FP64 oracle of that image (1e-5);
  every value of the full output finite and positive; no mass clipped away
  when the boxes are whole images, at most the containment's complement
  otherwise;
- smaller footprints (the box-copy path) equal to the same values;
- the autotuned shape and the HIP-graph replays (three, with other device
  work between them) equal to the eager result.
Child process: torch (device buffers) before the library."""
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

_CHILD = r"""
import sys, time
import numpy as np
import torch
root = sys.argv[1]
sys.path[:0] = [root, root + "/tests"]
import xtsrcmaps_amd as xa
from xtsrcmaps_amd import synthetic
ok = True
def check(name, cond, msg):
    global ok
    print(f"{name}: {msg} {'ok' if cond else 'FAILED'}", flush=True)
    ok = ok and bool(cond)

ctx = xa.Context(0)
dev = torch.device("cuda:0")
stream = torch.cuda.Stream()
torch.cuda.set_stream(stream)
ctx.set_stream(stream.cuda_stream)
B = synthetic.make_batch_inputs(synthetic.BATCH_CONFIGS["c2"])
inp = B.phys
S, E = inp.src_sph.shape[0], inp.cfg.n_energy
t = lambda a: torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device=dev)
king = {}
for side, (params, cos, loge) in inp.king_raw.items():
    king[side] = (t(ctx.normalize_rpsf(params, loge, inp.psf_scale)), t(cos), t(loge))
af = t(ctx.aeff_value(inp.costh, inp.logE, *inp.aeff_grid["front"]))
ab = t(ctx.aeff_value(inp.costh, inp.logE, *inp.aeff_grid["back"]))
ltf = t(xa.livetime_efficiency_factors(inp.logE, *inp.effic))
src, cube, wcube = t(inp.src_sph), t(inp.cube), t(inp.wcube)
C = len(inp.costh)
ecb = torch.empty((S, C), dtype=torch.float64, device=dev); wecb = torch.empty_like(ecb)
expo = torch.empty((S, E), dtype=torch.float64, device=dev)
upsf = torch.empty((S, 401, E), dtype=torch.float64, device=dev); parint = torch.empty_like(upsf)
torch.cuda.synchronize()
ctx.src_cosbins_dev(src, cube, inp.nside, ecb)
ctx.src_cosbins_dev(src, wcube, inp.nside, wecb)
ctx.exposure_dev(ecb, wecb, af, ab, ltf, expo)
ctx.psf_lookup_dev(king["front"], king["back"], t(inp.costh), t(inp.logE), af, ab, ecb, wecb,
                   ltf, expo, upsf, parint)
boxes, radius, total = ctx.model_map_footprints(B.wcs, B.src_begin, src, parint, 0.995)
full = all(boxes[s].h0 == 0 and boxes[s].w0 == 0 and
           boxes[s].h1 == B.wcs[i].naxis1 and boxes[s].w1 == B.wcs[i].naxis2
           for i in range(4) for s in range(B.src_begin[i], B.src_begin[i + 1]))
print(f"99.5% footprint radii {radius.min():.1f} .. {radius.max():.1f} deg; boxes whole images: "
      f"{full}; output {total:.4g} values", flush=True)
out = torch.empty(total, dtype=torch.float32, device=dev)
ctx.point_src_model_maps_batched_dev(B.wcs, B.src_begin, src, expo, upsf, parint, boxes, out,
                                     inp.rule)
torch.cuda.synchronize()
ctx.synchronize()
check("finite", bool(torch.isfinite(out).all()) and bool((out > 0).all()),
      f"{total} values finite and positive")
ref_out = out.clone()

def block(o, s, wcs):
    b = boxes_now[s]
    n = E * (b.h1 - b.h0) * (b.w1 - b.w0)
    return o[b.offset:b.offset + n].view(E, b.h1 - b.h0, b.w1 - b.w0)

# the oracle for two sources of every image
from oracle import oracle as orc
from pipelines import oracle_stages, oracle_rule, oracle_wcs
boxes_now = boxes
worst = 0.0
for i in range(4):
    picks = [int(B.src_begin[i]), int(B.src_begin[i + 1]) - 1]
    sub = type(inp)(**{**inp.__dict__, "src_sph": inp.src_sph[picks]})
    o = oracle_stages(orc, sub)
    w = B.wcs[i]
    ref = orc.point_src_model_map_wcs(w.naxis1, w.naxis2, oracle_wcs(orc, w), o["src"], o["expo"],
                                      o["upsf"], o["parint"], oracle_rule(inp))
    for k, s in enumerate(picks):
        got = block(out, s, w).double().cpu().numpy().transpose(1, 2, 0)
        b = boxes[s]
        r = ref[k, b.h0:b.h1, b.w0:b.w1, :]
        rel = float((np.abs(got - r) / np.abs(r)).max())
        worst = max(worst, rel)
check("oracle", worst < 1e-5, f"8 sources (2 per image) inside their footprints: max rel {worst:.2e}")

# smaller footprints: the box-copy path gives the same values
boxes_s, radius_s, total_s = ctx.model_map_footprints(B.wcs, B.src_begin, src, parint, 0.5)
part = sum(1 for i in range(4) for s in range(B.src_begin[i], B.src_begin[i + 1])
           if (boxes_s[s].h1 - boxes_s[s].h0) * (boxes_s[s].w1 - boxes_s[s].w0)
           < B.wcs[i].naxis1 * B.wcs[i].naxis2)
out_s = torch.empty(total_s, dtype=torch.float32, device=dev)
ctx.point_src_model_maps_batched_dev(B.wcs, B.src_begin, src, expo, upsf, parint, boxes_s, out_s,
                                     inp.rule)
torch.cuda.synchronize()
ctx.synchronize()
worst, clipped = 0.0, 0.0
img_of = np.searchsorted(B.src_begin, np.arange(S), side="right") - 1
for s in list(range(0, S, 97)) + [S - 1]:
    i = int(img_of[s]); w = B.wcs[i]
    boxes_now = boxes
    full_s = block(ref_out, s, w)
    bs = boxes_s[s]
    boxes_now = boxes_s
    got = block(out_s, s, w)
    sl = full_s[:, bs.h0:bs.h1, bs.w0:bs.w1]
    worst = max(worst, float(((got.double() - sl.double()).abs() / sl.double().abs()).max()))
    tot = full_s.double().sum(dim=(1, 2))
    clipped = max(clipped, float((1.0 - sl.double().sum(dim=(1, 2)) / tot).max()))
check("boxes", worst <= 2e-7, f"50% footprints ({part} of {S} sources clipped): values equal to "
      f"the whole-image ones to {worst:.1e}; largest in-image mass outside a box {clipped:.3f}")
check("clipped mass", clipped <= 0.5 + 0.02, "at most the containment's complement (+2%)")

# the autotuned shape, then graph replays with other device work between
ctx.set_tuning("autotune", 1)
t0 = time.perf_counter()
out2 = torch.empty_like(out)
ctx.point_src_model_maps_batched_dev(B.wcs, B.src_begin, src, expo, upsf, parint, boxes, out2,
                                     inp.rule)
ctx.synchronize()
print(f"autotune: {time.perf_counter() - t0:.2f} s, batch_sources {ctx.get_tuning('batch_sources')} "
      f"rowweight_blocks {ctx.get_tuning('rowweight_blocks')} batches {ctx.get_tuning('batches')}",
      flush=True)
def rel(a, b, chunk=1 << 28):
    worst = 0.0
    for k in range(0, a.numel(), chunk):
        x, y = a[k:k + chunk], b[k:k + chunk]
        worst = max(worst, float(((x - y).abs() / y.abs()).max()))
    return worst
check("autotune", rel(out2, ref_out) <= 2e-7, f"tuned result vs default {rel(out2, ref_out):.1e}")
ctx.set_tuning("graph", 1)
worst = 0.0
for k in range(4):
    out2.fill_(-1.0)
    torch.cuda.synchronize()
    ctx.point_src_model_maps_batched_dev(B.wcs, B.src_begin, src, expo, upsf, parint, boxes, out2,
                                         inp.rule)
    ctx.synchronize()
    worst = max(worst, rel(out2, ref_out))
    # other device work between replays: an eager model map of one image
    ctx.point_src_model_map_wcs(B.wcs[0], inp.src_sph[:3], expo[:3].cpu().numpy(),
                                upsf[:3].cpu().numpy(), parint[:3].cpu().numpy(), inp.rule)
caps, reps = ctx.get_tuning("graph_captures"), ctx.get_tuning("graph_replays")
check("graph", worst <= 2e-7 and reps >= 2,
      f"{caps} captures, {reps} replays (the first eager call between grows buffers and "
      f"forces one recapture) vs eager {worst:.1e}")
sys.exit(0 if ok else 1)
"""


def test_gpu_batched_engine_c2():
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    print("test_gpu_batched_engine_c2: four images, 2,000 sources (child process)", flush=True)
    r = subprocess.run([sys.executable, "-c", _CHILD, root], capture_output=True, text=True,
                       timeout=280)
    print(r.stdout[-3000:], flush=True)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
