"""GPU tests of the flux-weighted summed model (xts_point_src_summed_model_dev)
on the large-sky configuration (10,000 sources, 1000x1000 at 0.02 deg, 100
energies):
- the large sky's physics on a 160x160 window with 12 sources: the summed
  model against the FP64 oracle's per-source maps x flux, summed in FP64,
  per energy plane (max and median relative error; the plain FP32 path
  within 1e-5), and with the FP8 surrogate (trained on the device for the
  0.02-degree pixels) within 1e-4;
- two sources on the full 1000x1000 image against the oracle, per plane;
- misuse (flags, shapes, dtypes, a missing network, tuning values)
  refused before any launch; batches of 5 sources equal to one batch;
- the full 10,000 sources: every value finite and positive; the sums over
  two halves of the source list add up to the whole (FP32 accumulation
  order only, 1e-5); doubling every flux doubles the model (1e-6: the near
  pixels are added by atomics in no fixed order).
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
import bench
import xtsrcmaps_amd as xa
from xtsrcmaps_amd import surrogate as sg, synthetic
from oracle import oracle as orc
from pipelines import oracle_stages, oracle_rule, oracle_wcs
ok = True
def check(name, cond, msg):
    global ok
    print(f"{name}: {msg} {'ok' if cond else 'FAILED'}", flush=True)
    ok = ok and bool(cond)

def per_plane(got, ref):
    # got [E, H, W] float32, ref [E, H, W] float64
    rel = np.abs(got.astype(np.float64) - ref) / np.abs(ref)
    rel = rel.reshape(rel.shape[0], -1)
    return rel.max(1), np.median(rel, 1)

def oracle_sum(inp, pick, flux):
    sub = type(inp)(**{**inp.__dict__, "src_sph": inp.src_sph[pick]})
    o = oracle_stages(orc, sub)
    H, W = inp.wcs.naxis1, inp.wcs.naxis2
    maps = orc.point_src_model_map_wcs(H, W, oracle_wcs(orc, inp.wcs), o["src"], o["expo"],
                                       o["upsf"], o["parint"], oracle_rule(inp))  # [S, H, W, E]
    return np.einsum("shwe,s->ehw", maps, flux[pick])

def gpu_sum(ctx, inp, d, flux_t, pick=None, flags=0):
    E, H, W = len(inp.logE), inp.wcs.naxis1, inp.wcs.naxis2
    out = torch.empty((E, H, W), dtype=torch.float32, device="cuda")
    sel = (lambda t: t) if pick is None else (lambda t: t[pick].contiguous())
    ctx.point_src_summed_model_dev(inp.wcs, sel(d["src"]), sel(flux_t), sel(d["expo"]),
                                   sel(d["upsf"]), sel(d["parint"]), out, inp.rule, flags=flags)
    torch.cuda.synchronize()
    return out

# -- the window: every source against the oracle
cfg = synthetic.CONFIGS["c4w"]
inp = synthetic.make_inputs(cfg)
flux = synthetic.fluxes(cfg)
ctx, d = bench.setup_inputs(inp, with_out=False)
bench.psf_stages(ctx, inp, d)
torch.cuda.synchronize()
ft = torch.as_tensor(flux, dtype=torch.float64, device="cuda")

# -- misuse refused before any launch
def refused(fn):
    try:
        fn()
    except xa.XtsError:
        return True
    return False
E, H, W = len(inp.logE), inp.wcs.naxis1, inp.wcs.naxis2
out = torch.empty((E, H, W), dtype=torch.float32, device="cuda")
call = lambda o=out, f=ft, fl=0: ctx.point_src_summed_model_dev(
    inp.wcs, d["src"], f, d["expo"], d["upsf"], d["parint"], o, inp.rule, flags=fl)
check("refusals", all([
    refused(lambda: call(fl=xa.COMPUTE_F64)),                          # FP64 cubature
    refused(lambda: call(fl=xa.LAYOUT_SEHW)),                          # a per-source layout
    refused(lambda: call(fl=xa.SURROGATE)),                            # no network set yet
    refused(lambda: call(fl=xa.SURROGATE_FP8)),                        # FP8 without SURROGATE
    refused(lambda: call(o=out.double())),                             # output dtype
    refused(lambda: call(o=out[:, :, :W - 1])),                        # output shape
    refused(lambda: call(f=ft[:-1].contiguous())),                     # flux length
    refused(lambda: ctx.set_tuning("sum_group", 3)),
    refused(lambda: ctx.set_tuning("sum_batch_sources", -1)),
]), "bad flags, shapes, dtypes, a missing network and bad tuning values raise")
# batches of any size give the same model (up to the atomics' order)
ctx.set_tuning("sum_batch_sources", 5)
small = torch.empty_like(out)
call(o=small)
ctx.set_tuning("sum_batch_sources", 0)
call()
torch.cuda.synchronize()
e = float(((small - out).abs() / out).max())
check("batches", e <= 1e-6, f"batches of 5 sources: max rel {e:.1e} from one batch")
t0 = time.perf_counter()
ref = oracle_sum(inp, np.arange(cfg.n_src), flux)
print(f"window: oracle {time.perf_counter() - t0:.1f} s", flush=True)
mx, md = per_plane(gpu_sum(ctx, inp, d, ft).cpu().numpy(), ref)
print("window, plain: per-plane max " + " ".join(f"{x:.1e}" for x in mx[::11]) +
      f"; worst {mx.max():.2e}, median of medians {np.median(md):.2e}", flush=True)
check("window plain", mx.max() <= 1e-5, "summed model within 1e-5 of the oracle in every plane")
net = sg.fit(ctx, inp.wcs, inp.rule, seed=4)
print(f"surrogate for 0.02-degree pixels: final loss {net.final_loss:.3e}", flush=True)
for name, fl in (("bf16", xa.SURROGATE), ("fp8", xa.SURROGATE | xa.SURROGATE_FP8)):
    mx, md = per_plane(gpu_sum(ctx, inp, d, ft, flags=fl).cpu().numpy(), ref)
    print(f"window, {name} surrogate: per-plane max " + " ".join(f"{x:.1e}" for x in mx[::11]) +
          f"; worst {mx.max():.2e}, median of medians {np.median(md):.2e}", flush=True)
    check(f"window {name}", mx.max() <= 1e-4, f"{name} surrogate within 1e-4 in every plane")
del d, ctx
torch.cuda.empty_cache()

# -- the full image
cfg = synthetic.CONFIGS["c4"]
inp = synthetic.make_inputs(cfg)
flux = synthetic.fluxes(cfg)
ctx, d = bench.setup_inputs(inp, with_out=False)
bench.psf_stages(ctx, inp, d)
torch.cuda.synchronize()
ft = torch.as_tensor(flux, dtype=torch.float64, device="cuda")
pick = np.array([17, 5003])
t0 = time.perf_counter()
ref = oracle_sum(inp, pick, flux)
print(f"full image, 2 sources: oracle {time.perf_counter() - t0:.1f} s", flush=True)
mx, md = per_plane(gpu_sum(ctx, inp, d, ft, torch.as_tensor(pick, device="cuda")).cpu().numpy(),
                   ref)
print("full image, 2 sources: per-plane max " + " ".join(f"{x:.1e}" for x in mx[::11]) +
      f"; worst {mx.max():.2e}, median of medians {np.median(md):.2e}", flush=True)
check("subset", mx.max() <= 1e-5, "two sources on the full image within 1e-5 in every plane")
del ref
t0 = time.perf_counter()
whole = gpu_sum(ctx, inp, d, ft)
print(f"10,000 sources: {time.perf_counter() - t0:.2f} s", flush=True)
check("finite", bool(torch.isfinite(whole).all()) and float(whole.min()) > 0,
      "every value finite and positive")
S = cfg.n_src
a = gpu_sum(ctx, inp, d, ft, torch.arange(0, S // 2, device="cuda"))
b = gpu_sum(ctx, inp, d, ft, torch.arange(S // 2, S, device="cuda"))
e = float(((a + b - whole).abs() / whole).max())
check("additive", e <= 1e-5, f"two halves add up to the whole: max rel {e:.1e}")
del a, b
# every factor doubles exactly; the near pixels' atomic additions run in
# whatever order the hardware takes them, so equal to FP32 rounding
twice = gpu_sum(ctx, inp, d, 2.0 * ft)
e = float(((twice - 2.0 * whole).abs() / whole).max())
check("linear", e <= 1e-6, f"doubled fluxes give twice the model: max rel {e:.1e}")
sys.exit(0 if ok else 1)
"""


def test_gpu_summed_model_large_sky():
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    print("test_gpu_summed: window and full-image oracle checks, full-size properties (child)",
          flush=True)
    r = subprocess.run([sys.executable, "-c", _CHILD, root], capture_output=True, text=True,
                       timeout=280)
    print(r.stdout)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
