"""The no-allocation model map (xts_point_src_model_map_wcs_dev_noalloc)
through the C ABI against the existing device entry point
(xts_point_src_model_map_wcs_dev) on the same inputs: two calls in a row,
the second with fewer sources on a smaller image, each within 2e-7 relative
(the map integrals are summed with atomics, so the last bit may differ);
the device's free memory unchanged across both calls; a call that needs more
than the scratch refused before any launch.  Child process: torch (device
memory for the inputs and outputs) before the library."""
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

_CHILD = r"""
import ctypes as C
import sys
import numpy as np
import torch
sys.path.insert(0, sys.argv[1])
import bench
import xtsrcmaps_amd as xa
from xtsrcmaps_amd import capi
cfg, inp, ctx, d = bench.setup("c1", surrogate=False)
d["serial"] = True
bench.step(ctx, inp, d)                 # exposure and PSF on the device
torch.cuda.synchronize()
L = capi.lib()
flags = xa.LAYOUT_SEHW | xa.OUT_F32 | xa.COMPUTE_F32
rule = inp.rule
rp, N = capi.rule_ptr(rule)

def run(entry, wcs, S, out):
    p = capi.devs([("src", d["src"][:S].contiguous(), "float64", (S, 2)),
                   ("expo", d["expo"][:S].contiguous(), "float64", (S, cfg.n_energy)),
                   ("upsf", d["upsf"][:S].contiguous(), "float64", (S, 401, cfg.n_energy)),
                   ("parint", d["parint"][:S].contiguous(), "float64", (S, 401, cfg.n_energy)),
                   ("out", out, "float32", None, S * cfg.n_energy * wcs.naxis1 * wcs.naxis2)])
    keep.append(p)
    return entry(ctx.handle, C.byref(wcs), p[0], S, p[1], p[2], p[3], cfg.n_energy, rp, N,
                 flags, p[4])

keep = []
small = xa.Wcs.make((120, 90), (70.5, 40.5), (180.0, 0.0), (-0.05, 0.05), 0.0, "CAR")
cases = [(inp.wcs, cfg.n_src), (small, 30)]
refs, outs = [], []
for wcs, S in cases:
    ref = torch.empty((S, cfg.n_energy, wcs.naxis1, wcs.naxis2), dtype=torch.float32,
                      device="cuda")
    capi.check(run(L.xts_point_src_model_map_wcs_dev, wcs, S, ref))
    refs.append(ref)
    outs.append(torch.full_like(ref, -1.0))
torch.cuda.synchronize()
free0 = torch.cuda.mem_get_info()[0]
for (wcs, S), out in zip(cases, outs):
    print(f"no-alloc call: {S} sources, {wcs.naxis1}x{wcs.naxis2}", flush=True)
    capi.check(run(L.xts_point_src_model_map_wcs_dev_noalloc, wcs, S, out))
    torch.cuda.synchronize()
# nothing else touches device memory between the two readings (the
# comparisons below allocate torch temporaries)
free1 = torch.cuda.mem_get_info()[0]
print(f"device free memory before {free0} after {free1}", flush=True)
ok = free0 == free1
for (wcs, S), ref, out in zip(cases, refs, outs):
    rel = float(((out.double() - ref.double()).abs() / ref.double().abs()).max())
    print(f"  {S} sources: max rel vs xts_point_src_model_map_wcs_dev {rel:.2e}", flush=True)
    ok = ok and rel <= 2e-7 and bool(torch.isfinite(out).all())
# more than the scratch holds: 2,000 sources on 400x400 -> refused, nothing launched
big = xa.Wcs.make((400, 400), (200.5, 200.5), (180.0, 0.0), (-0.05, 0.05), 0.0, "CAR")
dummy = torch.empty(1, device="cuda")
rc = L.xts_point_src_model_map_wcs_dev_noalloc(
    ctx.handle, C.byref(big), capi._vp(d["src"].data_ptr()), 2000, capi._vp(d["expo"].data_ptr()),
    capi._vp(d["upsf"].data_ptr()), capi._vp(d["parint"].data_ptr()), cfg.n_energy, rp, N, flags,
    capi._vp(dummy.data_ptr()))
msg = L.xts_last_error().decode()
print(f"oversize call rc {rc}: {msg}", flush=True)
ok = ok and rc == 1 and "scratch" in msg
# the existing entry point still works afterwards and agrees
again = torch.empty_like(refs[0])
capi.check(run(L.xts_point_src_model_map_wcs_dev, cases[0][0], cases[0][1], again))
torch.cuda.synchronize()
ok = ok and float(((again.double() - refs[0].double()).abs() / refs[0].double().abs()).max()) <= 2e-7
sys.exit(0 if ok else 1)
"""


def test_gpu_no_alloc_model_map_matches_device_entry_point():
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    print("test_gpu_no_alloc_model_map: c1 then 30 sources on 120x90 (child process)",
          flush=True)
    r = subprocess.run([sys.executable, "-c", _CHILD, root], capture_output=True, text=True,
                       timeout=240)
    print(r.stdout[-2000:], flush=True)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-3000:]
