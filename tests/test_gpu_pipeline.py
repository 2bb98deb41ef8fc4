"""GPU parity of the pipelined model map (tuning key "pipeline_batches"): the
FP32 [S, E, H, W] map in source batches, one batch's row weights computed by
the same launch as the previous batch's map values.  Checked against the
FP64 oracle and against the passes over all sources at once, through the C
ABI.

Tolerances: 1e-5 relative per map value against the oracle (the bar of the
plain FP32 path); against the map of the passes over all sources, 2e-5 (two
FP32 evaluations within 1e-5 of the oracle each: the tile-frame row-weight
pass forms the pixel blend from FP64 corners once per tile, the default pass
per source from rounded corner offsets)."""
import sys

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

F32_SEHW = None


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    import xtsrcmaps_amd as xa
    global F32_SEHW
    F32_SEHW = xa.LAYOUT_SEHW | xa.OUT_F32 | xa.COMPUTE_F32
    c = xa.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def c1():
    from xtsrcmaps_amd import synthetic
    return synthetic.make_inputs(synthetic.CONFIGS["c1"])


def _rel(got, ref):
    return np.abs(got.astype(np.float64) - ref) / np.abs(ref)


def _maps(ctx, batches, *args, extra=0):
    """The map with the passes over all sources, then in `batches` batches
    (the batches' row weights by the tile-frame pass); extra: more flags."""
    ctx.set_tuning("pipeline_batches", 1)
    a = ctx.point_src_model_map_wcs(*args, flags=F32_SEHW | extra)
    ctx.set_tuning("pipeline_batches", batches)
    b = ctx.point_src_model_map_wcs(*args, flags=F32_SEHW | extra)
    ctx.set_tuning("pipeline_batches", 1)
    return a, b


def test_gpu_tile_frame_row_weights(ctx, orc, c1):
    """The row-weight pass that keeps each tile's pixel frame in LDS across
    runs of sources (tuning key "rowweight_frames"), in the map over all
    sources at once: within 1e-5 of the FP64 oracle on all 100 sources of the
    200x200 workload and 2e-5 of the default pass."""
    from pipelines import oracle_rule, oracle_stages, oracle_wcs
    _log("test_gpu_tile_frame_row_weights: 100 x 37 x 200 x 200, runs of 16 sources")
    o = oracle_stages(orc, c1)
    ref = orc.point_src_model_map_wcs(200, 200, oracle_wcs(orc, c1.wcs), o["src"], o["expo"],
                                      o["upsf"], o["parint"], oracle_rule(c1))
    args = (c1.wcs, o["src"], o["expo"], o["upsf"], o["parint"], c1.rule)
    ctx.set_tuning("pipeline_batches", 1)
    a = ctx.point_src_model_map_wcs(*args, flags=F32_SEHW)
    ctx.set_tuning("rowweight_frames", 16)
    b = ctx.point_src_model_map_wcs(*args, flags=F32_SEHW)
    ctx.set_tuning("rowweight_frames", 0)
    r_or = _rel(b.transpose(0, 2, 3, 1), ref)
    r_gpu = _rel(b, a.astype(np.float64))
    _log(f"  vs oracle: max {r_or.max():.3e} median {np.median(r_or):.3e}; "
         f"vs default pass: max {r_gpu.max():.3e}")
    assert np.isfinite(b).all() and r_or.max() < 1e-5 and r_gpu.max() < 2e-5


def test_gpu_pipeline_c1_vs_oracle(ctx, orc, c1):
    """All 100 sources of the 200x200 workload (37 energies, 8x8
    Gauss-Legendre) in 4 batches of 25: within 1e-5 of the FP64 oracle and
    2e-5 of the unbatched map, finite and positive."""
    from pipelines import oracle_rule, oracle_stages, oracle_wcs
    _log("test_gpu_pipeline_c1_vs_oracle: 100 x 37 x 200 x 200, 4 batches")
    o = oracle_stages(orc, c1)
    ref = orc.point_src_model_map_wcs(200, 200, oracle_wcs(orc, c1.wcs), o["src"], o["expo"],
                                      o["upsf"], o["parint"], oracle_rule(c1))
    a, b = _maps(ctx, 4, c1.wcs, o["src"], o["expo"], o["upsf"], o["parint"], c1.rule)
    assert np.isfinite(b).all() and (b > 0).all()
    r_or = _rel(b.transpose(0, 2, 3, 1), ref)
    r_gpu = _rel(b, a.astype(np.float64))
    _log(f"  vs oracle: max {r_or.max():.3e} median {np.median(r_or):.3e}; "
         f"vs unbatched: max {r_gpu.max():.3e}")
    assert r_or.max() < 1e-5 and r_gpu.max() < 2e-5


@pytest.mark.parametrize("proj", ["CAR", "AIT"])
def test_gpu_pipeline_ragged_shapes(ctx, orc, proj):
    """A 44 x 70 image (neither side a multiple of the 32 x 8 emit tile or the
    16 x 16 row-weight tile), 50 sources in 3 batches (17, 17, 16) inside,
    near the edges and outside the image, 45 energies (two energy chunks of
    the emit), a random PSF table: equal to the unbatched map and within 1e-5
    of the FP64 oracle."""
    import xtsrcmaps_amd as xa
    from pipelines import oracle_wcs
    _log(f"test_gpu_pipeline_ragged_shapes: 44x70 {proj}, 50 sources, 45 energies")
    H, W, S, E = 44, 70, 50, 45
    wcs = xa.Wcs.make((H, W), (20.5, 35.5), (210.0, 35.0), (-0.1, 0.08), 0.0, proj)
    rng = np.random.default_rng(5)
    pix = np.column_stack([rng.uniform(-6, H + 6, S), rng.uniform(-6, W + 6, S)])
    pix[:4] = [[0.2, 0.3], [H - 0.6, W - 0.4], [22.0, 35.0], [22.5, 35.5]]
    src = ctx.pix2sph(wcs, pix)
    upsf = rng.uniform(0.1, 1.0, (S, 401, E)) * np.exp(-np.linspace(0, 9, 401))[None, :, None]
    upsf, parint, _ = ctx.partial_total_integral(upsf)
    expo = rng.uniform(1e8, 1e9, (S, E))
    x8, w8 = np.polynomial.legendre.leggauss(8)
    rule = np.array([[u, v, wu * wv] for u, wu in zip(x8, w8) for v, wv in zip(x8, w8)])
    a, b = _maps(ctx, 3, wcs, src, expo, upsf, parint, rule)
    ref = orc.point_src_model_map_wcs(H, W, oracle_wcs(orc, wcs), src, expo, upsf, parint, rule)
    nz = ref != 0
    for name, m in (("unbatched", a), ("3 batches", b)):
        got = m.transpose(0, 2, 3, 1).astype(np.float64)
        r_or = np.abs(got[nz] - ref[nz]) / np.abs(ref[nz])
        _log(f"  {name} vs oracle: max {r_or.max():.3e}")
        assert np.all(got[~nz] == 0) and r_or.max() < 1e-5
    r_gpu = _rel(b[a != 0], a[a != 0].astype(np.float64))
    _log(f"  batched vs unbatched: max {r_gpu.max():.3e}")
    assert np.all(b[a == 0] == 0) and r_gpu.max() < 2e-5


def test_gpu_pipeline_split_halves(ctx, orc, c1):
    """The split model map (row-weight half, then the rest) in batches equals
    the one-call batched map; a tuning change between the halves cancels the
    pending half."""
    import xtsrcmaps_amd as xa
    from pipelines import oracle_stages
    _log("test_gpu_pipeline_split_halves: 40 sources of c1, 2 batches")
    S = 40
    o = oracle_stages(orc, c1, n_src=S)
    d = {k: torch.as_tensor(np.ascontiguousarray(o[k]), dtype=torch.float64, device="cuda")
         for k in ("src", "expo", "upsf", "parint")}
    E = d["expo"].shape[1]
    torch.cuda.synchronize()
    out1 = torch.empty((S, E, 200, 200), dtype=torch.float32, device="cuda")
    out2 = torch.empty_like(out1)
    ctx.set_tuning("pipeline_batches", 2)
    ctx.point_src_model_map_wcs_dev(c1.wcs, d["src"], d["expo"], d["upsf"], d["parint"], out1,
                                    c1.rule, flags=F32_SEHW)
    ctx.point_src_model_map_rows_dev(c1.wcs, d["src"], E, c1.rule, flags=F32_SEHW)
    ctx.point_src_model_map_finish_dev(c1.wcs, d["src"], d["expo"], d["upsf"], d["parint"], out2,
                                       c1.rule, flags=F32_SEHW)
    ctx.synchronize()
    a, b = out1.double().cpu().numpy(), out2.double().cpu().numpy()
    r = np.max(np.abs(a - b) / np.abs(a))
    _log(f"  split vs one call: max rel {r:.3e}")
    assert np.isfinite(b).all() and r < 2e-6
    ctx.point_src_model_map_rows_dev(c1.wcs, d["src"], E, c1.rule, flags=F32_SEHW)
    ctx.set_tuning("pipeline_batches", 1)
    with pytest.raises(xa.XtsError):
        ctx.point_src_model_map_finish_dev(c1.wcs, d["src"], d["expo"], d["upsf"], d["parint"],
                                           out2, c1.rule, flags=F32_SEHW)


def test_gpu_pipeline_surrogate(ctx, orc, c1):
    """The surrogate map (BF16 and FP8 networks, and the same pixels without
    the network) in 4 batches, the network
    pass of each batch on the context stream, or (tuning key
    "surrogate_beside", default 4) on the second stream as 2 resident blocks
    per CU (out-of-range values of the surrogate's tuning keys refused):
    equal to the one-pass surrogate map with the same network (a short fit on
    the 200x200 image) within 2e-6, and finite and positive."""
    import xtsrcmaps_amd as xa
    from pipelines import oracle_stages
    from xtsrcmaps_amd import surrogate as sg
    _log("test_gpu_pipeline_surrogate: 100 x 37 x 200 x 200, 4 batches")
    o = oracle_stages(orc, c1)
    sg.fit(ctx, c1.wcs, c1.rule, n_samples=1 << 18, steps=300, seed=3)
    for beside in (0, 2):
        ctx.set_tuning("surrogate_beside", beside)
        for extra in (xa.SURROGATE, xa.SURROGATE | xa.SURROGATE_FP8,
                      xa.SURROGATE | xa.SURROGATE_NO_NET):
            a, b = _maps(ctx, 4, c1.wcs, o["src"], o["expo"], o["upsf"], o["parint"], c1.rule,
                         extra=extra)
            r = np.max(_rel(b, a.astype(np.float64)))
            _log(f"  beside {beside}, flags {extra:#x}: batched vs one pass max rel {r:.3e}")
            assert np.isfinite(b).all() and (b > 0).all() and r < 2e-6
    ctx.set_tuning("surrogate_beside", 4)
    for key, bad in (("surrogate_beside", 9), ("surrogate_bf16_lo", 2)):
        with pytest.raises(xa.XtsError):
            ctx.set_tuning(key, bad)
