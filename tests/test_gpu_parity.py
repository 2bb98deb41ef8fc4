"""GPU parity: every stage of the HIP path against the FP64 oracle on the
same inputs, and against the reference's own fixtures.  Calls go through the
C ABI (xtsrcmaps_amd.capi).

Tolerances: FP64 stages 1e-10 relative (different but exact-in-FP64
summation orders); FP32 cubature 1e-5 relative per map value (the bar for the
plain FP32 path); BF16x3 contractions 5e-6 relative (FP32-level)."""
import os
import sys

import numpy as np
import pytest

from conftest import max_rel

pytestmark = pytest.mark.gpu


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    import xtsrcmaps_amd as xa
    c = xa.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def c0():
    from xtsrcmaps_amd import synthetic
    return synthetic.make_inputs(synthetic.CONFIGS["c0"])


@pytest.fixture(scope="module")
def c1():
    from xtsrcmaps_amd import synthetic
    return synthetic.make_inputs(synthetic.CONFIGS["c1"])


@pytest.fixture(scope="module")
def c0_oracle(orc, c0):
    from pipelines import oracle_stages
    return oracle_stages(orc, c0)


def _gpu_stages(ctx, inp, src, precision=0):
    import xtsrcmaps_amd as xa
    out = {}
    king = {}
    for side, (params, cos, loge) in inp.king_raw.items():
        king[side] = (ctx.normalize_rpsf(params, loge, inp.psf_scale), cos, loge)
    out["king"] = king
    out["aeff"] = {s: ctx.aeff_value(inp.costh, inp.logE, *g) for s, g in inp.aeff_grid.items()}
    out["ltf"] = xa.livetime_efficiency_factors(inp.logE, *inp.effic)
    out["ecb"] = ctx.src_cosbins(src, inp.cube, inp.nside)
    out["wecb"] = ctx.src_cosbins(src, inp.wcube, inp.nside)
    out["expo"] = ctx.exposure(out["ecb"], out["wecb"], out["aeff"]["front"],
                               out["aeff"]["back"], out["ltf"], precision)
    out["upsf"], out["parint"] = ctx.psf_lookup_table_and_partial_integrals(
        king["front"], king["back"], inp.costh, inp.logE, out["aeff"]["front"],
        out["aeff"]["back"], out["ecb"], out["wecb"], out["ltf"], out["expo"], precision)
    return out


# ------------------------------------------------------------------ (a), (b)
def test_gpu_irf_and_exposure_stages(ctx, c0, c0_oracle):
    _log("test_gpu_irf_and_exposure_stages: c0")
    g = _gpu_stages(ctx, c0, c0.src_sph)
    o = c0_oracle
    for side in ("front", "back"):
        assert max_rel(g["king"][side][0], o["king"][side][0]) < 1e-12
        assert max_rel(g["aeff"][side], o["aeff"][side], floor=1e-300) < 1e-13
    np.testing.assert_array_equal(g["ecb"], o["ecb"])
    np.testing.assert_array_equal(g["wecb"], o["wecb"])
    assert max_rel(g["expo"], o["expo"]) < 1e-12
    assert max_rel(g["upsf"], o["upsf"]) < 1e-10
    assert max_rel(g["parint"][:, 1:], o["parint"][:, 1:]) < 1e-10


def test_gpu_psf_stage_entry_points(ctx, orc, c1):
    """bilerp, corrected_exposure_psf, mean_psf and normalize one by one
    (psf.hxx:50-84) against the oracle, and their chain against the fused
    psf_lookup_table_and_partial_integrals."""
    _log("test_gpu_psf_stage_entry_points: c1 (100 sources)")
    g = _gpu_stages(ctx, c1, c1.src_sph)
    corr = {}
    for side in ("front", "back"):
        params, cos, loge = g["king"][side]
        kings = ctx.king(params, cos, loge)
        b = ctx.bilerp(c1.costh, c1.logE, cos, loge, kings)
        assert max_rel(b, orc.bilerp(c1.costh, c1.logE, cos, loge, kings), floor=1e-300) < 1e-13
        corr[side] = ctx.corrected_exposure_psf(b, g["aeff"][side], g["ecb"], g["wecb"], g["ltf"])
        ref = orc.corrected_exposure_psf(b, g["aeff"][side], g["ecb"], g["wecb"], g["ltf"])
        assert max_rel(corr[side], ref, floor=1e-300) < 1e-12
    mean = ctx.mean_psf(corr["front"], corr["back"], g["expo"])
    assert max_rel(mean, orc.mean_psf(corr["front"], corr["back"], g["expo"])) < 1e-15
    par, tot = orc.partial_total_integral(mean)
    normed = ctx.normalize(mean, tot)
    assert max_rel(normed, orc.normalize(mean, tot)) < 1e-15
    # the chain equals the fused entry point
    assert max_rel(normed, g["upsf"]) < 1e-11
    with pytest.raises(Exception):
        params, cos, loge = g["king"]["front"]
        ctx.bilerp(np.array([2.0]), c1.logE, cos, loge, ctx.king(params, cos, loge))


def test_gpu_src_cosbins_many_sources(ctx, orc, c1):
    _log("test_gpu_src_cosbins_many_sources")
    rng = np.random.default_rng(5)
    src = np.stack([rng.uniform(0, 360, 5000), np.degrees(np.arcsin(rng.uniform(-1, 1, 5000)))], 1)
    src[:4] = [[0.0, 90.0], [10.0, -90.0], [359.9999, 0.0], [45.0, 41.8103149]]
    np.testing.assert_array_equal(ctx.src_cosbins(src, c1.cube, 32),
                                  orc.src_cosbins(src, c1.cube, 32))


def test_gpu_bf16x3_contractions(ctx, c1, orc):
    from pipelines import oracle_stages
    _log("test_gpu_bf16x3_contractions: c1, 100 sources")
    o = oracle_stages(orc, c1)
    g = _gpu_stages(ctx, c1, c1.src_sph, precision=1)
    assert max_rel(g["expo"], o["expo"]) < 5e-6
    # the PSF table carries the contraction error (FP32 level) into every node
    assert max_rel(g["upsf"], o["upsf"]) < 1e-5
    rng = np.random.default_rng(0)
    A, B = rng.uniform(-1, 1, (77, 40)), rng.uniform(-1, 1, (40, 123)) * 1e8
    ref = A @ B
    got = ctx.contract(A, B, precision=1)
    assert np.max(np.abs(got - ref)) < 5e-6 * np.max(np.abs(ref))


# ------------------------------------------------------------------ golden chain
def test_gpu_reference_fixtures(ctx, golden):
    """The 3C279 fixtures of the reference, computed on the GPU."""
    _log("test_gpu_reference_fixtures")
    P, A, X, M = golden("irf_psf"), golden("irf_aeff"), golden("exposure_psf"), golden("mm_single_source")
    from oracle import oracle as orc
    import xtsrcmaps_amd as xa
    grids = {}
    for k in ("front", "back"):
        cos, loge, par, _ = orc.prepare_grid(P[k + "_e_lo"], P[k + "_e_hi"], P[k + "_c_lo"],
                                             P[k + "_c_hi"], P[k + "_raw"])
        nrm = ctx.normalize_rpsf(par, loge, P[k + "_scale"])
        grids[k] = (nrm, cos, loge)
    assert max_rel(ctx.king(*grids["front"]), P["king_front"]) < 1e-6
    logE = np.log10(X["energies"])
    costh = xa.costhetas(40, 0.0, True)
    ltf = xa.livetime_efficiency_factors(logE, A["front_effic_p0"], A["front_effic_p1"])
    aeff = {}
    for k in ("front", "back"):
        cos, loge, par, minc = orc.prepare_grid(A[k + "_e_lo"], A[k + "_e_hi"], A[k + "_c_lo"],
                                                A[k + "_c_hi"], A[k + "_raw"])
        aeff[k] = ctx.aeff_value(costh, logE, cos, loge, par[..., 0], minc)
    ecb, wecb = X["ecb"].astype(float), X["wecb"].astype(float)
    expo = ctx.exposure(ecb, wecb, aeff["front"], aeff["back"], ltf)
    assert max_rel(expo, X["exposure"]) < 5e-6
    upsf, parint = ctx.psf_lookup_table_and_partial_integrals(
        grids["front"], grids["back"], costh, logE, aeff["front"], aeff["back"], ecb, wecb, ltf,
        expo)
    assert max_rel(upsf[:, 0, :], X["upsf_peak"]) < 1e-6
    # single-source AIT cubature maps (fixture layout [W, H, E])
    i = int(np.argmin(np.hypot(X["src_dirs"][:, 0] - M["src_sph"][0],
                               X["src_dirs"][:, 1] - M["src_sph"][1])))
    w = xa.Wcs.make(M["naxis"], M["crpix"], M["crval"], M["cdelt"], float(M["crota2"]),
                    str(M["proj"]))
    lat = M["lattice"]
    g3 = np.array([[-1, -1, 1], [-1, 1, 1], [1, 1, 1], [1, -1, 1]], float)
    g3[:, :2] /= np.sqrt(3.0)
    for name, rule in (("mm_sqd3", g3), ("mm_sqd11", M["rule_square_deg11"]),
                       ("mm_asqd12_single_src", M["rule_asym_square_d12"])):
        for flags in (xa.COMPUTE_F64, xa.COMPUTE_F32):
            m = ctx.cubature_convolve_map_psf(w, np.array([M["src_sph"]]), upsf[i:i + 1], rule,
                                              flags=flags)[0]
            got = m.transpose(1, 0, 2)[np.ix_(lat, lat)]
            assert max_rel(got, M[name]) < 1e-5, (name, flags)


# ------------------------------------------------------------------ (c) model map
def _oracle_model_map(orc, inp, stages, n_src=None):
    from pipelines import oracle_rule, oracle_wcs
    S = stages["upsf"].shape[0] if n_src is None else n_src
    w = oracle_wcs(orc, inp.wcs)
    return orc.point_src_model_map_wcs(inp.wcs.naxis1, inp.wcs.naxis2, w, stages["src"][:S],
                                       stages["expo"][:S], stages["upsf"][:S],
                                       stages["parint"][:S], oracle_rule(inp), details=True)


def test_gpu_map_matches_gtsrcmaps_through_the_writer(ctx, orc, golden, tmp_path):
    """FermiTools' canonical source (analysis/canonical_srcmap.fits) on the
    GPU path: FP32 cubature within 1e-5 of the FP64 oracle, and written as
    float32 through the FITS writer, within the oracle's own agreement with
    gtsrcmaps' cube (median 2e-5, 99th percentile 1.5e-3, max 3e-3)."""
    import xtsrcmaps_amd as xa
    import fits_reader as fr
    from pipelines import canonical_source, canonical_cube
    _log("test_gpu_map_matches_gtsrcmaps: 4FGL J1118.2-0415, 100x100 AIT, 38 energies")
    c = canonical_source(orc, golden)
    w = c["wcs"]
    ow = orc.make_wcs(w["naxis"], w["crpix"], w["crval"], w["cdelt"], w["crota2"], w["proj"])
    H, W = w["naxis"]
    ref = orc.point_src_model_map_wcs(H, W, ow, c["src"], c["expo"], c["upsf"], c["parint"],
                                      c["rule"])
    wcs = xa.Wcs(naxis1=H, naxis2=W, crpix1=w["crpix"][0], crpix2=w["crpix"][1],
                 crval1=w["crval"][0], crval2=w["crval"][1], cdelt1=w["cdelt"][0],
                 cdelt2=w["cdelt"][1], crota2=w["crota2"], proj=xa.PROJ_AIT, is_galactic=0)
    got = ctx.point_src_model_map_wcs(wcs, c["src"], c["expo"], c["upsf"], c["parint"], c["rule"],
                                      flags=xa.LAYOUT_SHWE | xa.OUT_F64 | xa.COMPUTE_F32)
    assert max_rel(got, ref) < 1e-5
    path = tmp_path / "srcmap.fits"
    xa.write_src_model(path, got, [c["name"]], float32=True, wcs=wcs)
    (h, d), = fr.split_hdus(path.read_bytes())[1:]
    v = fr.verify(h, d)
    assert v["hdu_sum_is_minus_zero"] and v["datasum_ok"]
    ours = fr.image(h, d).astype(np.float64)
    can = canonical_cube(golden).astype(np.float64)
    rel = np.abs(ours - can) / np.abs(can)
    _log(f"  vs gtsrcmaps: median {np.median(rel):.2e} p99 {np.percentile(rel, 99):.2e} "
         f"max {rel.max():.2e}")
    assert np.median(rel) < 2e-5 and np.percentile(rel, 99) < 1.5e-3 and rel.max() < 3e-3


def test_gpu_model_map_c0_fp64_and_fp32(ctx, orc, c0, c0_oracle):
    import xtsrcmaps_amd as xa
    _log("test_gpu_model_map_c0: FP64 and FP32 cubature vs oracle")
    ref, det = _oracle_model_map(orc, c0, c0_oracle)
    o = c0_oracle
    m64 = ctx.point_src_model_map_wcs(c0.wcs, o["src"], o["expo"], o["upsf"], o["parint"],
                                      c0.rule, flags=xa.LAYOUT_SHWE | xa.OUT_F64 | xa.COMPUTE_F64)
    assert max_rel(m64, ref) < 1e-10
    m32 = ctx.point_src_model_map_wcs(c0.wcs, o["src"], o["expo"], o["upsf"], o["parint"],
                                      c0.rule, flags=xa.LAYOUT_SHWE | xa.OUT_F64 | xa.COMPUTE_F32)
    rel = np.abs(m32 - ref) / np.abs(ref)
    _log(f"  c0 FP32 cubature: max rel {rel.max():.3e} median {np.median(rel):.3e}")
    assert rel.max() < 1e-5
    # the parts: solid angle, boundary radius, correction factor
    from pipelines import oracle_wcs
    w = oracle_wcs(orc, c0.wcs)
    assert max_rel(ctx.solid_angle(c0.wcs), orc.solid_angle(100, 100, w)) < 1e-12
    rad, fov = ctx.psf_boundary_radius(c0.wcs, o["src"])
    np.testing.assert_allclose(rad, det["radius"], rtol=1e-13)
    np.testing.assert_array_equal(fov, det["in_fov"].astype(bool))
    cf = ctx.map_correction_factor(det["inv_mapinteg"], det["radius"], det["in_fov"], o["upsf"],
                                   o["parint"])
    assert max_rel(cf, det["cf"]) < 1e-12


def test_gpu_model_map_c1_subset_fp32(ctx, orc, c1):
    import xtsrcmaps_amd as xa
    from pipelines import oracle_stages
    _log("test_gpu_model_map_c1_subset: 4 sources, 200x200, 8x8 GL, FP32")
    o = oracle_stages(orc, c1, n_src=4)
    ref, _ = _oracle_model_map(orc, c1, o)
    got = ctx.point_src_model_map_wcs(c1.wcs, o["src"], o["expo"], o["upsf"], o["parint"],
                                      c1.rule, flags=xa.LAYOUT_SEHW | xa.OUT_F32 | xa.COMPUTE_F32)
    got = got.transpose(0, 2, 3, 1).astype(np.float64)
    rel = np.abs(got - ref) / np.abs(ref)
    _log(f"  c1 subset: max rel {rel.max():.3e} median {np.median(rel):.3e}")
    assert rel.max() < 1e-5


def test_gpu_layouts_and_raw_cubature_agree(ctx, c1):
    import xtsrcmaps_amd as xa
    _log("test_gpu_layouts_and_raw_cubature_agree")
    rng = np.random.default_rng(2)
    S, E = 3, 41                                # two energy chunks
    upsf = rng.uniform(0.5, 2.0, (S, 401, E))
    src = c1.src_sph[:S]
    a = ctx.cubature_convolve_map_psf(c1.wcs, src, upsf, c1.rule,
                                      flags=xa.LAYOUT_SHWE | xa.OUT_F32)
    b = ctx.cubature_convolve_map_psf(c1.wcs, src, upsf, c1.rule,
                                      flags=xa.LAYOUT_SEHW | xa.OUT_F32)
    np.testing.assert_array_equal(a, b.transpose(0, 2, 3, 1))
    # constant table -> constant map (weights sum to 1)
    const = np.full((S, 401, E), 3.25)
    m = ctx.cubature_convolve_map_psf(c1.wcs, src, const, c1.rule, flags=xa.OUT_F64)
    np.testing.assert_allclose(m, 3.25, rtol=2e-6)


@pytest.mark.parametrize("proj", ["CAR", "AIT"])
def test_gpu_non_square_images(ctx, orc, c1, proj):
    """H != W (the reference loops h < Nw, quirk Q3: this build loops h < Nh),
    unequal CDELTs, an oblique field, a 3 x 4 tensor rule (tensor loop with
    nu != nv) and the degree-7 list rule, both layouts, FP32 and FP64
    cubature, against the FP64 oracle."""
    import xtsrcmaps_amd as xa
    from pipelines import DEG7, oracle_wcs
    _log(f"test_gpu_non_square_images: 48x72 {proj}")
    H, W = 48, 72
    wcs = xa.Wcs.make((H, W), (20.5, 40.5), (210.0, 35.0), (-0.11, 0.08), 0.0, proj)
    w = oracle_wcs(orc, wcs)
    # sources inside, near an edge and outside the image
    src = ctx.pix2sph(wcs, np.array([[24.0, 30.0], [2.0, 70.0], [60.0, 10.0]]))
    rng = np.random.default_rng(21)
    S, E = 3, 9
    upsf = rng.uniform(0.1, 1.0, (S, 401, E)) * np.exp(-np.linspace(0, 8, 401))[None, :, None]
    upsf, parint, _ = ctx.partial_total_integral(upsf)
    expo = rng.uniform(1e8, 1e9, (S, E))
    x3, w3 = np.polynomial.legendre.leggauss(3)
    x4, w4 = np.polynomial.legendre.leggauss(4)
    rule34 = np.array([[u, v, wu * wv] for u, wu in zip(x3, w3) for v, wv in zip(x4, w4)])
    for rule in (rule34, DEG7):
        ref = orc.point_src_model_map_wcs(H, W, w, src, expo, upsf, parint, rule)
        for flags, tol in ((xa.LAYOUT_SHWE | xa.OUT_F64 | xa.COMPUTE_F64, 1e-10),
                           (xa.LAYOUT_SHWE | xa.OUT_F64 | xa.COMPUTE_F32, 1e-5),
                           (xa.LAYOUT_SEHW | xa.OUT_F32 | xa.COMPUTE_F32, 1e-5)):
            got = ctx.point_src_model_map_wcs(wcs, src, expo, upsf, parint, rule, flags=flags)
            if flags & xa.LAYOUT_SEHW:
                got = got.transpose(0, 2, 3, 1)
            r = max_rel(got, ref)
            _log(f"  rule {len(rule)} flags {flags:#x}: max rel {r:.2e}")
            assert got.shape == (S, H, W, E) and r < tol


def test_gpu_edge_cases(ctx, orc, c0):
    import xtsrcmaps_amd as xa
    from pipelines import oracle_wcs
    _log("test_gpu_edge_cases: out-of-FOV source, source on a pixel centre, E=1, 1-point rule")
    w = oracle_wcs(orc, c0.wcs)
    src = np.array([[180.0 - 0.1 * 60, 0.0],          # outside the 100x100 0.1 deg image
                    [180.05, 0.05],                   # on a pixel centre
                    [180.0, 4.98]])                   # inside the 3.5 px border
    rng = np.random.default_rng(9)
    upsf = rng.uniform(0.1, 1.0, (3, 401, 1))
    upsf, parint, _ = ctx.partial_total_integral(upsf)
    expo = np.array([[2.0], [3.0], [4.0]])
    rule = np.array([[0.0, 0.0, 4.0]])
    ref, det = orc.point_src_model_map_wcs(100, 100, w, src, expo, upsf, parint, rule,
                                           details=True)
    assert list(det["in_fov"]) == [0, 1, 0]
    got = ctx.point_src_model_map_wcs(c0.wcs, src, expo, upsf, parint, rule,
                                      flags=xa.OUT_F64 | xa.COMPUTE_F64)
    assert max_rel(got, ref) < 1e-10
    with pytest.raises(xa.XtsError):
        ctx.point_src_model_map_wcs(c0.wcs, src, expo, upsf, parint, rule, flags=0x40)


def test_gpu_model_map_c1_full_vs_oracle(ctx, orc, c1):
    """All 100 sources of c1 at full size (200x200, 37 energies, 8x8
    Gauss-Legendre) against the FP64 oracle of the same workload, stages
    included: every value within 1e-5; finite and positive."""
    import xtsrcmaps_amd as xa
    from pipelines import oracle_stages
    _log("test_gpu_model_map_c1_full_vs_oracle: 100 x 37 x 200 x 200")
    o = oracle_stages(orc, c1)
    ref, _ = _oracle_model_map(orc, c1, o)
    g = _gpu_stages(ctx, c1, c1.src_sph)
    assert max_rel(g["upsf"], o["upsf"]) < 1e-10 and max_rel(g["expo"], o["expo"]) < 1e-10
    m32 = ctx.point_src_model_map_wcs(c1.wcs, c1.src_sph, g["expo"], g["upsf"], g["parint"],
                                      c1.rule, flags=xa.LAYOUT_SEHW | xa.OUT_F32)
    assert np.isfinite(m32).all() and (m32 > 0).all()
    got = m32.transpose(0, 2, 3, 1).astype(np.float64)
    rel = np.abs(got - ref) / np.abs(ref)
    _log(f"  c1 full: max rel {rel.max():.3e} median {np.median(rel):.3e}")
    assert rel.max() < 1e-5


def test_gpu_correction_factor_exact_radius_flag(ctx, orc, c1, golden):
    """XTS_CF_EXACT_RADIUS: the map scaled by the FermiTools-consistent
    correction factor instead of the as-coded one, i.e. the oracle's map times
    cf_exact / cf_coded per (source, energy), on a c1 subset (a third of c1's
    sources lie within 1 degree of an edge) and on the gtsrcmaps canonical
    source (AIT, degree-7 rule)."""
    import xtsrcmaps_amd as xa
    from pipelines import canonical_source, oracle_stages, oracle_wcs
    _log("test_gpu_correction_factor_exact_radius_flag")
    o = oracle_stages(orc, c1, n_src=12)
    cases = [(c1.wcs, o["src"], o["expo"], o["upsf"], o["parint"], c1.rule)]
    cs = canonical_source(orc, golden)
    cases.append((xa.Wcs.make(cs["wcs"]["naxis"], cs["wcs"]["crpix"], cs["wcs"]["crval"],
                              cs["wcs"]["cdelt"], cs["wcs"]["crota2"], cs["wcs"]["proj"]),
                  cs["src"], cs["expo"], cs["upsf"], cs["parint"], cs["rule"]))
    for wcs, src, expo, upsf, parint, rule in cases:
        w = oracle_wcs(orc, wcs)
        H, W = wcs.naxis1, wcs.naxis2
        ref, det = orc.point_src_model_map_wcs(H, W, w, src, expo, upsf, parint, rule,
                                               details=True)
        cfx = orc.map_correction_factor(det["inv_mapinteg"], det["radius"], det["in_fov"], upsf,
                                        parint, exact=True)
        scale = np.where(det["cf"] != 0, cfx / np.where(det["cf"] != 0, det["cf"], 1), 0)
        expect = ref * scale[:, None, None, :]
        got = ctx.point_src_model_map_wcs(wcs, src, expo, upsf, parint, rule,
                                          flags=xa.OUT_F64 | xa.COMPUTE_F32 | xa.CF_EXACT_RADIUS)
        nz = expect != 0
        rel = np.abs(got[nz] - expect[nz]) / np.abs(expect[nz])
        _log(f"  {H}x{W}: exact-radius map max rel {rel.max():.3e}; "
             f"factor changed by up to {np.max(np.abs(scale - 1)):.2e}")
        assert rel.max() < 1e-5 and np.all(got[~nz] == 0)


_C3P_CHILD = r"""
import sys
import numpy as np
import torch
sys.path.insert(0, sys.argv[1])
import bench
import xtsrcmaps_amd as xa
cfg, inp, ctx, d = bench.setup("c3p", surrogate=False)
bench.step(ctx, inp, d)
torch.cuda.synchronize()
out = d["out"]
ok = bool(torch.isfinite(out).all()) and bool((out > 0).all())
# 32 sources spread over the list against the FP64 oracle of the same
# workload (its own exposure and mean PSF, the same image and rule)
sys.path.insert(0, sys.argv[1] + "/tests")
from oracle import oracle as orc
from pipelines import oracle_stages, oracle_rule, oracle_wcs
pick = np.linspace(0, cfg.n_src - 1, 32).round().astype(int)
sub = type(inp)(**{**inp.__dict__, "src_sph": inp.src_sph[pick]})
o = oracle_stages(orc, sub)
ref = orc.point_src_model_map_wcs(cfg.npix, cfg.npix, oracle_wcs(orc, inp.wcs), o["src"],
                                  o["expo"], o["upsf"], o["parint"], oracle_rule(inp))
got = out[torch.as_tensor(pick, device=out.device)].double().cpu().numpy().transpose(0, 2, 3, 1)
rel = np.abs(got - ref) / np.abs(ref)
print(f"finite_positive={ok} 32-source max_rel={rel.max():.3e} median={np.median(rel):.3e}",
      flush=True)
sys.exit(0 if ok and rel.max() < 1e-5 else 1)
"""


_SPLIT_CHILD = r"""
import sys
import numpy as np
import torch
sys.path.insert(0, sys.argv[1])
import bench
import xtsrcmaps_amd as xa
cfg, inp, ctx, d = bench.setup("c1", surrogate=False)
flags = xa.LAYOUT_SEHW | xa.OUT_F32 | xa.COMPUTE_F32
E = cfg.n_energy
for _ in range(2):
    bench.step(ctx, inp, d)          # row weights beside the PSF on a second stream
torch.cuda.synchronize()
a = d["out"].clone()
ctx.point_src_model_map_wcs_dev(inp.wcs, d["src"], d["expo"], d["upsf"], d["parint"], d["out"],
                                inp.rule, flags=flags)
torch.cuda.synchronize()
b = d["out"].clone()
ctx.point_src_model_map_rows_dev(inp.wcs, d["src"], E, inp.rule, flags=flags)
ctx.point_src_model_map_finish_dev(inp.wcs, d["src"], d["expo"], d["upsf"], d["parint"],
                                   d["out"], inp.rule, flags=flags)
torch.cuda.synchronize()
c = d["out"].clone()
rel = lambda x, y: float(((x - y).abs() / y.abs()).max())
print(f"two-stream vs one call {rel(a, b):.2e}, split vs one call {rel(c, b):.2e}", flush=True)
# the second half without a first half with the same arguments is refused
# before any launch
refused = 0
try:
    ctx.point_src_model_map_finish_dev(inp.wcs, d["src"], d["expo"], d["upsf"], d["parint"],
                                       d["out"], inp.rule, flags=flags)
except Exception:
    refused += 1
ctx.point_src_model_map_rows_dev(inp.wcs, d["src"], E, inp.rule, flags=flags)
try:
    ctx.point_src_model_map_finish_dev(inp.wcs, d["src"][:50], d["expo"][:50], d["upsf"][:50],
                                       d["parint"][:50], d["out"], inp.rule, flags=flags)
except Exception:
    refused += 1
# any other model map between the halves, or a change of stream, cancels the
# pending first half: its second half is refused before any launch
ctx.point_src_model_map_rows_dev(inp.wcs, d["src"], E, inp.rule, flags=flags)
ctx.point_src_model_map_wcs_dev(inp.wcs, d["src"], d["expo"], d["upsf"], d["parint"], d["out"],
                                inp.rule, flags=flags)
try:
    ctx.point_src_model_map_finish_dev(inp.wcs, d["src"], d["expo"], d["upsf"], d["parint"],
                                       d["out"], inp.rule, flags=flags)
except Exception:
    refused += 1
ctx.point_src_model_map_rows_dev(inp.wcs, d["src"], E, inp.rule, flags=flags)
ctx.set_stream(d["stream"].cuda_stream)
try:
    ctx.point_src_model_map_finish_dev(inp.wcs, d["src"], d["expo"], d["upsf"], d["parint"],
                                       d["out"], inp.rule, flags=flags)
except Exception:
    refused += 1
# so does a call that rebuilds the geometry slots (the solid angle of another
# image, the diffuse map)
for other in ("solid_angle", "diffuse"):
    ctx.point_src_model_map_rows_dev(inp.wcs, d["src"], E, inp.rule, flags=flags)
    if other == "solid_angle":
        ctx.solid_angle(xa.Wcs.make((16, 12), (8.5, 6.5), (10.0, 20.0), (-0.1, 0.1)))
    else:
        cube = xa.Wcs.make((36, 19), (18.5, 10.0), (0.0, 0.0), (-10.0, 10.0), galactic=True)
        ctx.diffuse_src_model_map_wcs(xa.Wcs.make((16, 12), (8.5, 6.5), (10.0, 20.0),
                                                  (-0.5, 0.5), galactic=True),
                                      cube, np.ones((2, 19, 36), np.float32),
                                      np.ones((401, 2)), np.ones(2), 5.0)
    try:
        ctx.point_src_model_map_finish_dev(inp.wcs, d["src"], d["expo"], d["upsf"], d["parint"],
                                           d["out"], inp.rule, flags=flags)
    except Exception:
        refused += 1
torch.cuda.synchronize()
print(f"refused {refused}", flush=True)
ok = rel(a, b) < 2e-7 and rel(c, b) < 2e-7 and refused == 6
sys.exit(0 if ok else 1)
"""


def test_gpu_model_map_split_halves():
    """The model map in two calls (row weights, then table, factor and emit)
    equals the one-call entry point, also with the PSF stages computed on a
    second context and stream between the halves, as the bench runs them;
    atomically accumulated map integrals may differ in the last bit, hence
    2e-7.  A second half without a matching first half is refused, and so is
    one after another model map, a change of stream, or a call that rebuilds
    the geometry (solid angle, diffuse map).  (Child process: torch before
    the library.)"""
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    _log("test_gpu_model_map_split_halves: c1, two streams (child process)")
    r = subprocess.run([sys.executable, "-c", _SPLIT_CHILD, root], capture_output=True,
                       text=True, timeout=240)
    _log("  " + r.stdout.strip()[-300:])
    assert r.returncode == 0, r.stdout[-1000:] + r.stderr[-2000:]


def test_gpu_model_map_c3p_full_size_properties():
    """2,000 sources on a 400x400 image, 37 energies: 1.18e10 values, a 47 GB
    FP32 map, through the device entry points as the bench runs them.  Every
    value finite and positive; 32 sources spread over the list within 1e-5 of
    the FP64 oracle of the same workload.  (Large-S paths: pool segments, grid sizes, 64-bit
    offsets.)  Run in a child process that imports torch before the library,
    as bench.py does: the two must share torch's HIP runtime."""
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    _log("test_gpu_model_map_c3p_full: 2000 x 37 x 400 x 400 (child process)")
    r = subprocess.run([sys.executable, "-c", _C3P_CHILD, root], capture_output=True, text=True,
                       timeout=240)
    _log("  " + r.stdout.strip()[-300:])
    assert r.returncode == 0, r.stderr[-2000:]
