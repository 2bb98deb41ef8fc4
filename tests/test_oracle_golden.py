"""Pin the FP64 oracle against the reference's own test vectors.

Every fixture under tests/golden/ was extracted from the reference tree by
tests/golden/make_golden.py (data only).  Tolerances are set by the float32
precision of the reference's header fixtures; the achieved floors are far
below the reference tests' own rtol (1e-5, king 1e-2).
"""
import numpy as np
import pytest

from conftest import max_rel, near_relative


def test_separations_match_reference_table(orc, golden):
    # psf/psf.hxx:24-31 vs tests/test_king.cxx:26-108 (rtol 1e-5 there)
    sep = orc.separations()
    exp = golden("irf_psf")["separations"]
    assert sep[0] == 0.0 and exp[0] == 0.0
    assert max_rel(sep[1:], exp[1:]) < 1e-5
    assert sep[1] == pytest.approx(1e-4, rel=1e-15)
    assert sep[400] == pytest.approx(70.0, rel=1e-13)


def test_prepare_grid_matches_psf3_rpsf(psf_grids, golden):
    P = golden("irf_psf")
    for k in ("front", "back"):
        np.testing.assert_array_equal(psf_grids[k]["raw"], P[k + "_rpsf"])


def test_normalize_rpsf_matches_fixture(psf_grids, golden):
    # test_irf.psf.cpp:113-129: fixture holds normalised NCORE, unscaled sigmas
    P = golden("irf_psf")
    for k in ("front", "back"):
        raw, nrm = psf_grids[k]["raw"], psf_grids[k]["norm"]
        sf = nrm[..., 2] / raw[..., 2]
        np.testing.assert_allclose(nrm[..., 3] / raw[..., 3], sf, rtol=1e-14)
        cmp = nrm.copy()
        cmp[..., 2] /= sf
        cmp[..., 3] /= sf
        assert max_rel(cmp, P[k + "_norm"]) < 1e-6


def test_king_table_matches_fixture(orc, psf_grids, golden):
    K = orc.king(psf_grids["front"]["norm"])
    assert max_rel(K, golden("irf_psf")["king_front"]) < 1e-6


def test_livetime_efficiency_factors(chain, golden):
    A = golden("irf_aeff")
    assert near_relative(chain["ltf"][0], A["livetime_factor1"]).all()
    assert near_relative(chain["ltf"][1], A["livetime_factor2"]).all()


def test_costhetas_match_fixture(chain, golden):
    exp = golden("exposure_psf")["exp_costhetas"]
    # the fixture is printed to 6 significant digits
    assert max_rel(chain["costh"], exp) < 1e-5


def test_aeff_value_matches_fixture(chain, golden):
    X = golden("exposure_psf")
    for k in ("front", "back"):
        assert near_relative(chain["aeff"][k], X["aeff_%s_c_e" % k]).all()


def test_exposure_contractions_match_fixture(chain, golden):
    X = golden("exposure_psf")
    assert max_rel(chain["expo"], X["exposure"]) < 5e-6
    for n, p in zip(("lef", "lwf", "leb", "lwb"), chain["parts"]):
        assert max_rel(p, X["exposure_" + n]) < 5e-6


def test_exposure_pre_ltf_contractions(orc, chain, golden):
    A = golden("irf_aeff")
    for side in ("front", "back"):
        e = chain["ecb"] @ chain["aeff"][side]
        w = chain["wecb"] @ chain["aeff"][side]
        # printed to 6 significant digits in the reference header
        assert max_rel(e, A[side + "_exp_aeff"]) < 1e-5
        assert max_rel(w, A[side + "_wexp_aeff"]) < 1e-5


def test_mean_psf_peak_matches_fixture(chain, golden):
    # uPsf_peak_SE.bin = normalised uPsf[:, 0, :]; un-normalised differs by
    # ~1e-3, so the radial normalisation is exercised.
    peak = golden("exposure_psf")["upsf_peak"]
    assert max_rel(chain["upsf"][:, 0, :], peak) < 1e-6


def test_partial_integral_is_cumulative_trapezoid(chain):
    # psf.partial_total_integral.cxx:23-48 integrates the linear interpolant
    # of y = 2 pi sin(theta) psf exactly per segment: a cumulative trapezoid.
    from scipy.integrate import cumulative_trapezoid
    from oracle import oracle
    th = np.deg2rad(oracle.separations())
    upsf, parint = chain["upsf"], chain["parint"]
    for s in (0, 17, 262):
        y = 2 * np.pi * np.sin(th)[:, None] * upsf[s]
        ct = cumulative_trapezoid(y, th, axis=0, initial=0.0)
        # upsf is normalised by the same total, so ct ends at 1
        np.testing.assert_allclose(ct[-1], 1.0, rtol=1e-12)
        np.testing.assert_allclose(parint[s], ct, rtol=1e-10, atol=1e-14)


def _single_source_lut(chain, golden):
    M, X = golden("mm_single_source"), golden("exposure_psf")
    dirs = X["src_dirs"]
    i = int(np.argmin(np.hypot(dirs[:, 0] - M["src_sph"][0], dirs[:, 1] - M["src_sph"][1])))
    return M, chain["upsf"][i:i + 1]


def _lattice_points(orc, M, w, rule=None, npix=None):
    lat = M["lattice"]
    pts = orc.square_ptswts(100, 100, w, rule) if rule is not None else None
    if npix is not None:
        pts = orc.naive_ptswts(100, 100, w, npix)
    return np.ascontiguousarray(pts[np.ix_(lat, lat)])


def _deg3_rule():
    r = np.array([[-1, -1, 1], [-1, 1, 1], [1, 1, 1], [1, -1, 1]], float)
    r[:, :2] /= np.sqrt(3.0)
    return r


@pytest.fixture(scope="module")
def mm_maps(orc, chain, golden):
    M, lut = _single_source_lut(chain, golden)
    w = orc.make_wcs(M["naxis"], M["crpix"], M["crval"], M["cdelt"],
                     float(M["crota2"]), str(M["proj"]))
    src = np.array([M["src_sph"]])
    rules = {"mm_sqd3": _deg3_rule(), "mm_sqd11": M["rule_square_deg11"],
             "mm_asqd12_single_src": M["rule_asym_square_d12"]}
    out = {}
    n = len(M["lattice"])
    for name, rule in rules.items():
        pts = _lattice_points(orc, M, w, rule=rule)
        # the fixture layout is [W, H, E] (test.mm.psf_value.single.cpp:161-174)
        out[name] = orc.cubature_convolve_map_psf(n, n, src, lut, w, pts)[0].transpose(1, 0, 2)
    return M, w, src, lut, rules, out


def test_pixel_cubature_maps_match_fixtures(mm_maps):
    M, _, _, _, _, out = mm_maps
    for name, m in out.items():
        assert max_rel(m, M[name]) < 1e-5, name


def test_rule_difference_maps_pin_point_placement(mm_maps):
    # The LUT error (float32 inputs) cancels in rule differences, so these
    # pin the cubature point placement and weights much more tightly.
    M, _, _, _, _, out = mm_maps
    for a, b in (("mm_sqd3", "mm_sqd11"), ("mm_sqd11", "mm_asqd12_single_src")):
        fd, od = M[a] - M[b], out[a] - out[b]
        scale = np.max(np.abs(fd))
        assert scale > 1e-9 * np.max(np.abs(M[a]))
        assert np.max(np.abs(od - fd)) < 2e-2 * scale


def test_planted_rule_transpose_is_detected(orc, mm_maps):
    # A plausible mistake -- u and v swapped in the bilinear blend -- moves
    # the asymmetric d12 map beyond the pinned tolerance.
    M, w, src, lut, rules, out = mm_maps
    bad = rules["mm_asqd12_single_src"][:, [1, 0, 2]]
    n = len(M["lattice"])
    pts = _lattice_points(orc, M, w, rule=bad)
    m = orc.cubature_convolve_map_psf(n, n, src, lut, w, pts)[0].transpose(1, 0, 2)
    fd = M["mm_asqd12_single_src"] - M["mm_sqd11"]
    assert np.max(np.abs((m - out["mm_sqd11"]) - fd)) > 5e-2 * np.max(np.abs(fd))


def test_midpoint_4096_searchsep_matches_fixture(orc, chain, golden):
    M, lut = _single_source_lut(chain, golden)
    w = orc.make_wcs(M["naxis"], M["crpix"], M["crval"], M["cdelt"],
                     float(M["crota2"]), str(M["proj"]))
    lat = M["lattice"][::2]
    pts = np.ascontiguousarray(orc.naive_ptswts(100, 100, w, 64)[np.ix_(lat, lat)])
    n = len(lat)
    m = orc.cubature_convolve_map_psf(n, n, np.array([M["src_sph"]]), lut, w, pts,
                                      searchsep=True)[0].transpose(1, 0, 2)
    assert max_rel(m, M["mm_psf_value_4096"][::2, ::2]) < 1e-5


def test_fermitools_fov_and_boundary_radius(orc, golden):
    """FermiTools' own in-bounds flags and PSF radii for the 263 sources of
    the 3C279 ROI (ft_mm_withinBounds.bin, ft_mm_psf_radius.bin).  The
    flags must agree exactly.  FermiTools computes the radius its own way
    while model_map.psf_boundary_radius.cxx:34-59 takes the nearest of the
    half-pixel edge points, so the radii agree only to 5e-4 relative (a
    loose pin of the edge geometry and of the AIT projection)."""
    M, X, F = golden("mm_single_source"), golden("exposure_psf"), golden("fermitools_boundary")
    w = orc.make_wcs(M["naxis"], M["crpix"], M["crval"], M["cdelt"], float(M["crota2"]),
                     str(M["proj"]))
    nh, nw = int(M["naxis"][0]), int(M["naxis"][1])
    rad, fov = orc.psf_boundary_radius(nh, nw, X["src_dirs"], w)
    np.testing.assert_array_equal(fov.astype(bool), F["within_bounds"].astype(bool))
    inside = fov.astype(bool)
    assert inside.sum() == 47
    rel = (rad[inside] - F["psf_radius"][inside]) / F["psf_radius"][inside]
    assert np.max(np.abs(rel)) < 5e-4
