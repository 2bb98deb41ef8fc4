"""Pin the oracle where no reference fixture does: closed forms, library
routines and invariants that a dropped term, wrong sign or wrong index
would break."""
import numpy as np
import pytest


# ------------------------------------------------------------------ HEALPix
def test_ang2pix_base_face_centres(orc):
    # nside = 1: the 12 base pixels, NESTED numbering = face number.
    for f in range(12):
        ring = f // 4
        z = (2.0 / 3.0, 0.0, -2.0 / 3.0)[ring]
        phi = (np.pi / 4 + (f % 4) * np.pi / 2) if ring != 1 else (f % 4) * np.pi / 2
        assert orc.ang2pix_nest(np.arccos(z), phi, 1) == f


def test_ang2pix_nested_hierarchy_and_equal_area(orc):
    rng = np.random.default_rng(7)
    n = 20000
    z = rng.uniform(-1, 1, n)
    phi = rng.uniform(0, 2 * np.pi, n)
    th = np.arccos(z)
    for nside in (2, 4, 8, 16, 32):
        for k in range(0, n, 997):
            child = orc.ang2pix_nest(th[k], phi[k], nside * 2)
            parent = orc.ang2pix_nest(th[k], phi[k], nside)
            assert child >> 2 == parent
    # equal area: uniform points fill the 48 nside-2 pixels evenly
    pix = np.array([orc.ang2pix_nest(th[k], phi[k], 2) for k in range(n)])
    counts = np.bincount(pix, minlength=48)
    assert counts.min() > 0.8 * n / 48 and counts.max() < 1.2 * n / 48
    assert pix.max() == 47


def test_ang2pix_polar_caps_and_phi_wrap(orc):
    # phi and phi + 2 pi land in the same pixel; near-pole branch is consistent
    for th in (1e-4, 0.005, 0.3, 1.2, np.pi - 0.005):
        for ph in (0.1, 2.0, 4.0):
            assert orc.ang2pix_nest(th, ph, 32) == orc.ang2pix_nest(th, ph + 2 * np.pi, 32)
    assert orc.ang2pix_nest(1e-6, 0.3, 32) // (32 * 32) == 0
    assert orc.ang2pix_nest(np.pi - 1e-6, 0.3, 32) // (32 * 32) == 8


# ------------------------------------------------------------------ WCS
def test_car_plate_carree_closed_form(orc):
    w = orc.make_wcs((100, 100), (50.5, 50.5), (180.0, 0.0), (-0.1, 0.1), 0.0, "CAR")
    for p1, p2 in ((1.0, 1.0), (50.5, 50.5), (100.0, 37.25), (0.5, 100.5)):
        lon, lat = orc.pix2sph(w, p1, p2)
        assert lon == pytest.approx(180.0 - 0.1 * (p1 - 50.5), abs=1e-12)
        assert lat == pytest.approx(0.1 * (p2 - 50.5), abs=1e-12)


@pytest.mark.parametrize("proj,crval", [("CAR", (180.0, 0.0)), ("CAR", (30.0, 45.0)),
                                        ("AIT", (193.98, -5.82)), ("AIT", (10.0, 60.0))])
def test_wcs_round_trip_and_reference_point(orc, proj, crval):
    w = orc.make_wcs((200, 200), (100.5, 100.5), crval, (-0.05, 0.05), 0.0, proj)
    lon, lat = orc.pix2sph(w, 100.5, 100.5)
    assert lat == pytest.approx(crval[1], abs=1e-10)
    assert (lon - crval[0] + 180) % 360 - 180 == pytest.approx(0.0, abs=1e-10)
    rng = np.random.default_rng(1)
    for p1, p2 in rng.uniform(0.5, 200.5, (50, 2)):
        lo, la = orc.pix2sph(w, p1, p2)
        q1, q2 = orc.sph2pix(w, lo, la)
        assert q1 == pytest.approx(p1, abs=1e-8) and q2 == pytest.approx(p2, abs=1e-8)


def test_crota2_rotates_intermediate_frame(orc):
    # 90 degrees of CROTA2 with equal |cdelt| maps a step along axis 1 onto
    # the latitude axis (Greisen & Calabretta 2002, eq. 189-190).
    w = orc.make_wcs((100, 100), (50.5, 50.5), (180.0, 0.0), (-0.1, 0.1), 90.0, "CAR")
    lon, lat = orc.pix2sph(w, 60.5, 50.5)
    assert lon == pytest.approx(180.0, abs=1e-9)
    assert lat == pytest.approx(-1.0, abs=1e-9)


def test_solid_angle_close_to_exact_car_area(orc):
    # Exact plate-carree area of the quad the reference uses (pixel centres
    # (h+1, w+1) .. (h+2, w+2)): d_alpha (sin d2 - sin d1).  The flat-triangle
    # formula (math/solid_angle.hpp:42-79) agrees to O(pixel^2).
    w = orc.make_wcs((40, 40), (20.5, 20.5), (180.0, 30.0 * 0), (-0.1, 0.1), 0.0, "CAR")
    om = orc.solid_angle(40, 40, w)
    for h, ww in ((0, 0), (10, 31), (39, 39)):
        lat1 = 0.1 * (ww + 1 - 20.5)
        lat2 = lat1 + 0.1
        exact = np.deg2rad(0.1) * (np.sin(np.deg2rad(lat2)) - np.sin(np.deg2rad(lat1)))
        assert om[h, ww] == pytest.approx(exact, rel=1e-5)


# ------------------------------------------------------------------ quadrature / King
def test_legendre_matches_numpy(orc):
    for n in (8, 64):
        x, w = orc.legendre_rw(n)
        xr, wr = np.polynomial.legendre.leggauss(n)
        order = np.argsort(x)
        np.testing.assert_allclose(x[order], xr, atol=1e-14)
        np.testing.assert_allclose(w[order], wr, rtol=1e-11)


def test_psf_base_integral_is_planar_king_integral(orc):
    from scipy.integrate import quad
    p = np.array([1.3, 0.2, 0.02, 0.06, 2.1, 3.3])
    for r in (0.5, 3.0, 90.0):
        num, _ = quad(lambda t: orc.evaluate_king(t, 1.0, p) * 2 * np.pi * t,
                      0.0, np.deg2rad(r), limit=400, epsabs=0, epsrel=1e-12)
        assert orc.psf_base_integral(r, 1.0, p) == pytest.approx(num, rel=1e-9)


def test_normalize_rpsf_unit_integral(orc):
    # After normalisation the King integral is 1: by 64-point GL on the sphere
    # below 120 MeV, by the analytic planar form above (irf.normalize_rpsf.cpp).
    from scipy.integrate import quad
    par = np.zeros((1, 2, 6))
    par[0, :] = [1.0, 0.1, 0.05, 0.2, 1.8, 2.5]
    loge = np.array([1.5, 3.0])          # 31.6 MeV and 1 GeV
    nrm = orc.normalize_rpsf(par, loge, np.array([1.0, 0.0, 0.0], np.float32))
    sph, _ = quad(lambda v: orc.evaluate_king(np.deg2rad(v), 1.0, nrm[0, 0])
                  * np.sin(np.deg2rad(v)) * 2 * np.pi * np.deg2rad(1.0), 0, 90, limit=400)
    assert sph == pytest.approx(1.0, rel=1e-6)
    assert orc.psf_base_integral(90.0, 1.0, nrm[0, 1]) == pytest.approx(1.0, rel=1e-12)


def test_fast_separation_lower_index_brackets(orc):
    sep = orc.separations()
    for r in (1.0, 2.5, 7.3, 33.0, 69.9):
        d = orc.fast_separation_lower_index(r)
        assert sep[d] <= r * (1 + 1e-12) and r < sep[d + 1] * (1 + 1e-12)
    # as coded, radii below 1 degree map to index 1
    assert orc.fast_separation_lower_index(0.5) == 1


# ------------------------------------------------------------------ cubature
def _small_image(orc, proj="CAR", n=12, cdelt=0.05):
    return orc.make_wcs((n, n), (n / 2 + 0.5, n / 2 + 0.5), (180.0, 0.0),
                        (-cdelt, cdelt), 0.0, proj)


def test_cubature_of_constant_lut_is_constant(orc):
    w = _small_image(orc)
    rule = np.array([[0.3, -0.2, 1.5], [-0.7, 0.4, 2.5]])
    pts = orc.square_ptswts(12, 12, w, rule)
    lut = np.full((1, 401, 3), 2.5)
    src = np.array([[180.0, 0.0]])
    m = orc.cubature_convolve_map_psf(12, 12, src, lut, w, pts)
    np.testing.assert_allclose(m, 2.5, rtol=1e-13)


def test_logsep_equals_searchsep_inside_grid(orc):
    # Both are linear interpolation on the same node grid; the logsep index
    # and fraction formulas must land on the same bracket.
    w = _small_image(orc, n=16, cdelt=0.3)
    rng = np.random.default_rng(3)
    lut = rng.uniform(0.5, 2.0, (1, 401, 4))
    x, wt = orc.legendre_rw(4)
    rule = np.array([[a, b, wa * wb] for a, wa in zip(x, wt) for b, wb in zip(x, wt)])
    pts = orc.square_ptswts(16, 16, w, rule)
    src = np.array([[179.0, 1.3]])  # off-grid: no separation below 1e-4 deg
    a = orc.cubature_convolve_map_psf(16, 16, src, lut, w, pts)
    b = orc.cubature_convolve_map_psf(16, 16, src, lut, w, pts, searchsep=True)
    np.testing.assert_allclose(a, b, rtol=1e-11)


def test_brute_force_pixel_integral(orc):
    # An 8x8 Gauss-Legendre pixel cubature of a smooth LUT agrees with a
    # 48x48 midpoint brute force on the same pixels.
    w = _small_image(orc, n=6, cdelt=0.1)
    sep = orc.separations()
    lut = (1.0 / (1.0 + (sep / 0.3) ** 2))[None, :, None] * np.ones((1, 1, 2))
    x, wt = orc.legendre_rw(8)
    rule = np.array([[a, b, wa * wb] for a, wa in zip(x, wt) for b, wb in zip(x, wt)])
    src = np.array([[180.02, 0.13]])
    gl = orc.cubature_convolve_map_psf(6, 6, src, lut, w, orc.square_ptswts(6, 6, w, rule))
    bf = orc.cubature_convolve_map_psf(6, 6, src, lut, w, orc.naive_ptswts(6, 6, w, 48),
                                       searchsep=True)
    np.testing.assert_allclose(gl, bf, rtol=2e-3)


def test_point_src_model_map_pipeline_identities(orc):
    # With a constant LUT the raw cubature is the constant, so the final map is
    # c * Omega * exposure * CF, and CF = min(1, integral) / sum_in_radius.
    n = 20
    w = _small_image(orc, n=n, cdelt=0.1)
    sep = orc.separations()
    upsf = np.ones((1, 401, 2))
    parint, _ = orc.partial_total_integral(upsf)
    expo = np.array([[3.0, 5.0]])
    src = np.array([[180.0, 0.0]])
    rule = np.array([[0.0, 0.0, 4.0]])
    out, det = orc.point_src_model_map_wcs(n, n, w, src, expo, upsf, parint, rule,
                                           details=True)
    om = orc.solid_angle(n, n, w)
    rad, fov = orc.psf_boundary_radius(n, n, src, w)
    assert fov[0] == 1 and rad[0] == pytest.approx(1.0, rel=1e-6)
    d = orc.fast_separation_lower_index(rad[0])
    # pixel-centre separations within the radius
    cen = np.array([[orc.pix2dir(w, h + 1.0, ww + 1.0) for ww in range(n)] for h in range(n)])
    sd = orc.sph2dir(180.0, 0.0)
    rads = 2 * np.arcsin(0.5 * np.linalg.norm(cen - sd, axis=2))
    inside = rads <= np.deg2rad(rad[0])
    t1, t2 = np.deg2rad(sep[d]), np.deg2rad(sep[d + 1])
    seg = np.pi * (np.sin(t1) + np.sin(t2)) * (t2 - t1)
    integ = min(1.0, parint[0, d, 0] + seg)
    cf = integ / np.sum(om[inside])
    np.testing.assert_allclose(det["cf"][0], cf, rtol=1e-12)
    np.testing.assert_allclose(out[0, :, :, 1], om * 5.0 * cf, rtol=1e-12)
