"""Independent checks of the three parts of the oracle that the reference's
fixtures do not cover: the solid angle's neighbour-centre quad
(math/solid_angle.hpp:42-79, Q10), the correction factor's radius index and
segment for radii below 1 degree (psf.fast_separation_lower_index.cxx:5-22,
model_map.map_correction_factor.cxx:24-50, Q5/Q6), and the logsep
interpolation outside [1e-4, 70] degrees (convolve_pixel_psf_logsep.hpp:
46-54, Q2).  Each expected value comes from mathematics the test computes
itself (spherical excess, an exact cumulative integral of a piecewise-linear
integrand, the segment's straight line), chosen so that a plausible mistake
in the oracle -- a corner quad, one triangle, a shifted axis, the wrong node
or segment, clamping instead of extrapolating -- fails the tolerance.  Also
the FermiTools-consistent exact-radius integral against FermiTools' own
values for the 3C279 ROI (tests/expected/ft_mm_upsf_radius_integ.bin)."""
import numpy as np
import pytest

from conftest import max_rel

D2R = np.pi / 180.0


def _sep():
    """psf/psf.hxx:24-31 (the table itself is pinned by separations_expected)."""
    s = np.zeros(401)
    s[1:] = 1e-4 * np.exp(np.arange(400) * np.log(7e5) / 399.0)
    return s


# ------------------------------------------------------------- solid angle
def _tri_excess(a, b, c):
    """Spherical excess of the geodesic triangle abc (Oosterom & Strackee 1983)."""
    num = abs(np.dot(a, np.cross(b, c)))
    den = 1.0 + np.dot(a, b) + np.dot(b, c) + np.dot(c, a)
    return 2.0 * np.arctan2(num, den)


def _quad_area(orc, w, p1, p2, dp=1.0):
    """Geodesic quad of the directions at FITS coordinates (p1, p2),
    (p1, p2+dp), (p1+dp, p2+dp), (p1+dp, p2), split along the second diagonal."""
    A = orc.pix2dir(w, p1, p2)
    B = orc.pix2dir(w, p1, p2 + dp)
    Cc = orc.pix2dir(w, p1 + dp, p2 + dp)
    D = orc.pix2dir(w, p1 + dp, p2)
    return _tri_excess(A, B, D) + _tri_excess(Cc, B, D), (A, B, Cc, D)


def _flat_tri2(a, b, d):
    ab, ad = a - b, a - d
    ang = lambda x, y: 2 * np.arcsin(0.5 * np.linalg.norm(x - y))
    c = np.cross(ab / np.linalg.norm(ab), ad / np.linalg.norm(ad))
    return ang(a, b) * ang(a, d) * np.linalg.norm(c)


WCS_CASES = {
    # rotated CAR 40 degrees off its native equator: the pixel area (cos of
    # the native latitude) varies along both image axes
    "car_dec60_rot30": dict(naxis=(24, 18), crpix=(12.5, -800.0), crval=(40.0, 20.0),
                            cdelt=(-0.05, 0.05), crota2=30.0, proj="CAR"),
    # AIT 30-35 degrees from its reference point (sheared pixels), as the
    # 3C279 ROI lies off its own reference point
    "ait_south": dict(naxis=(20, 26), crpix=(-600.0, 700.0), crval=(195.0, -5.8),
                      cdelt=(-0.05, 0.05), crota2=0.0, proj="AIT"),
}


@pytest.mark.parametrize("case", sorted(WCS_CASES))
def test_solid_angle_is_the_neighbour_centre_quad_area(orc, case):
    c = WCS_CASES[case]
    w = orc.make_wcs(c["naxis"], c["crpix"], c["crval"], c["cdelt"], c["crota2"], c["proj"])
    nh, nw = c["naxis"]
    om = orc.solid_angle(nh, nw, w)
    exact = np.zeros_like(om)
    corner = np.zeros_like(om)
    one_tri = np.zeros_like(om)
    shifted = np.zeros_like(om)
    for h in range(nh):
        for k in range(nw):
            exact[h, k], (A, B, Cc, D) = _quad_area(orc, w, 1.0 + h, 1.0 + k)
            corner[h, k], _ = _quad_area(orc, w, 0.5 + h, 0.5 + k)
            shifted[h, k], _ = _quad_area(orc, w, 1.0 + h, 0.0 + k)
            one_tri[h, k] = _flat_tri2(A, B, D)
    # the flat two-triangle quad of arc-length sides equals the spherical
    # quad to O(pixel^2) ~ 1e-7 at 0.05 degrees
    assert max_rel(om, exact) < 1e-6
    if c["proj"] == "AIT":
        # Hammer-Aitoff is equal-area: every quad of neighbouring pixel
        # centres covers |cdelt1 cdelt2| of the plane, hence that solid angle
        area = abs(c["cdelt"][0] * c["cdelt"][1]) * D2R ** 2
        assert max_rel(om, np.full_like(om, area)) < 1e-6
        mistakes = (one_tri,)   # a shifted quad has the same area here
    else:
        mistakes = (corner, one_tri, shifted)
    # the tolerance separates the reference's quad from plausible mistakes
    for bad in mistakes:
        assert max_rel(bad, exact) > 2e-5


# ------------------------------------------------- correction-factor radius
def _table(E=3):
    sep = _sep()
    sig = np.array([0.02, 0.7, 9.0])[:E]
    upsf = np.exp(-np.outer(sep, 1.0 / sig)) * (1.0 + 0.3 * np.cos(np.outer(sep, [3.0, 1.0, 0.2])[:, :E]))
    th = sep * D2R
    y = 2 * np.pi * np.sin(th)[:, None] * upsf
    # exact cumulative integral of the piecewise-linear interpolant of y
    P = np.zeros_like(y)
    P[1:] = np.cumsum(0.5 * (y[1:] + y[:-1]) * np.diff(th)[:, None], axis=0)
    return sep, th, upsf, y, P


RADII = [2e-4, 0.0031, 0.05, 0.5, 0.999, 1.0, 1.5, 7.3, 33.3, 69.9]


def test_correction_factor_index_and_segment_as_coded(orc):
    sep, th, upsf, y, P = _table()
    S, E = len(RADII), upsf.shape[1]
    U, PI = np.repeat(upsf[None], S, 0), np.repeat(P[None], S, 0)
    rng = np.random.default_rng(5)
    invmi = rng.uniform(0.5, 2.0, (S, E))
    fov = np.ones(S, np.uint8)
    cf = orc.map_correction_factor(invmi, np.array(RADII), fov, U, PI)
    for s, R in enumerate(RADII):
        # as coded: below 1 degree the node index is 1 (Q5), else the node at or
        # below R; the whole next segment is added (Q6): the integral to node d+1
        d = 1 if R < 1.0 else int(np.searchsorted(sep, R, side="right") - 1)
        d = min(d, 399)
        expect = np.minimum(P[d + 1], 1.0) * invmi[s]
        np.testing.assert_allclose(cf[s], expect, rtol=1e-13)
    # outside the field of view the factor is 1
    cf0 = orc.map_correction_factor(invmi, np.array(RADII), np.zeros(S, np.uint8), U, PI)
    assert np.all(cf0 == 1.0)


def test_correction_factor_exact_radius_variant(orc):
    sep, th, upsf, y, P = _table()
    S, E = len(RADII), upsf.shape[1]
    U, PI = np.repeat(upsf[None], S, 0), np.repeat(P[None], S, 0)
    got = orc.psf_radius_integral(np.array(RADII), U, PI, exact=True)
    for s, R in enumerate(RADII):
        d = int(np.searchsorted(sep, R, side="right") - 1)
        tr = R * D2R
        yr = y[d] + (y[d + 1] - y[d]) * (tr - th[d]) / (th[d + 1] - th[d])
        expect = P[d] + 0.5 * (y[d] + yr) * (tr - th[d])
        np.testing.assert_allclose(got[s], expect, rtol=1e-12)
    # at a node both variants agree with the cumulative integral there
    node = np.array([sep[37], sep[250]])
    a = orc.psf_radius_integral(node, U[:2], PI[:2], exact=True)
    np.testing.assert_allclose(a, P[[37, 250]], rtol=1e-12)


def test_exact_radius_integral_matches_fermitools(orc, golden, chain):
    """FermiTools' PSF integral within each in-FOV source's radius for the
    3C279 ROI (47 sources x 38 energies) against the exact-radius integral
    of the oracle's mean PSF built from the reference's IRF and livetime
    fixtures.  FermiTools' radius differs from the reference's edge-point
    radius by up to 4.1e-4 relative, which bounds the agreement."""
    M, X, F = golden("mm_single_source"), golden("exposure_psf"), golden("fermitools_boundary")
    ch = chain
    w = orc.make_wcs(M["naxis"], M["crpix"], M["crval"], M["cdelt"], float(M["crota2"]),
                     str(M["proj"]))
    nh, nw = int(M["naxis"][0]), int(M["naxis"][1])
    rad, fov = orc.psf_boundary_radius(nh, nw, X["src_dirs"], w)
    inside = fov.astype(bool)
    ft = F["upsf_radius_integ"].reshape(47, 38)
    got = orc.psf_radius_integral(rad[inside], ch["upsf"][inside], ch["parint"][inside],
                                  exact=True)
    rel = np.abs(got - ft) / np.abs(ft)
    assert np.median(rel) < 1e-6
    assert rel.max() < 1e-3
    # as coded (the whole segment, index 1 below 1 degree) it is far off
    coded = orc.psf_radius_integral(rad[inside], ch["upsf"][inside], ch["parint"][inside])
    assert (np.abs(coded - ft) / np.abs(ft)).max() > 1e-2


# ------------------------------------------------ logsep outside the grid
def _point_pixels(seps_deg):
    """One single-point pixel per separation: the point at angle s from the
    source direction (1, 0, 0) along the equator, weight 1."""
    s = np.asarray(seps_deg) * D2R
    pts = np.zeros((1, len(s), 4))
    pts[0, :, 0], pts[0, :, 1], pts[0, :, 3] = np.cos(s), np.sin(s), 1.0
    return pts


def test_logsep_extrapolates_the_edge_segments(orc):
    sep = _sep()
    # a table with a kink at every node, in particular at sep[1] and sep[399]
    k = np.arange(401, dtype=float)
    L = np.stack([1.0 + 0.5 * k + 0.01 * k ** 2, 50.0 - 0.1 * k + 0.003 * k ** 1.5], axis=1)
    below = [0.0, 3e-5, 9.9e-5]
    above = [70.5, 85.0, 120.0]
    inside = [1.5e-4, 0.3, 69.0]
    ss = below + above + inside
    pts = _point_pixels(ss)
    w = orc.make_wcs((1, len(ss)), (1, 1), (0, 0), (1, 1))   # not read: points are given
    m = orc.cubature_convolve_map_psf(1, len(ss), np.array([[0.0, 0.0]]), L[None], w, pts)[0, 0]
    for j, s in enumerate(ss):
        if s < 1e-4:       # the line of segment [1, 2], extended below sep[1]
            exp = L[1] + (s - sep[1]) * (L[2] - L[1]) / (sep[2] - sep[1])
            wrong = L[0] + (s - sep[0]) * (L[1] - L[0]) / (sep[1] - sep[0])
        elif s > 70.0:     # the line of segment [399, 400], extended past 70 degrees
            exp = L[399] + (s - sep[399]) * (L[400] - L[399]) / (sep[400] - sep[399])
            wrong = L[400]
        else:
            i = int(np.searchsorted(sep, s, side="right"))
            exp = L[i - 1] + (s - sep[i - 1]) * (L[i] - L[i - 1]) / (sep[i] - sep[i - 1])
            wrong = None
        np.testing.assert_allclose(m[j], exp, rtol=1e-10, atol=1e-12)
        if wrong is not None:
            assert np.max(np.abs(wrong - exp) / np.abs(exp)) > 1e-4
