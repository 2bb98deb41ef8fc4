"""Independent checks of the oracle's diffuse-source model map
(model_map/model_map.diffuse_src_model_map_wcs.cxx:8-97 and the steps it
names), each against mathematics the test computes itself:

* the J2000 galactic frame against published positions (the galactic
  centre, the galactic and celestial poles) and round trips;
* the crop and resampled grid: every model-map pixel centre lies inside
  the crop, the grid WCS places its pixel u at the cube's a + (u - 1) step,
  and the resample (math/trilerp.hpp:9-56) reproduces a trilinear function
  of (plane, axis 1, axis 2) exactly, and across the longitude seam equals
  the resample of the same sky from a cube rolled by half a turn;
* the kernel (psf/psf.make_kernel.cpp:8-90) against a table linear in
  separation, for which the node blend is exact, so the normalised kernel
  is (a + b s) / sum with s the separation the test computes from the
  pinned WCS; swapped blend weights, an off-centre kernel or transposed
  axes miss by orders of magnitude;
* the direct convolution against scipy.signal.convolve2d;
* the sampling against a grid linear in its pixel coordinates (bilinear
  interpolation exact), times the pinned solid angle and the exposure.
"""
import numpy as np
import pytest
from scipy.signal import convolve2d

from conftest import max_rel


def _sep():
    s = np.zeros(401)
    s[1:] = 1e-4 * np.exp(np.arange(400) * np.log(7e5) / 399.0)
    return s


def _cube_wcs(orc, n1=360, n2=181, step=1.0, crval1=0.0):
    """All-sky galactic CAR cube, longitude axis reversed as in the galactic
    diffuse MapCube."""
    return orc.make_wcs((n1, n2), ((n1 + 1) / 2.0, (n2 + 1) / 2.0), (crval1, 0.0),
                        (-step, step), 0.0, "CAR")


def _roi(orc, crval, n=(40, 36), step=0.1, proj="CAR", crota2=0.0):
    return orc.make_wcs(n, ((n[0] + 1) / 2.0, (n[1] + 1) / 2.0), crval, (-step, step),
                        crota2, proj)


# ------------------------------------------------------------ frames
def test_galactic_frame_known_positions(orc):
    # the galactic centre, (l, b) = (0, 0), at RA 17h45m37.224s, Dec -28d56m10.23s (J2000)
    l, b = orc.cel2gal(266.40510, -28.936175)
    assert min(l, 360.0 - l) < 2e-4 and abs(b) < 2e-4
    # the north galactic pole, and the celestial pole at l = 122.93192, b = 27.12825
    assert abs(orc.cel2gal(192.85948, 27.12825)[1] - 90.0) < 1e-9
    l, b = orc.cel2gal(37.0, 90.0)
    assert abs(l - 122.93192) < 1e-9 and abs(b - 27.12825) < 1e-9
    assert abs(orc.gal2cel(0.0, 90.0)[1] - 27.12825) < 1e-9
    rng = np.random.default_rng(1)
    for ra, dec in zip(rng.uniform(0, 360, 50), np.degrees(np.arcsin(rng.uniform(-0.99, 0.99, 50)))):
        l, b = orc.cel2gal(ra, dec)
        r2, d2 = orc.gal2cel(l, b)
        assert abs(d2 - dec) < 1e-10 and abs((r2 - ra + 180.0) % 360.0 - 180.0) < 1e-9


# ------------------------------------------------------------ crop, grid, resample
@pytest.mark.parametrize("case", ["same", "cel_over_gal", "seam", "rotated_ait"])
def test_geometry_covers_the_model_map(orc, case):
    cube = _cube_wcs(orc)
    roi_gal = 1
    if case == "same":
        roi = _roi(orc, (30.0, 10.0))
    elif case == "cel_over_gal":
        roi, roi_gal = _roi(orc, (193.98, -5.82), n=(50, 44), step=0.2), 0
    elif case == "seam":
        roi = _roi(orc, (180.0, -20.0), n=(60, 40), step=0.25)
    else:
        roi, roi_gal = _roi(orc, (83.6, 22.0), proj="AIT", crota2=17.0), 0
    g = orc.diffuse_geometry(roi, roi_gal, cube, 1, 0.25)
    assert g.wrap == 1 and g.frame == (0 if roi_gal else 1)
    for h in range(1, roi.naxis1 + 1, 3):
        for w in range(1, roi.naxis2 + 1, 3):
            lon, lat = orc.pix2sph(roi, h, w)
            if not roi_gal:
                lon, lat = orc.cel2gal(lon, lat)
            q1, q2 = orc.sph2pix(cube, lon, lat)
            q1 += 360.0 * np.round((g.a1 + 0.5 * g.n1 - q1) / 360.0)
            assert g.a1 <= q1 <= g.a1 + g.n1 - 1 and g.a2 <= q2 <= g.a2 + g.n2 - 1
    # grid pixel u is cube pixel a + (u - 1) step
    for u1, u2 in ((1.0, 1.0), (g.m1, g.m2), (0.37 * g.m1, 0.81 * g.m2)):
        a = orc.pix2dir(g.grid, u1, u2)
        b = orc.pix2dir(cube, g.a1 + (u1 - 1) * g.step1, g.a2 + (u2 - 1) * g.step2)
        assert np.linalg.norm(a - b) < 1e-12
    # the kernel centre is the model map's reference direction
    lon, lat = roi.crval1, roi.crval2
    if not roi_gal:
        lon, lat = orc.cel2gal(lon, lat)
    assert np.linalg.norm(orc.pix2dir(g.grid, g.kc1, g.kc2) - orc.sph2dir(lon, lat)) < 1e-12
    # transform lengths: the smallest 2^a 3^b 5^c >= m + c
    def smooth(n):
        for f in (2, 3, 5):
            while n % f == 0:
                n //= f
        return n == 1
    c = (g.nk - 1) // 2
    for n, m in ((g.fft_p, g.m1), (g.fft_q, g.m2)):
        assert n >= m + c and smooth(n) and not any(smooth(k) for k in range(m + c, n))
    assert g.nk % 2 == 1 and g.nk in (min(g.m1, g.m2), min(g.m1, g.m2) - 1)


def test_resample_is_exact_on_trilinear_cubes(orc):
    N1, N2, Ec = 200, 120, 5
    cube = orc.make_wcs((N1, N2), (80.0, 50.0), (30.0, 5.0), (-0.5, 0.5), 0.0, "CAR")
    roi = _roi(orc, (30.0, 5.0), n=(30, 26), step=0.2)
    g = orc.diffuse_geometry(roi, 1, cube, 1, 0.2)
    assert g.wrap == 0 and g.m1 > g.n1
    k, y, x = np.meshgrid(np.arange(Ec), np.arange(1, N2 + 1), np.arange(1, N1 + 1),
                          indexing="ij")
    c = [0.7, 0.1, 0.02, -0.03, 0.004, -0.002, 0.001, 3e-4]

    def f(k, x, y):
        return (c[0] + c[1] * k + c[2] * x + c[3] * y + c[4] * k * x + c[5] * x * y +
                c[6] * k * y + c[7] * k * x * y)
    data = f(k, x, y).astype(np.float32)
    E = 9
    out = orc.diffuse_resample(data, cube, g, E)
    kk = np.arange(E) * (Ec - 1) / (E - 1)
    xx = g.a1 + np.arange(g.m1) * (g.n1 - 1) / (g.m1 - 1)
    yy = g.a2 + np.arange(g.m2) * (g.n2 - 1) / (g.m2 - 1)
    K, X, Y = np.meshgrid(kk, xx, yy, indexing="ij")
    # the cube holds float32 values of f; trilinear blending of those is f
    # up to their rounding
    assert max_rel(out, f(K, X, Y)) < 1e-6


def test_resample_across_the_seam_equals_a_rolled_cube(orc):
    rng = np.random.default_rng(3)
    N1, N2, Ec = 360, 181, 3
    data = rng.uniform(0.5, 1.5, (Ec, N2, N1)).astype(np.float32)
    cube = _cube_wcs(orc, N1, N2)
    # the same sky with longitude 180 at the centre column
    rolled = np.roll(data, N1 // 2, axis=2)
    cube2 = _cube_wcs(orc, N1, N2, crval1=180.0)
    roi = _roi(orc, (180.0, 12.0), n=(44, 30), step=0.3)
    g1 = orc.diffuse_geometry(roi, 1, cube, 1, 0.3)
    g2 = orc.diffuse_geometry(roi, 1, cube2, 1, 0.3)
    assert g1.a1 < 1 or g1.a1 + g1.n1 - 1 > N1          # the crop crosses the seam
    assert 1 <= g2.a1 and g2.a1 + g2.n1 - 1 <= N1       # and the rolled one does not
    assert (g1.n1, g1.n2, g1.m1, g1.m2) == (g2.n1, g2.n2, g2.m1, g2.m2)
    r1 = orc.diffuse_resample(data, cube, g1, 4)
    r2 = orc.diffuse_resample(rolled, cube2, g2, 4)
    assert np.array_equal(r1, r2)


# ------------------------------------------------------------ kernel
@pytest.mark.parametrize("proj", ["CAR", "AIT"])
def test_kernel_linear_table(orc, proj):
    grid = orc.make_wcs((61, 61), (30.2, 33.7), (40.0, 50.0), (-0.07, 0.05), 0.0, proj)
    nk, kc1, kc2 = 15, 30.6, 28.3
    sep = _sep()
    a, b = np.array([3.0, 1.0]), np.array([-0.5, 2.0])
    upsf = a[None, :] + b[None, :] * sep[:, None]
    K = orc.psf_make_kernel(nk, grid, kc1, kc2, upsf)
    c = (nk - 1) // 2
    d0 = orc.pix2dir(grid, kc1, kc2)
    s = np.zeros((nk, nk))
    for h in range(nk):
        for w in range(nk):
            d = orc.pix2dir(grid, kc1 + h - c, kc2 + w - c)
            s[h, w] = 2.0 * np.degrees(np.arcsin(0.5 * np.linalg.norm(d - d0)))
    for e in range(2):
        ref = a[e] + b[e] * s
        assert max_rel(K[e], ref / ref.sum()) < 1e-12
    assert abs(K.sum(axis=(1, 2)) - 1).max() < 1e-14


def test_direct_convolution_against_scipy(orc):
    rng = np.random.default_rng(4)
    img = rng.uniform(-1, 1, (3, 23, 17))
    ker = rng.uniform(0, 1, (3, 9, 9))
    out = orc.convolve_same(img, ker)
    for e in range(3):
        full = convolve2d(img[e], ker[e], mode="full")
        assert np.allclose(out[e], full[4:4 + 23, 4:4 + 17], rtol=0, atol=1e-13)


# ------------------------------------------------------------ sampling
def test_sampling_linear_grid(orc):
    cube = _cube_wcs(orc)
    roi = _roi(orc, (193.98, -5.82), n=(32, 28), step=0.2)
    g = orc.diffuse_geometry(roi, 0, cube, 1, 0.25)
    E = 2
    al, be, ga = np.array([5.0, 2.0]), np.array([0.01, -0.02]), np.array([0.03, 0.005])
    I, J = np.meshgrid(np.arange(g.m1), np.arange(g.m2), indexing="ij")
    conv = al[:, None, None] + be[:, None, None] * I + ga[:, None, None] * J
    om = orc.solid_angle(roi.naxis1, roi.naxis2, roi)
    expo = np.array([2.0, 3.0])
    out = orc.diffuse_sample(conv, g, roi, om, expo)
    ref = np.zeros_like(out)
    for h in range(roi.naxis1):
        for w in range(roi.naxis2):
            lon, lat = orc.cel2gal(*orc.pix2sph(roi, h + 1, w + 1))
            v1, v2 = orc.sph2pix(g.grid, lon, lat)
            v1 += g.period * np.round((0.5 * (g.m1 + 1) - v1) / g.period)
            ref[:, h, w] = (al + be * (v1 - 1) + ga * (v2 - 1)) * om[h, w] * expo
    assert max_rel(out, ref) < 1e-12


def test_composite_is_the_steps(orc):
    rng = np.random.default_rng(6)
    cube = _cube_wcs(orc, 180, 91, 2.0)
    data = rng.uniform(0.5, 1.5, (4, 91, 180)).astype(np.float32)
    roi = _roi(orc, (150.0, 30.0), n=(20, 18), step=0.5)
    sep = _sep()
    E = 3
    upsf = np.exp(-sep[:, None] / np.array([0.5, 1.0, 2.0])[None, :])
    expo = np.array([1e9, 2e9, 3e9])
    out = orc.diffuse_src_model_map_wcs(roi, 0, cube, 1, data, upsf, expo, 0.5)
    g = orc.diffuse_geometry(roi, 0, cube, 1, 0.5)
    img = orc.diffuse_resample(data, cube, g, E)
    K = orc.psf_make_kernel(g.nk, g.grid, g.kc1, g.kc2, upsf)
    conv = orc.convolve_same(img, K)
    ref = orc.diffuse_sample(conv, g, roi, orc.solid_angle(20, 18, roi), expo)
    assert np.array_equal(out, ref) and (out > 0).all()


# ------------------------------------------------------------ the library's plan
@pytest.mark.parametrize("case", ["same", "cel_over_gal", "seam", "rotated_ait", "pole"])
def test_library_geometry_equals_oracle(orc, case):
    """xts_diffuse_geometry (host code of the library) against the oracle's."""
    import xtsrcmaps_amd as xa
    cube = xa.Wcs.make((360, 181), (180.5, 91.0), (0.0, 0.0), (-1.0, 1.0), 0.0, "CAR",
                       galactic=True)
    roi = {"same": xa.Wcs.make((60, 50), (30.5, 25.5), (30.0, 10.0), (-0.25, 0.25), 0.0, "CAR",
                               galactic=True),
           "cel_over_gal": xa.Wcs.make((56, 64), (28.5, 32.5), (193.98, -5.82), (-0.2, 0.2)),
           "seam": xa.Wcs.make((70, 40), (35.5, 20.5), (180.0, -20.0), (-0.3, 0.3), 0.0, "CAR",
                               galactic=True),
           "rotated_ait": xa.Wcs.make((48, 52), (24.0, 27.0), (83.6, 22.0), (-0.25, 0.25), 17.0,
                                      "AIT"),
           "pole": xa.Wcs.make((40, 40), (20.5, 20.5), (0.0, 88.0), (-0.25, 0.25), 0.0, "AIT",
                               galactic=True)}[case]
    g = xa.diffuse_geometry(roi, cube, 0.4, 0.7)

    def ow(w):
        return orc.make_wcs((w.naxis1, w.naxis2), (w.crpix1, w.crpix2), (w.crval1, w.crval2),
                            (w.cdelt1, w.cdelt2), w.crota2, "CAR" if w.proj == 0 else "AIT")
    o = orc.diffuse_geometry(ow(roi), roi.is_galactic, ow(cube), 1, 0.4, 0.7)
    for a, b in (("crop_a1", "a1"), ("crop_a2", "a2"), ("crop_n1", "n1"), ("crop_n2", "n2"),
                 ("m1", "m1"), ("m2", "m2"), ("wrap", "wrap"), ("nk", "nk"),
                 ("fft_p", "fft_p"), ("fft_q", "fft_q")):
        assert g[a] == getattr(o, b), a
    for a in ("step1", "step2", "kc1", "kc2"):
        assert abs(g[a] - getattr(o, a)) < 1e-9, a
    for f in ("crpix1", "crpix2", "cdelt1", "cdelt2"):
        assert abs(getattr(g["grid"], f) - getattr(o.grid, f)) < 1e-9, f
    if case == "pole":
        assert g["crop_n1"] == 360 and g["crop_a2"] + g["crop_n2"] - 1 == 181


def test_library_rejects_bad_diffuse_arguments():
    import xtsrcmaps_amd as xa
    cube = xa.Wcs.make((360, 181), (180.5, 91.0), (0.0, 0.0), (-1.0, 1.0), 0.0, "CAR",
                       galactic=True)
    roi = xa.Wcs.make((60, 50), (30.5, 25.5), (30.0, 10.0), (-0.25, 0.25), 0.0, "CAR",
                      galactic=True)
    for target, margin in ((0.0, 0.0), (-1.0, 0.0), (0.5, -1.0), (0.001, 0.0)):
        with pytest.raises(xa.XtsError):
            xa.diffuse_geometry(roi, cube, target, margin)
    # a map outside a cube that does not wrap
    part = xa.Wcs.make((100, 60), (50.5, 30.5), (20.0, 0.0), (-1.0, 1.0), 0.0, "CAR",
                       galactic=True)
    far = xa.Wcs.make((20, 20), (10.5, 10.5), (200.0, 0.0), (-0.25, 0.25), 0.0, "CAR",
                      galactic=True)
    with pytest.raises(xa.XtsError):
        xa.diffuse_geometry(far, part, 0.5)
