"""GPU parity of the diffuse-source model map
(model_map/model_map.diffuse_src_model_map_wcs.cxx:8-97) and its two
reference-shaped steps, through the C ABI:

* FFT::convolve_map_psf (xts_fft_convolve_map_psf) against scipy's FFT
  convolution (an independent transform) at sizes whose transforms use
  every radix (2, 3, 4, 5), within 1e-12 of the largest output value;
* Psf::make_kernel (xts_psf_make_kernel) against the FP64 oracle, 1e-12;
* the whole map against the FP64 oracle (direct convolution) on small
  cubes: same frame, a celestial map over a galactic cube, a map across the
  longitude seam, a rotated AIT map, a map holding the galactic pole, a map
  partly outside a cube that does not wrap, an all-sky AIT cube, a cube of
  one plane at one energy; within 1e-10 of each value
  and 1e-11 of the largest (the model map around the pole, where a CAR
  cube's longitude stretches without bound, takes FP64 rounding of the
  pixel positions to 1e-12), zeros where the oracle has zeros, bitwise
  repeatable, and the FP32 output within 1e-6 of the FP64;
* the benchmark's shape (400 x 400 at 0.05 deg, 37 energies, over a
  2880 x 1441 cube with 28 planes) on three energies against the oracle's
  resample, kernel and sampling around scipy's FFT convolution, 1e-9.

All arithmetic is FP64, so the bars are set by FFT rounding, not by the
method."""
import sys

import numpy as np
import pytest
import torch
from scipy.signal import fftconvolve

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


def _ow(orc, w):
    """The oracle's WCS of an xts_wcs."""
    return orc.make_wcs((w.naxis1, w.naxis2), (w.crpix1, w.crpix2), (w.crval1, w.crval2),
                        (w.cdelt1, w.cdelt2), w.crota2, "CAR" if w.proj == 0 else "AIT")


def _sep():
    s = np.zeros(401)
    s[1:] = 1e-4 * np.exp(np.arange(400) * np.log(7e5) / 399.0)
    return s


@pytest.mark.parametrize("shape", [(2, 17, 23, 5, 7), (3, 64, 48, 33, 33), (1, 100, 7, 1, 1),
                                   (2, 250, 311, 101, 61), (1, 1000, 1200, 999, 3)])
def test_gpu_fft_convolve_map_psf(ctx, shape):
    C, H, W, KH, KW = shape
    _log(f"test_gpu_fft_convolve_map_psf: {shape}")
    rng = np.random.default_rng(sum(shape))
    img = rng.uniform(-1, 1, (C, H, W))
    ker = rng.uniform(0, 1, (C, KH, KW))
    out = ctx.fft_convolve_map_psf(img, ker)
    assert out.shape == (C, H + KH - 1, W + KW - 1)
    for c in range(C):
        ref = fftconvolve(img[c], ker[c], mode="full")
        err = np.abs(out[c] - ref).max() / np.abs(ref).max()
        _log(f"  plane {c}: max err / max {err:.2e}")
        assert err < 1e-12


@pytest.mark.parametrize("proj,nk", [("CAR", 15), ("AIT", 101)])
def test_gpu_psf_make_kernel(ctx, orc, proj, nk):
    import xtsrcmaps_amd as xa
    _log(f"test_gpu_psf_make_kernel: {proj} {nk}")
    grid = xa.Wcs.make((200, 180), (100.3, 91.7), (40.0, 50.0), (-0.07, 0.05), 12.0, proj)
    sep = _sep()
    upsf = np.exp(-sep[:, None] / np.array([0.05, 0.3, 2.0])[None, :]) + 1e-3
    k = ctx.psf_make_kernel(nk, grid, 97.6, 88.2, upsf)
    ref = orc.psf_make_kernel(nk, _ow(orc, grid), 97.6, 88.2, upsf)
    r = np.abs(k - ref).max() / np.abs(ref).max()
    _log(f"  max err / max {r:.2e}")
    assert r < 1e-12 and abs(k.sum(axis=(1, 2)) - 1).max() < 1e-13


def _case(xa, name):
    """(roi, roi_gal, cube, cube planes, target, margin)."""
    cube = xa.Wcs.make((360, 181), (180.5, 91.0), (0.0, 0.0), (-1.0, 1.0), 0.0, "CAR",
                       galactic=True)
    if name == "same":
        roi = xa.Wcs.make((60, 50), (30.5, 25.5), (30.0, 10.0), (-0.25, 0.25), 0.0, "CAR",
                          galactic=True)
        return roi, cube, 0.5, 0.0
    if name == "cel_over_gal":
        roi = xa.Wcs.make((56, 64), (28.5, 32.5), (193.98, -5.82), (-0.2, 0.2), 0.0, "CAR")
        return roi, cube, 0.4, 1.5
    if name == "seam":
        roi = xa.Wcs.make((70, 40), (35.5, 20.5), (180.0, -20.0), (-0.3, 0.3), 0.0, "CAR",
                          galactic=True)
        return roi, cube, 0.6, 0.0
    if name == "rotated_ait":
        roi = xa.Wcs.make((48, 52), (24.0, 27.0), (83.6, 22.0), (-0.25, 0.25), 17.0, "AIT")
        return roi, cube, 0.5, 0.0
    if name == "pole":
        roi = xa.Wcs.make((40, 40), (20.5, 20.5), (0.0, 88.0), (-0.25, 0.25), 0.0, "AIT",
                          galactic=True)
        return roi, cube, 1.0, 0.0
    if name == "ait_cube":
        # an all-sky AIT cube (no longitude wrap in its pixels) under a
        # celestial map
        ait = xa.Wcs.make((360, 180), (180.5, 90.5), (0.0, 0.0), (-1.0, 1.0), 0.0, "AIT",
                          galactic=True)
        roi = xa.Wcs.make((44, 40), (22.5, 20.5), (266.4, -28.9), (-0.25, 0.25), 0.0, "CAR")
        return roi, ait, 0.5, 0.0
    if name == "one_plane":
        roi = xa.Wcs.make((30, 26), (15.5, 13.5), (100.0, -40.0), (-0.5, 0.5), 0.0, "CAR",
                          galactic=True)
        return roi, cube, 1.0, 0.0
    # a cube covering 100 x 60 deg that does not wrap, the map across its edge
    part = xa.Wcs.make((100, 60), (50.5, 30.5), (20.0, 0.0), (-1.0, 1.0), 0.0, "CAR",
                       galactic=True)
    roi = xa.Wcs.make((60, 40), (30.5, 20.5), (68.0, 5.0), (-0.25, 0.25), 0.0, "CAR",
                      galactic=True)
    return roi, part, 0.5, 0.0


@pytest.mark.parametrize("name", ["same", "cel_over_gal", "seam", "rotated_ait", "pole",
                                  "partial", "ait_cube", "one_plane"])
def test_gpu_diffuse_map_vs_oracle(ctx, orc, name):
    import xtsrcmaps_amd as xa
    roi, cube, target, margin = _case(xa, name)
    g = xa.diffuse_geometry(roi, cube, target, margin)
    _log(f"test_gpu_diffuse_map_vs_oracle[{name}]: grid {g['m1']}x{g['m2']}, kernel {g['nk']}, "
         f"transforms {g['fft_p']}x{g['fft_q']}")
    rng = np.random.default_rng(11)
    planes, E = (1, 1) if name == "one_plane" else (5, 7)
    lat = cube.cdelt2 * (np.arange(1, cube.naxis2 + 1) - cube.crpix2)
    data = (np.exp(-np.abs(lat) / 10.0)[None, :, None] *
            rng.uniform(0.5, 1.5, (planes, cube.naxis2, cube.naxis1))).astype(np.float32)
    sep = _sep()
    upsf = (1.0 + (sep[:, None] / np.linspace(0.3, 3.0, E)[None, :]) ** 2 / 5.0) ** -2.5
    expo = np.linspace(1e10, 3e10, E)
    if name == "ait_cube":
        # AIT pixels outside the ellipse hold no sky
        y = np.arange(1, cube.naxis2 + 1)[:, None] - cube.crpix2
        x = np.arange(1, cube.naxis1 + 1)[None, :] - cube.crpix1
        data *= ((x / 180.0) ** 2 + (y / 90.0) ** 2 <= 1.0)[None].astype(np.float32)
    got = ctx.diffuse_src_model_map_wcs(roi, cube, data, upsf, expo, target, margin)
    ref = orc.diffuse_src_model_map_wcs(_ow(orc, roi), roi.is_galactic, _ow(orc, cube),
                                        cube.is_galactic, data, upsf, expo, target, margin)
    nz = ref != 0
    r = np.abs(got[nz] - ref[nz]) / np.abs(ref[nz])
    rmax = np.abs(got - ref).max() / np.abs(ref).max()
    _log(f"  {nz.sum()} of {nz.size} nonzero; max rel {r.max():.2e}, max err / max {rmax:.2e}")
    assert nz.any() and np.all(got[~nz] == 0)
    assert r.max() < 1e-10 and rmax < 1e-11
    if name == "partial":
        assert (~nz).any()
    # device entry point: bitwise repeatable, FP32 output within 1e-6
    dev = torch.device("cuda")
    t_cube = torch.as_tensor(data, device=dev)
    t_u = torch.as_tensor(upsf, device=dev)
    t_e = torch.as_tensor(expo, device=dev)
    o64 = torch.empty((E, roi.naxis1, roi.naxis2), dtype=torch.float64, device=dev)
    o32 = torch.empty((E, roi.naxis1, roi.naxis2), dtype=torch.float32, device=dev)
    torch.cuda.synchronize()
    ctx.diffuse_src_model_map_wcs_dev(roi, cube, t_cube, t_u, t_e, target, o64, margin)
    ctx.diffuse_src_model_map_wcs_dev(roi, cube, t_cube, t_u, t_e, target, o32, margin,
                                      flags=xa.OUT_F32)
    ctx.synchronize()
    a = o64.cpu().numpy()
    assert np.array_equal(a, got)
    b = o32.double().cpu().numpy()
    assert (np.abs(b[nz] - a[nz]) / np.abs(a[nz])).max() < 1e-6


def test_gpu_diffuse_benchmark_shape(ctx, orc):
    import xtsrcmaps_amd as xa
    from xtsrcmaps_amd import synthetic as sy
    cfg = sy.DIFFUSE_CONFIGS["diffuse"]
    roi, cube_wcs, cube, upsf, expo = sy.diffuse_inputs(cfg)
    E = cfg.n_energy
    g = xa.diffuse_geometry(roi, cube_wcs, cfg.target)
    _log(f"test_gpu_diffuse_benchmark_shape: grid {g['m1']}x{g['m2']}, kernel {g['nk']}")
    dev = torch.device("cuda")
    out = torch.empty((E, roi.naxis1, roi.naxis2), dtype=torch.float64, device=dev)
    t_cube = torch.as_tensor(cube, device=dev)
    t_u = torch.as_tensor(upsf, device=dev)
    t_e = torch.as_tensor(expo, device=dev)
    torch.cuda.synchronize()
    ctx.diffuse_src_model_map_wcs_dev(roi, cube_wcs, t_cube, t_u, t_e, cfg.target, out)
    ctx.synchronize()
    got = out.cpu().numpy()
    og = orc.diffuse_geometry(_ow(orc, roi), 0, _ow(orc, cube_wcs), 1, cfg.target)
    assert (og.m1, og.m2, og.nk) == (g["m1"], g["m2"], g["nk"])
    img = orc.diffuse_resample(cube, _ow(orc, cube_wcs), og, E)
    planes = [0, E // 2, E - 1]
    K = orc.psf_make_kernel(og.nk, og.grid, og.kc1, og.kc2, upsf[:, planes])
    c = (og.nk - 1) // 2
    conv = np.stack([fftconvolve(img[e], K[i], mode="full")[c:c + og.m1, c:c + og.m2]
                     for i, e in enumerate(planes)])
    om = orc.solid_angle(roi.naxis1, roi.naxis2, _ow(orc, roi))
    ref = orc.diffuse_sample(conv, og, _ow(orc, roi), om, expo[planes])
    r = np.abs(got[planes] - ref) / np.abs(ref)
    _log(f"  planes {planes}: max rel {r.max():.2e}")
    assert np.isfinite(got).all() and (got > 0).all() and r.max() < 1e-9


def test_gpu_diffuse_map_from_mapcube_file(ctx, tmp_path):
    """A MapCube FITS file (primary image, ENERGIES table) through
    xts_read_mapcube into the diffuse map: bit for bit the map of the same
    cube handed over as an array."""
    import xtsrcmaps_amd as xa
    from test_fits_mapcube import _mapcube
    _log("test_gpu_diffuse_map_from_mapcube_file")
    rng = np.random.default_rng(8)
    data = rng.uniform(0.5, 1.5, (3, 181, 360)).astype(np.float32)
    f = tmp_path / "cube.fits"
    _mapcube(f, data, energies=[1e2, 1e3, 1e4])
    w, cube, energies = xa.read_mapcube(f)
    assert w.is_galactic == 1 and energies is not None and np.array_equal(cube, data)
    roi = xa.Wcs.make((40, 36), (20.5, 18.5), (193.98, -5.82), (-0.25, 0.25), 0.0, "CAR")
    E = 4
    upsf = (1.0 + (_sep()[:, None] / np.linspace(0.5, 2.0, E)[None, :]) ** 2 / 5.0) ** -2.5
    expo = np.full(E, 1e10)
    a = ctx.diffuse_src_model_map_wcs(roi, w, cube, upsf, expo, 0.5)
    b = ctx.diffuse_src_model_map_wcs(roi, w, data, upsf, expo, 0.5)
    assert np.array_equal(a, b) and np.isfinite(a).all() and (a > 0).all()
