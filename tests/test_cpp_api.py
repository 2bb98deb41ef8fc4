"""The reference's C++ API (Fermi::Exposure / Psf / ModelMap / Irf /
SkyGeom / Tensor) as implemented by this library: the main.cxx pipeline and
the stage-by-stage forms run by tests/cpp/fermi_pipeline, compared with the
FP64 oracle on the same synthetic inputs.

Tolerances as in test_gpu_parity: FP64 stages 1e-10 relative, the FP32
cubature 1e-5 relative per map value."""
import os
import subprocess

import numpy as np
import pytest

from conftest import max_rel

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "cpp", "fermi_pipeline")


def test_driver_built_and_usage():
    """The driver links against the library; without arguments it prints its
    usage and exits 2 before touching a device."""
    if not os.path.exists(DRIVER):
        pytest.fail(f"{DRIVER} missing: run python -m xtsrcmaps_amd.build")
    r = subprocess.run([DRIVER], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "usage" in r.stderr


@pytest.mark.gpu
def test_cpp_pipeline_matches_oracle(orc, tmp_path):
    from pipelines import DEG7, oracle_stages, oracle_wcs, read_xtsa, write_xtsa, xtsa_inputs
    from xtsrcmaps_amd import synthetic
    cfg = synthetic.Config("cpp", 3, 64, 0.1, n_energy=9, rule="deg7", seed=11,
                           src_mode="uniform")
    inp = synthetic.make_inputs(cfg)
    # Obs::ExposureCubeData holds the livetime cube as float (obs_types.hxx),
    # so the oracle gets the same float-rounded cube
    inp.cube = inp.cube.astype(np.float32).astype(np.float64)
    inp.wcube = inp.wcube.astype(np.float32).astype(np.float64)
    write_xtsa(tmp_path / "in.xtsa", xtsa_inputs(inp))
    print("test_cpp_pipeline_matches_oracle: running the C++ driver", flush=True)
    r = subprocess.run([DRIVER, str(tmp_path / "in.xtsa"), str(tmp_path / "out.xtsa"),
                        str(tmp_path / "srcmaps.fits")],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    got = read_xtsa(tmp_path / "out.xtsa")
    # Fermi::fits::write_src_model of the same map: [E, W, H] cubes, FP64
    import fits_reader as fr
    hdus = fr.split_hdus((tmp_path / "srcmaps.fits").read_bytes())
    assert len(hdus) == 1 + got["map"].shape[0]
    for s, (h, d) in enumerate(hdus[1:]):
        assert fr.cards(h)["EXTNAME"] == f"src{s}"
        np.testing.assert_array_equal(fr.image(h, d), got["map"][s].transpose(2, 1, 0))
        assert fr.verify(h, d)["hdu_sum_is_minus_zero"]

    st = oracle_stages(orc, inp)
    assert max_rel(got["exposure"], st["expo"]) < 1e-12
    np.testing.assert_array_equal(got["ecb"], st["ecb"])
    assert max_rel(got["upsf"], st["upsf"]) < 1e-10
    assert max_rel(got["parint"][:, 1:], st["parint"][:, 1:]) < 1e-10
    # stage-by-stage mean PSF equals the fused build
    assert max_rel(got["chain_upsf"], got["upsf"]) < 1e-11
    assert max_rel(got["chain_parint"][:, 1:], got["parint"][:, 1:]) < 1e-11

    w = oracle_wcs(orc, inp.wcs)
    ref_map, det = orc.point_src_model_map_wcs(64, 64, w, st["src"], st["expo"], st["upsf"],
                                               st["parint"], DEG7, details=True)
    assert max_rel(got["map"], ref_map) < 1e-5
    assert max_rel(got["inv_map_integral"], det["inv_mapinteg"]) < 1e-5
    assert max_rel(got["correction"], det["cf"]) < 1e-5
    # the reference's stage sequence equals the fused model map
    assert max_rel(got["chain_map"], got["map"]) < 1e-5
    radius, in_fov = orc.psf_boundary_radius(64, 64, st["src"], w)
    assert max_rel(got["radius"], radius) < 1e-12
    np.testing.assert_array_equal(got["in_fov"].astype(bool), np.asarray(in_fov, bool))
    # SkyGeom on the first pixel row
    for j in range(64):
        lon, lat = orc.pix2sph(w, 1.0, j + 1.0)
        assert abs(got["row_sph"][j, 0] - lon) < 1e-9 and abs(got["row_sph"][j, 1] - lat) < 1e-9
    np.testing.assert_allclose(got["row_pix"][:, 0], 1.0, atol=1e-8)
    np.testing.assert_allclose(got["row_pix"][:, 1], np.arange(1, 65), atol=1e-8)


@pytest.mark.gpu
def test_cpp_pipeline_from_fits_inputs(orc, tmp_path):
    """main.cxx's pipeline with the IRFs read by Irf::load_aeff / load_psf from
    the P8R3 files and the livetime cubes by fits::read_expcube from a FITS
    file, against the oracle fed the same tables (read_irf_pars output, which
    test_fits_readers pins to the reference's fixtures)."""
    import fits_reader as fr
    import xtsrcmaps_amd as xa
    from pipelines import DEG7, oracle_wcs, read_xtsa, write_xtsa
    gold = os.path.join(ROOT, "tests", "golden")
    aeff_f = os.path.join(gold, "aeff_P8R3_SOURCE_V2_FB.fits")
    psf_f = os.path.join(gold, "psf_P8R3_SOURCE_V2_FB.fits")
    rng = np.random.default_rng(12)
    nside, nbr = 8, 40
    npix = 12 * nside * nside
    cubes = {}
    tabs = []
    for name in ("EXPOSURE", "WEIGHTED_EXPOSURE"):
        cos = rng.uniform(0, 1e7, (npix, nbr)).astype(np.float32)
        keys = {"NSIDE": nside, "NBRBINS": nbr, "COSMIN": 0.0, "ORDERING": "NESTED",
                "COORDSYS": "EQU", "THETABIN": "SQRT(1-COSTHETA)"}
        tabs.append((name, [("COSBINS", cos), ("RA", np.zeros((npix, 1), np.float32)),
                            ("DEC", np.zeros((npix, 1), np.float32))], keys))
        cubes[name] = cos.astype(np.float64)
    fr.write_bintables(tmp_path / "ltcube.fits", tabs)
    wcs = xa.Wcs.make((48, 48), (24.5, 24.5), (150.0, 20.0), (-0.1, 0.1), 0.0, "CAR")
    logE = np.linspace(2.0, 5.5, 11)
    src = np.array([[150.3, 20.2], [149.1, 19.0], [151.0, 21.5]])
    write_xtsa(tmp_path / "in.xtsa", {
        "wcs": [wcs.naxis1, wcs.naxis2, wcs.crpix1, wcs.crpix2, wcs.crval1, wcs.crval2,
                wcs.cdelt1, wcs.cdelt2, wcs.crota2, wcs.proj],
        "logE": logE, "src_sph": src})
    print("test_cpp_pipeline_from_fits_inputs: running the C++ driver", flush=True)
    r = subprocess.run([DRIVER, "--fits", aeff_f, psf_f, str(tmp_path / "ltcube.fits"),
                        str(tmp_path / "in.xtsa"), str(tmp_path / "out.xtsa")],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    got = read_xtsa(tmp_path / "out.xtsa")

    # the oracle on the same tables
    def grid(path, table):
        ext, off, rows = xa.read_irf_pars(path, table)
        cols = [rows[0][o:o + e] for e, o in zip(ext, off)]
        raw = np.stack([c.reshape(ext[2], ext[0]) for c in cols[4:]])
        return orc.prepare_grid(cols[0], cols[1], cols[2], cols[3], raw)
    aeff, king = {}, {}
    costh = orc.costhetas(nbr, 0.0, True)
    for side in ("front", "back"):
        cos, loge, par, minc = grid(aeff_f, f"EFFECTIVE AREA_{side.upper()}")
        if side == "front":
            np.testing.assert_array_equal(got["aeff_front_params"], par)
            np.testing.assert_array_equal(got["aeff_front_cosths"], cos)
            np.testing.assert_array_equal(got["aeff_front_logEs"], loge)
        aeff[side] = orc.aeff_value(costh, logE, cos, loge, par[..., 0], minc)
        cos, loge, par, _ = grid(psf_f, f"RPSF_{side.upper()}")
        _, _, scale = xa.read_irf_pars(psf_f, f"PSF_SCALING_PARAMS_{side.upper()}")
        king[side] = (orc.normalize_rpsf(par, loge, scale[0]), cos, loge)
    assert max_rel(got["psf_back_params"], king["back"][0]) < 1e-12
    _, _, effic = xa.read_irf_pars(aeff_f, "EFFICIENCY_PARAMS_FRONT")
    ltf = orc.ltf(logE, effic[0], effic[1])
    np.testing.assert_allclose(got["ltf"], ltf, rtol=1e-14)
    ecb = orc.src_cosbins(src, cubes["EXPOSURE"], nside)
    wecb = orc.src_cosbins(src, cubes["WEIGHTED_EXPOSURE"], nside)
    expo = orc.exposure(ecb, wecb, aeff["front"], aeff["back"], ltf)
    assert max_rel(got["exposure"], expo) < 1e-12
    upsf, parint = orc.psf_lookup_table_and_partial_integrals(
        king["front"], king["back"], costh, logE, aeff["front"], aeff["back"], ecb, wecb, ltf,
        expo)
    assert max_rel(got["upsf"], upsf) < 1e-10
    ref = orc.point_src_model_map_wcs(48, 48, oracle_wcs(orc, wcs), src, expo, upsf, parint,
                                      DEG7)
    assert max_rel(got["map"], ref) < 1e-5


DIFFUSE_DRIVER = os.path.join(ROOT, "tests", "cpp", "diffuse_driver")


def _read_arrays(path):
    """int64 rank, int64 extents, float64 data, repeated."""
    buf = open(path, "rb").read()
    out, pos = [], 0
    while pos < len(buf):
        rank = int(np.frombuffer(buf, np.int64, 1, pos)[0])
        shape = tuple(int(x) for x in np.frombuffer(buf, np.int64, rank, pos + 8))
        pos += 8 * (1 + rank)
        n = int(np.prod(shape))
        out.append(np.frombuffer(buf, np.float64, n, pos).reshape(shape))
        pos += 8 * n
    return out


def test_diffuse_driver_built_and_usage():
    if not os.path.exists(DIFFUSE_DRIVER):
        pytest.fail(f"{DIFFUSE_DRIVER} missing: run python -m xtsrcmaps_amd.build")
    r = subprocess.run([DIFFUSE_DRIVER], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "usage" in r.stderr


@pytest.mark.gpu
def test_cpp_diffuse_steps(tmp_path):
    """Fermi::ModelMap::diffuse_src_model_map_wcs, Psf::make_kernel and
    FFT::convolve_map_psf through the C++ layer (tests/cpp/diffuse_driver)
    equal the C ABI's results on the same inputs; the convolution equals
    scipy's within 1e-12 of its largest value."""
    import xtsrcmaps_amd as xa
    from scipy.signal import fftconvolve
    from test_fits_mapcube import _mapcube
    rng = np.random.default_rng(9)
    data = rng.uniform(0.5, 1.5, (3, 181, 360)).astype(np.float32)
    f = tmp_path / "cube.fits"
    _mapcube(f, data, energies=[1e2, 1e3, 1e4])
    print("test_cpp_diffuse_steps: running the C++ driver", flush=True)
    r = subprocess.run([DIFFUSE_DRIVER, str(f), str(tmp_path / "out.bin")],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    upsf, expo, cmap, kernel, img, ker, conv = _read_arrays(tmp_path / "out.bin")
    roi = xa.Wcs.make((40, 36), (20.5, 18.5), (193.98, -5.82), (-0.25, 0.25), 0.0, "CAR")
    w, cube, _ = xa.read_mapcube(f)
    with xa.Context(0) as ctx:
        ref = ctx.diffuse_src_model_map_wcs(roi, w, cube, upsf[0], expo[0], 0.5)
        kref = ctx.psf_make_kernel(15, roi, roi.crpix1, roi.crpix2, upsf[0])
        cref = ctx.fft_convolve_map_psf(img, ker)
    assert cmap.shape == (1, 40, 36, 4)
    np.testing.assert_array_equal(cmap[0], ref.transpose(1, 2, 0))
    np.testing.assert_array_equal(kernel, kref)
    np.testing.assert_array_equal(conv, cref)
    for c in range(2):
        s = fftconvolve(img[c], ker[c], mode="full")
        assert np.abs(conv[c] - s).max() < 1e-12 * np.abs(s).max()
