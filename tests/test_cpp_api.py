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
