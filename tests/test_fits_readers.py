"""The FITS readers of the IRF tables and the livetime cube through the C ABI
(fits/fits.read_irf_pars.cxx:12-126, fits/fits.read_expcube.cxx:8-116).
The P8R3 IRF files are the ones the reference's tests read
(tests/test_fitsfuncs.cxx:62-154); the values they must yield are the
reference's own header fixtures (tests/fits/aeff_expected.hxx,
psf_expected.hxx), kept in tests/golden/irf_*.npz.  Host only."""
import os

import numpy as np
import pytest

import fits_reader as fr

HERE = os.path.dirname(os.path.abspath(__file__))
AEFF = os.path.join(HERE, "golden", "aeff_P8R3_SOURCE_V2_FB.fits")
PSF = os.path.join(HERE, "golden", "psf_P8R3_SOURCE_V2_FB.fits")


def _grid(extents, offsets, row):
    cols = [row[o:o + e] for e, o in zip(extents, offsets)]
    mc, me = extents[2], extents[0]
    return cols[0], cols[1], cols[2], cols[3], np.stack([c.reshape(mc, me) for c in cols[4:]])


@pytest.mark.parametrize("side", ["front", "back"])
def test_irf_tables_match_reference_fixtures(golden, side):
    import xtsrcmaps_amd as xa
    A, P = golden("irf_aeff"), golden("irf_psf")
    S = side.upper()
    ext, off, rows = xa.read_irf_pars(AEFF, f"EFFECTIVE AREA_{S}")
    assert list(ext) == [74, 74, 32, 32, 2368] and rows.shape == (1, 2580)
    e_lo, e_hi, c_lo, c_hi, raw = _grid(ext, off, rows[0])
    for got, key in ((e_lo, "e_lo"), (e_hi, "e_hi"), (c_lo, "c_lo"), (c_hi, "c_hi"),
                     (raw, "raw")):
        np.testing.assert_array_equal(got, A[f"{side}_{key}"])
    ext, off, rows = xa.read_irf_pars(AEFF, f"EFFICIENCY_PARAMS_{S}")
    assert list(ext) == [6] and rows.shape == (2, 6)
    np.testing.assert_array_equal(rows[0], A[f"{side}_effic_p0"])
    np.testing.assert_array_equal(rows[1], A[f"{side}_effic_p1"])
    ext, _, rows = xa.read_irf_pars(AEFF, f"PHI_DEPENDENCE_{S}")
    assert list(ext) == [23, 23, 8, 8, 184, 184] and rows.shape == (1, 430)

    ext, off, rows = xa.read_irf_pars(PSF, f"RPSF_{S}")
    assert list(ext) == [23, 23, 8, 8] + [184] * 6
    assert list(off) == [0, 23, 46, 54, 62, 246, 430, 614, 798, 982]
    e_lo, e_hi, c_lo, c_hi, raw = _grid(ext, off, rows[0])
    for got, key in ((e_lo, "e_lo"), (e_hi, "e_hi"), (c_lo, "c_lo"), (c_hi, "c_hi"),
                     (raw, "raw")):
        np.testing.assert_array_equal(got, P[f"{side}_{key}"])
    _, _, rows = xa.read_irf_pars(PSF, f"psf_scaling_params_{side}")   # case-insensitive
    np.testing.assert_array_equal(rows[0], P[f"{side}_scale"])
    ext, _, _ = xa.read_irf_pars(PSF, f"FISHEYE_CORRECTION_{S}")
    assert list(ext) == [46, 46, 16, 16, 736, 736, 736]


def test_prepared_grids_match_fixture_through_the_oracle(orc, golden):
    """The tables read here, gridded by the oracle's prepare_grid, give the
    reference's padded King grid (tests/expected/psf3_rpsf.*.bin)."""
    import xtsrcmaps_amd as xa
    P = golden("irf_psf")
    for side in ("front", "back"):
        ext, off, rows = xa.read_irf_pars(PSF, f"RPSF_{side.upper()}")
        e_lo, e_hi, c_lo, c_hi, raw = _grid(ext, off, rows[0])
        _, _, par, _ = orc.prepare_grid(e_lo, e_hi, c_lo, c_hi, raw)
        np.testing.assert_array_equal(par, P[f"{side}_rpsf"])


def test_irf_reader_errors(tmp_path):
    import xtsrcmaps_amd as xa
    with pytest.raises(xa.XtsError):
        xa.read_irf_pars(PSF, "NO_SUCH_TABLE")
    with pytest.raises(xa.XtsError):
        xa.read_irf_pars(tmp_path / "missing.fits", "RPSF_FRONT")
    p = tmp_path / "ints.fits"
    fr.write_bintables(p, [("T", [("A", np.zeros((1, 3), np.float32)),
                                  ("B", np.zeros((1, 2), np.int32))], {})])
    with pytest.raises(xa.XtsError):              # a non-float32 column
        xa.read_irf_pars(p, "T")
    (tmp_path / "junk.fits").write_bytes(b"x" * 2880)
    with pytest.raises(xa.XtsError):
        xa.read_irf_pars(tmp_path / "junk.fits", "T")


def _ltcube(path, nside, nbr, rng, ordering="NESTED", coordsys="EQU", thetabin="SQRT(1-COSTHETA)",
            dtype=np.float32):
    npix = 12 * nside * nside
    tabs = []
    cubes = {}
    for name in ("EXPOSURE", "WEIGHTED_EXPOSURE"):
        cos = rng.uniform(0, 1e6, (npix, nbr)).astype(dtype)
        ra = rng.uniform(0, 360, (npix, 1)).astype(dtype)
        dec = rng.uniform(-90, 90, (npix, 1)).astype(dtype)
        keys = {"NSIDE": nside, "NBRBINS": nbr, "COSMIN": 0.0, "ORDERING": ordering,
                "COORDSYS": coordsys, "THETABIN": thetabin}
        tabs.append((name, [("COSBINS", cos), ("RA", ra), ("DEC", dec)], keys))
        cubes[name] = (cos, ra[:, 0], dec[:, 0])
    fr.write_bintables(path, tabs)
    return cubes


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_expcube_round_trip(tmp_path, dtype):
    import xtsrcmaps_amd as xa
    rng = np.random.default_rng(5)
    p = tmp_path / "ltcube.fits"
    cubes = _ltcube(p, 4, 40, rng, dtype=dtype)
    for name, (cos, ra, dec) in cubes.items():
        d = xa.read_expcube(p, name)
        assert (d["nside"], d["nbrbins"], d["cosmin"]) == (4, 40, 0.0)
        assert d["ordering"] == "NESTED" and d["coordsys"] == "EQU" and d["thetabin"]
        np.testing.assert_array_equal(d["cosbins"], cos.astype(np.float32))
        np.testing.assert_array_equal(d["ra"], ra.astype(np.float32))
        np.testing.assert_array_equal(d["dec"], dec.astype(np.float32))
    _ltcube(p, 2, 8, rng, thetabin="COSTHETA")
    assert not xa.read_expcube(p, "EXPOSURE")["thetabin"]


def test_expcube_rejects_what_the_reference_rejects(tmp_path):
    import xtsrcmaps_amd as xa
    rng = np.random.default_rng(6)
    for kw in ({"ordering": "RING"}, {"coordsys": "GAL"}):
        p = tmp_path / f"bad_{list(kw.values())[0]}.fits"
        _ltcube(p, 2, 8, rng, **kw)
        with pytest.raises(xa.XtsError):
            xa.read_expcube(p, "EXPOSURE")
    p = tmp_path / "nokey.fits"
    fr.write_bintables(p, [("EXPOSURE", [("COSBINS", np.zeros((48, 8), np.float32)),
                                         ("RA", np.zeros((48, 1), np.float32)),
                                         ("DEC", np.zeros((48, 1), np.float32))],
                            {"NSIDE": 2, "NBRBINS": 8})])
    with pytest.raises(xa.XtsError):
        xa.read_expcube(p, "EXPOSURE")
