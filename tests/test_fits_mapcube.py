"""The MapCube reader (xts_read_mapcube, the input of
fits/fits.read_allsky_cropped.cpp:150-216): a FITS file written here card by
card (primary image of big-endian floats, an ENERGIES binary table) read back
bit for bit, with its WCS and energies; malformed files refused."""
import numpy as np
import pytest


def _card(key, value):
    if isinstance(value, str):
        v = "'" + value.replace("'", "''").ljust(8) + "'"
        return f"{key:<8}= {v:<70}"[:80]
    if isinstance(value, bool):
        return f"{key:<8}= {'T' if value else 'F':>20}".ljust(80)
    return f"{key:<8}= {value!r:>20}".ljust(80)


def _block(cards, data=b""):
    h = "".join(cards + ["END".ljust(80)])
    h += " " * (-len(h) % 2880)
    return h.encode("ascii") + data + b"\0" * (-len(data) % 2880)


def _mapcube(path, data, ctype=("GLON-CAR", "GLAT-CAR"), energies=None, bitpix=-32,
             truncate=0, crota2=None):
    planes, n2, n1 = data.shape
    cards = [_card("SIMPLE", True), _card("BITPIX", bitpix), _card("NAXIS", 3),
             _card("NAXIS1", n1), _card("NAXIS2", n2), _card("NAXIS3", planes),
             _card("EXTEND", True), _card("CTYPE1", ctype[0]), _card("CTYPE2", ctype[1]),
             _card("CRPIX1", (n1 + 1) / 2.0), _card("CRPIX2", (n2 + 1) / 2.0),
             _card("CRVAL1", 0.0), _card("CRVAL2", 0.0), _card("CDELT1", -1.0),
             _card("CDELT2", 1.0)]
    if crota2 is not None:
        cards.append(_card("CROTA2", crota2))
    dt = ">f4" if bitpix == -32 else ">f8"
    raw = data.astype(dt).tobytes()
    out = _block(cards, raw)
    if energies is not None:
        e = np.asarray(energies, ">f8").tobytes()
        tcards = [_card("XTENSION", "BINTABLE"), _card("BITPIX", 8), _card("NAXIS", 2),
                  _card("NAXIS1", 8), _card("NAXIS2", len(energies)), _card("PCOUNT", 0),
                  _card("GCOUNT", 1), _card("TFIELDS", 1), _card("TTYPE1", "Energy"),
                  _card("TFORM1", "D"), _card("EXTNAME", "ENERGIES")]
        out += _block(tcards, e)
    if truncate:
        out = out[:-truncate]
    path.write_bytes(out)


def test_mapcube_round_trip(tmp_path):
    import xtsrcmaps_amd as xa
    rng = np.random.default_rng(2)
    data = rng.uniform(1e-8, 1e-3, (4, 9, 16)).astype(np.float32)
    en = [100.0, 1e3, 1e4, 1e5]
    f = tmp_path / "cube.fits"
    _mapcube(f, data, energies=en, crota2=0.0)
    w, got, energies = xa.read_mapcube(f)
    assert got.dtype == np.float32 and np.array_equal(got, data)
    assert np.array_equal(energies, en)
    assert (w.naxis1, w.naxis2, w.proj, w.is_galactic) == (16, 9, xa.PROJ_CAR, 1)
    assert (w.crpix1, w.crpix2, w.cdelt1, w.cdelt2, w.crota2) == (8.5, 5.0, -1.0, 1.0, 0.0)


def test_mapcube_celestial_ait_f64_without_energies(tmp_path):
    import xtsrcmaps_amd as xa
    data = np.arange(2 * 5 * 7, dtype=np.float64).reshape(2, 5, 7) * 0.5
    f = tmp_path / "cel.fits"
    _mapcube(f, data, ctype=("RA---AIT", "DEC--AIT"), bitpix=-64)
    w, got, energies = xa.read_mapcube(f)
    assert np.array_equal(got, data.astype(np.float32)) and energies is None
    assert (w.proj, w.is_galactic, w.crota2) == (xa.PROJ_AIT, 0, 0.0)


@pytest.mark.parametrize("kind", ["truncated", "tan", "bitpix16"])
def test_mapcube_refuses_malformed(tmp_path, kind):
    import xtsrcmaps_amd as xa
    data = np.ones((2, 4, 6), np.float32)
    f = tmp_path / "bad.fits"
    if kind == "truncated":
        _mapcube(f, data, truncate=2900)
    elif kind == "tan":
        _mapcube(f, data, ctype=("RA---TAN", "DEC--TAN"))
    else:
        _mapcube(f, data.astype(np.int16), bitpix=16)
    with pytest.raises(xa.XtsError):
        xa.read_mapcube(f)
