"""Extract the reference's own test vectors into small committed fixtures.

Run once, in a container where the reference tree is mounted:

    python tests/golden/make_golden.py /root/reference

It reads only DATA: the numeric initialiser lists in the reference's
``tests/fits/*.hxx`` files (parsed as numbers, no code is compiled or run),
the raw little-endian binaries in ``tests/expected/*.bin``, the FITS header
cards of ``analysis/3C279_binned_ccube.fits``, two HDUs of FermiTools'
``analysis/canonical_srcmap.fits``, the P8R3 IRF FITS files (copied) and
the source position in ``analysis/single_point.xml``.
``--canonical-only`` / ``--irf-only`` write only those fixtures.  Outputs are ``.npz`` files next to this
script; large maps are subsampled on a fixed pixel lattice so the fixtures
stay small.
"""
from __future__ import annotations

import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

_NUM = re.compile(r"[-+]?(?:\d+\.\d*|\.\d+|\d+)(?:[eE][-+]?\d+)?")


def _strip_comments(text: str) -> str:
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return re.sub(r"//[^\n]*", "", text)


def _block(text: str, name: str, occurrence: int = 0) -> str:
    """Balanced-brace initialiser that follows the `occurrence`-th `name`."""
    pos = -1
    for _ in range(occurrence + 1):
        m = re.compile(r"\b" + re.escape(name) + r"\b").search(text, pos + 1)
        if m is None:
            raise KeyError(name)
        pos = m.start()
    start = text.index("{", pos)
    depth = 0
    for i in range(start, len(text)):
        if text[i] == "{":
            depth += 1
        elif text[i] == "}":
            depth -= 1
            if depth == 0:
                return text[start + 1 : i]
    raise ValueError(name)


def _top_groups(block: str) -> list[str]:
    """Split an initialiser body into its top-level { ... } groups."""
    out, depth, start = [], 0, None
    for i, ch in enumerate(block):
        if ch == "{":
            if depth == 0:
                start = i + 1
            depth += 1
        elif ch == "}":
            depth -= 1
            if depth == 0:
                out.append(block[start:i])
    return out


def _nums(s: str, dtype=np.float64) -> np.ndarray:
    return np.array([float(x) for x in _NUM.findall(s)], dtype=dtype)


def _vec(text, name, occ=0, dtype=np.float64):
    return _nums(_block(text, name, occ), dtype)


def _table(text, name):
    """A FITS-row struct {extents}, {offsets}, {rowdata...} as float32."""
    g = _top_groups(_block(text, name))
    ext = _nums(g[0], np.float64).astype(np.int64)
    off = _nums(g[1], np.float64).astype(np.int64)
    rows = [_nums(x, np.float32) for x in g[2:]]
    return ext, off, rows


def _grid_arrays(ext, off, row):
    """Split a raw IRF row into its energy/costheta edges and the params."""
    me0, mc0 = int(ext[0]), int(ext[2])
    e_lo = row[off[0] : off[0] + me0]
    e_hi = row[off[1] : off[1] + me0]
    c_lo = row[off[2] : off[2] + mc0]
    c_hi = row[off[3] : off[3] + mc0]
    ng = len(ext) - 4
    raw = np.concatenate([row[off[4 + p] : off[4 + p] + me0 * mc0] for p in range(ng)])
    return e_lo, e_hi, c_lo, c_hi, raw.reshape(ng, mc0, me0)


def _read(ref, name, dtype=np.float64):
    return np.fromfile(os.path.join(ref, "tests", "expected", name), dtype=dtype)


def _fits_header(path):
    cards = {}
    with open(path, "rb") as f:
        while True:
            blk = f.read(2880)
            if not blk:
                break
            for i in range(0, 2880, 80):
                card = blk[i : i + 80].decode("ascii", "replace")
                key = card[:8].strip()
                if key == "END":
                    return cards
                if card[8:10] == "= ":
                    val = card[10:].split("/")[0].strip()
                    cards[key] = val.strip("'").strip()
    return cards


def main(ref: str) -> None:
    fits = os.path.join(ref, "tests", "fits")
    rd = lambda n: _strip_comments(open(os.path.join(fits, n)).read())
    psf_t, aeff_t = rd("psf_expected.hxx"), rd("aeff_expected.hxx")
    cos_t, sep_t = rd("exposure_cosbins_expected.hxx"), rd("separations_expected.hxx")
    ene_t, dir_t = rd("ccube_energies_expected.hxx"), rd("src_dirs_expected.hxx")

    # ---- PSF IRF: raw rows, scaling, prepared and normalised grids, King table
    out = {}
    for side in ("FRONT", "BACK"):
        ext, off, rows = _table(psf_t, "RPSF_" + side)
        e_lo, e_hi, c_lo, c_hi, raw = _grid_arrays(ext, off, rows[0])
        _, _, srows = _table(psf_t, "PSF_SCALING_PARAMS_" + side)
        k = side.lower()
        out.update({f"{k}_e_lo": e_lo, f"{k}_e_hi": e_hi, f"{k}_c_lo": c_lo,
                    f"{k}_c_hi": c_hi, f"{k}_raw": raw, f"{k}_scale": srows[0]})
        out[f"{k}_rpsf"] = _read(ref, f"psf3_rpsf.{k}.bin").reshape(10, 25, 6)
        out[f"{k}_norm"] = _read(ref, f"psf3_normalized_rpsf.{k}.bin").reshape(10, 25, 6)
    out["king_front"] = _read(ref, "king_front.bin").reshape(401, 10, 25)
    out["separations"] = _vec(sep_t, "expected_separations")
    np.savez_compressed(os.path.join(HERE, "irf_psf.npz"), **out)

    # ---- Aeff IRF, efficiency (LTF) parameters and their expected products
    out = {}
    for side in ("FRONT", "BACK"):
        ext, off, rows = _table(aeff_t, "EFFECTIVE_AREA_" + side)
        e_lo, e_hi, c_lo, c_hi, raw = _grid_arrays(ext, off, rows[0])
        k = side.lower()
        out.update({f"{k}_e_lo": e_lo, f"{k}_e_hi": e_hi, f"{k}_c_lo": c_lo,
                    f"{k}_c_hi": c_hi, f"{k}_raw": raw})
        _, _, erows = _table(aeff_t, "EFFICIENCY_PARAMS_" + side)
        out[f"{k}_effic_p0"], out[f"{k}_effic_p1"] = erows[0], erows[1]
    out["livetime_factor1"] = _vec(aeff_t, "livetime_factor1")
    out["livetime_factor2"] = _vec(aeff_t, "livetime_factor2")
    out["front_exp_aeff"] = _vec(aeff_t, "exp_aeff", 0).reshape(263, 38)
    out["front_wexp_aeff"] = _vec(aeff_t, "wexp_aeff", 0).reshape(263, 38)
    out["back_exp_aeff"] = _vec(aeff_t, "exp_aeff", 1).reshape(263, 38)
    out["back_wexp_aeff"] = _vec(aeff_t, "wexp_aeff", 1).reshape(263, 38)
    np.savez_compressed(os.path.join(HERE, "irf_aeff.npz"), **out)

    # ---- Exposure inputs/outputs for the 263 3C279 sources
    out = {
        "exp_costhetas": _vec(cos_t, "exp_costhetas", dtype=np.float32),
        "ecb": _vec(cos_t, "exposure_source_cosine_bins", dtype=np.float32).reshape(263, 40),
        "wecb": _vec(cos_t, "weighted_exposure_source_cosine_bins", dtype=np.float32).reshape(263, 40),
        "aeff_front_c_e": _vec(cos_t, "aeff_front_c_e", dtype=np.float32).reshape(40, 38),
        "aeff_back_c_e": _vec(cos_t, "aeff_back_c_e", dtype=np.float32).reshape(40, 38),
        "energies": _vec(ene_t, "expected"),
        "src_dirs": _vec(dir_t, "dirs").reshape(263, 2),
        "exposure": _read(ref, "exposure.bin").reshape(263, 38),
    }
    for n in ("lef", "lwf", "leb", "lwb"):
        out["exposure_" + n] = _read(ref, f"exposure_{n}.bin").reshape(263, 38)
    out["upsf_peak"] = _read(ref, "uPsf_peak_SE.bin").reshape(263, 38)
    np.savez_compressed(os.path.join(HERE, "exposure_psf.npz"), **out)

    # ---- Single-source pixel-cubature maps on the 3C279 AIT ROI ([W, H, E])
    hdr = _fits_header(os.path.join(ref, "analysis", "3C279_binned_ccube.fits"))
    xml = open(os.path.join(ref, "analysis", "single_point.xml")).read()
    ra = float(re.search(r'name="RA"[^>]*value="([^"]+)"', xml).group(1))
    dec = float(re.search(r'name="DEC"[^>]*value="([^"]+)"', xml).group(1))
    lattice = np.arange(0, 100, 7)
    out = {
        "naxis": np.array([int(hdr["NAXIS1"]), int(hdr["NAXIS2"])]),
        "crpix": np.array([float(hdr["CRPIX1"]), float(hdr["CRPIX2"])]),
        "crval": np.array([float(hdr["CRVAL1"]), float(hdr["CRVAL2"])]),
        "cdelt": np.array([float(hdr["CDELT1"]), float(hdr["CDELT2"])]),
        "crota2": np.array(float(hdr.get("CROTA2", "0"))),
        "proj": np.array(hdr["CTYPE1"][-3:]),
        "src_sph": np.array([ra, dec]),
        "lattice": lattice,
    }
    for n in ("mm_psf_value_4096", "mm_sqd3.double", "mm_sqd11.double",
              "mm_asqd12_single_src.double"):
        m = _read(ref, n + ".bin").reshape(100, 100, 38)
        out[n.split(".")[0]] = m[np.ix_(lattice, lattice)]
    cub = open(os.path.join(ref, "xtsrcmaps", "model_map", "mm_cubature", "cubature.hpp")).read()
    cub = _strip_comments(cub)
    for n in ("square_deg7", "square_deg11", "asym_square_d12"):
        out["rule_" + n] = _nums(_block(cub, n)).reshape(-1, 3)
    np.savez_compressed(os.path.join(HERE, "mm_single_source.npz"), **out)

    # ---- FermiTools' own boundary/integral quantities for the 263 sources
    out = {
        "within_bounds": _read(ref, "ft_mm_withinBounds.bin", np.uint8),
        "psf_radius": _read(ref, "ft_mm_psf_radius.bin"),
        "upsf_radius_integ": _read(ref, "ft_mm_upsf_radius_integ.bin"),
        "correction_factor": _read(ref, "ft_mm_correction_factor.bin"),
        "map_integrals": _read(ref, "ft_mm_mapIntegrals.bin"),
    }
    np.savez_compressed(os.path.join(HERE, "fermitools_boundary.npz"), **out)


def _fits_hdus(path):
    """(header bytes, data bytes) of every HDU of a FITS file, read as bytes."""
    buf = open(path, "rb").read()
    pos, out = 0, []
    while pos < len(buf):
        start = pos
        hdr = {}
        while True:
            block = buf[pos:pos + 2880]
            pos += 2880
            cards = [block[i:i + 80].decode("ascii") for i in range(0, 2880, 80)]
            for c in cards:
                if c[8:10] == "= ":
                    hdr[c[:8].strip()] = c[10:].split("/")[0].strip().strip("'").strip()
            if any(c.startswith("END     ") or c.rstrip() == "END" for c in cards):
                break
        header = buf[start:pos]
        naxis = int(hdr.get("NAXIS", 0))
        n = abs(int(hdr["BITPIX"])) // 8 if naxis else 0
        for k in range(naxis):
            n *= int(hdr[f"NAXIS{k + 1}"])
        n += int(hdr.get("PCOUNT", 0))
        n = (n + 2879) // 2880 * 2880
        out.append((header, buf[pos:pos + n]))
        pos += n
    return out


def irf_files(ref: str) -> None:
    """The P8R3 SOURCE IRF files the reference's tests read
    (config/config.hxx:28-29), copied as data for the FITS readers."""
    for name in ("aeff_P8R3_SOURCE_V2_FB.fits", "psf_P8R3_SOURCE_V2_FB.fits"):
        data = open(os.path.join(ref, "analysis", name), "rb").read()
        open(os.path.join(HERE, name), "wb").write(data)


def canonical_srcmap(ref: str) -> None:
    """FermiTools' gtsrcmaps output for the 3C279 ROI: the EBOUNDS table HDU
    (bytes, for its CHECKSUM/DATASUM cards) and the first source's image HDU
    (header bytes and the float32 [38, 100, 100] cube)."""
    hdus = _fits_hdus(os.path.join(ref, "analysis", "canonical_srcmap.fits"))
    eb_h, eb_d = hdus[2]
    src_h, src_d = hdus[3]
    cube = np.frombuffer(src_d[:38 * 100 * 100 * 4], dtype=">f4").reshape(38, 100, 100)
    out = {
        "ebounds_hdu": np.frombuffer(eb_h + eb_d, dtype=np.uint8),
        "src_header": np.frombuffer(src_h, dtype=np.uint8),
        "src_cube": cube.astype(np.float32),
    }
    np.savez_compressed(os.path.join(HERE, "canonical_srcmap.npz"), **out)


if __name__ == "__main__":
    args = [a for a in sys.argv[1:] if not a.startswith("--")]
    ref = args[0] if args else "/root/reference"
    if "--canonical-only" not in sys.argv and "--irf-only" not in sys.argv:
        main(ref)
    if "--irf-only" not in sys.argv:
        canonical_srcmap(ref)
    if "--canonical-only" not in sys.argv:
        irf_files(ref)
