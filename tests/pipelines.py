"""The same synthetic configuration run through the oracle and through the
HIP product, stage by stage (test helper)."""
from __future__ import annotations

import numpy as np

from xtsrcmaps_amd import synthetic


def oracle_stages(orc, inp: synthetic.Inputs, n_src: int | None = None, rule=None):
    """FP64 oracle of (a), (b) and the inputs of (c)."""
    src = inp.src_sph if n_src is None else inp.src_sph[:n_src]
    out = {"src": src}
    king = {}
    for side, (params, cos, loge) in inp.king_raw.items():
        king[side] = (orc.normalize_rpsf(params, loge, np.array(inp.psf_scale, np.float32)),
                      cos, loge)
    out["king"] = king
    aeff = {}
    for side, (cos, loge, params, minc) in inp.aeff_grid.items():
        aeff[side] = orc.aeff_value(inp.costh, inp.logE, cos, loge, params, minc)
    out["aeff"] = aeff
    out["ltf"] = orc.ltf(inp.logE, inp.effic[0], inp.effic[1])
    out["ecb"] = orc.src_cosbins(src, inp.cube, inp.nside)
    out["wecb"] = orc.src_cosbins(src, inp.wcube, inp.nside)
    out["expo"] = orc.exposure(out["ecb"], out["wecb"], aeff["front"], aeff["back"], out["ltf"])
    out["upsf"], out["parint"] = orc.psf_lookup_table_and_partial_integrals(
        king["front"], king["back"], inp.costh, inp.logE, aeff["front"], aeff["back"],
        out["ecb"], out["wecb"], out["ltf"], out["expo"])
    return out


def oracle_rule(inp: synthetic.Inputs):
    return inp.rule if inp.rule is not None else DEG7


# Festa-Sommariva degree-7 rule, identical to the library's built-in table
DEG7 = np.array([
    [3.9383136101878902e-01, -3.6699793434042738e-01, 5.2092862230230441e-01],
    [8.0196847620659251e-01, 8.0992592975809186e-01, 2.3740240561379911e-01],
    [-1.8327327664247231e-02, 9.2596092593108370e-01, 2.4166897208389651e-01],
    [3.6699793434042738e-01, 3.9383136101878902e-01, 5.2092862230230441e-01],
    [-8.0992592975809186e-01, 8.0196847620659251e-01, 2.3740240561379911e-01],
    [-9.2596092593108370e-01, -1.8327327664247231e-02, 2.4166897208389651e-01],
    [-3.9383136101878902e-01, 3.6699793434042738e-01, 5.2092862230230441e-01],
    [-8.0196847620659251e-01, -8.0992592975809186e-01, 2.3740240561379911e-01],
    [1.8327327664247231e-02, -9.2596092593108370e-01, 2.4166897208389651e-01],
    [-3.6699793434042738e-01, -3.9383136101878902e-01, 5.2092862230230441e-01],
    [8.0992592975809186e-01, -8.0196847620659251e-01, 2.3740240561379911e-01],
    [9.2596092593108370e-01, 1.8327327664247231e-02, 2.4166897208389651e-01],
])


def oracle_wcs(orc, wcs):
    return orc.make_wcs((wcs.naxis1, wcs.naxis2), (wcs.crpix1, wcs.crpix2),
                        (wcs.crval1, wcs.crval2), (wcs.cdelt1, wcs.cdelt2), wcs.crota2,
                        "CAR" if wcs.proj == 0 else "AIT")
