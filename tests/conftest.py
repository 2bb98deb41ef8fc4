import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def near_relative(a, b, rtol=1e-5, atol=1e-5):
    """|a - b| <= atol + rtol |b|, the reference's NearRelative
    (tests/fermi_tests.hxx:23-39)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return np.abs(a - b) <= atol + rtol * np.abs(b)


def max_rel(a, b, floor=1e-300):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), floor)))


@pytest.fixture(scope="session")
def golden():
    def load(name):
        return np.load(os.path.join(GOLDEN, name + ".npz"))
    return load


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle
    oracle.lib()
    return oracle


# the 3C279 ROI chain from the reference's IRF and livetime fixtures
@pytest.fixture(scope="session")
def psf_grids(orc, golden):
    P = golden("irf_psf")
    out = {}
    for k in ("front", "back"):
        cos, loge, par, _ = orc.prepare_grid(P[k + "_e_lo"], P[k + "_e_hi"],
                                             P[k + "_c_lo"], P[k + "_c_hi"], P[k + "_raw"])
        out[k] = dict(raw=par, norm=orc.normalize_rpsf(par, loge, P[k + "_scale"]),
                      cos=cos, loge=loge)
    return out


@pytest.fixture(scope="session")
def chain(orc, golden, psf_grids):
    """The 3C279 exposure + mean-PSF chain for all 263 sources."""
    A, X = golden("irf_aeff"), golden("exposure_psf")
    logE = np.log10(X["energies"])
    costh = orc.costhetas(40, 0.0, True)
    ltf = orc.ltf(logE, A["front_effic_p0"], A["front_effic_p1"])
    aeff = {}
    for k in ("front", "back"):
        cos, loge, par, minc = orc.prepare_grid(A[k + "_e_lo"], A[k + "_e_hi"],
                                                A[k + "_c_lo"], A[k + "_c_hi"], A[k + "_raw"])
        aeff[k] = orc.aeff_value(costh, logE, cos, loge, par[..., 0], minc)
    ecb, wecb = X["ecb"].astype(float), X["wecb"].astype(float)
    expo, parts = orc.exposure(ecb, wecb, aeff["front"], aeff["back"], ltf, parts=True)
    g = psf_grids
    upsf, parint = orc.psf_lookup_table_and_partial_integrals(
        (g["front"]["norm"], g["front"]["cos"], g["front"]["loge"]),
        (g["back"]["norm"], g["back"]["cos"], g["back"]["loge"]),
        costh, logE, aeff["front"], aeff["back"], ecb, wecb, ltf, expo)
    return dict(logE=logE, costh=costh, ltf=ltf, aeff=aeff, expo=expo, parts=parts,
                upsf=upsf, parint=parint, ecb=ecb, wecb=wecb)
