"""ctypes wrapper around the FP64 CPU oracle (``libxts_oracle.so``).

TEST INFRASTRUCTURE ONLY: the oracle is the checker the HIP product path is
compared against.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` import this module.  Each wrapper names
the C function (and through it the reference file:line it restates).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libxts_oracle.so")
ND = 401
PROJ = {"CAR": 0, "AIT": 1}

_lib = None

_d = C.POINTER(C.c_double)
_f = C.POINTER(C.c_float)
_u8 = C.POINTER(C.c_uint8)
_l = C.c_long


class Wcs(C.Structure):
    _fields_ = [("naxis1", C.c_long), ("naxis2", C.c_long)] + [
        (n, C.c_double)
        for n in ("crpix1", "crpix2", "crval1", "crval2", "cdelt1", "cdelt2", "crota2")
    ] + [("proj", C.c_int)] + [
        (n, C.c_double)
        for n in ("m11", "m12", "m21", "m22", "i11", "i12", "i21", "i22",
                  "alphap", "deltap", "phip")
    ]


def build(force: bool = False) -> str:
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-B" if force else "-s", "-C", HERE],
                       check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = C.CDLL(LIB_PATH)
        L = _lib
        L.or_evaluate_king.restype = C.c_double
        L.or_psf_base_integral.restype = C.c_double
        L.or_sep_step.restype = C.c_double
        L.or_dir_diff.restype = C.c_double
        L.or_ang2pix_nest.restype = C.c_uint64
        L.or_ang2pix_nest.argtypes = [C.c_double, C.c_double, C.c_int64]
        L.or_fast_separation_lower_index.argtypes = [C.c_double]
        L.or_evaluate_king.argtypes = [C.c_double, C.c_double, _d]
        L.or_psf_base_integral.argtypes = [C.c_double, C.c_double, _d]
        L.or_normalize_rpsf.argtypes = [_d, _d, _l, _l, C.c_float, C.c_float, C.c_float]
        L.or_pix2sph.argtypes = [C.POINTER(Wcs), C.c_double, C.c_double, _d, _d]
        L.or_sph2pix.argtypes = [C.POINTER(Wcs), C.c_double, C.c_double, _d, _d]
        L.or_pix2dir.argtypes = [C.POINTER(Wcs), C.c_double, C.c_double, _d]
        L.or_sph2dir.argtypes = [C.c_double, C.c_double, _d]
    return _lib


def _p(a, kind=_d):
    return a.ctypes.data_as(kind)


def _c(a, dtype=np.float64):
    return np.ascontiguousarray(a, dtype=dtype)


# ---------------------------------------------------------------- constants
def separations() -> np.ndarray:
    out = np.zeros(ND)
    lib().or_separations(_p(out))
    return out


def fast_separation_lower_index(v: float) -> int:
    return lib().or_fast_separation_lower_index(float(v))


# ---------------------------------------------------------------- IRF math
def evaluate_king(sep_rad, sf, pars6):
    p = _c(pars6)
    return lib().or_evaluate_king(float(sep_rad), float(sf), _p(p))


def psf_base_integral(radius_deg, sf, pars6):
    p = _c(pars6)
    return lib().or_psf_base_integral(float(radius_deg), float(sf), _p(p))


def legendre_rw(n, eps=1e-15):
    x, w = np.zeros(n), np.zeros(n)
    lib().or_legendre_rw(C.c_int(n), C.c_double(eps), _p(x), _p(w))
    return x, w


def prepare_grid(e_lo, e_hi, c_lo, c_hi, raw):
    """raw is [Ng, Mc0, Me0] float32 -> (cosths, logEs, params[Mc,Me,Ng], minc)."""
    raw = _c(raw, np.float32)
    ng, mc0, me0 = raw.shape
    cos = np.zeros(mc0 + 2)
    loge = np.zeros(me0 + 2)
    par = np.zeros((mc0 + 2, me0 + 2, ng))
    minc = C.c_double()
    lib().or_prepare_grid(_p(_c(e_lo, np.float32), _f), _p(_c(e_hi, np.float32), _f), _l(me0),
                          _p(_c(c_lo, np.float32), _f), _p(_c(c_hi, np.float32), _f), _l(mc0),
                          _p(raw, _f), _l(ng), _p(cos), _p(loge), _p(par), C.byref(minc))
    return cos, loge, par, minc.value


def normalize_rpsf(params, logEs, scale3):
    p = _c(params).copy()
    mc, me, _ = p.shape
    lib().or_normalize_rpsf(_p(p), _p(_c(logEs)), mc, me,
                            float(scale3[0]), float(scale3[1]), float(scale3[2]))
    return p


def ltf(logEs, p0, p1):
    E = len(logEs)
    out = np.zeros((2, E))
    lib().or_ltf(_p(_c(logEs)), _l(E), _p(_c(p0, np.float32), _f),
                 _p(_c(p1, np.float32), _f), _p(out))
    return out


def aeff_value(costh, logE, pc, pe, params, minc):
    C_, E = len(costh), len(logE)
    out = np.zeros((C_, E))
    par = _c(params).reshape(len(pc), len(pe))
    rc = lib().or_aeff_value(_p(_c(costh)), _l(C_), _p(_c(logE)), _l(E),
                             _p(_c(pc)), _l(len(pc)), _p(_c(pe)), _l(len(pe)),
                             _p(par), C.c_double(minc), _p(out))
    if rc:
        raise ValueError("aeff_value: sample outside the IRF grid")
    return out


# ---------------------------------------------------------------- exposure
def costhetas(nbins, cosmin, thetabin=True):
    out = np.zeros(nbins)
    lib().or_costhetas(_l(nbins), C.c_double(cosmin), C.c_int(int(thetabin)), _p(out))
    return out


def ang2pix_nest(theta, phi, nside):
    return lib().or_ang2pix_nest(float(theta), float(phi), int(nside))


def src_cosbins(src_sph, cube, nside):
    src = _c(src_sph)
    cube = _c(cube)
    S, Cn = src.shape[0], cube.shape[1]
    out = np.zeros((S, Cn))
    lib().or_src_cosbins(_p(src), _l(S), _p(cube), C.c_int64(nside), _l(Cn), _p(out))
    return out


def exposure(ecb, wecb, af, ab, ltf_, parts=False):
    ecb, wecb, af, ab, ltf_ = map(_c, (ecb, wecb, af, ab, ltf_))
    S, Cn = ecb.shape
    E = af.shape[1]
    out = np.zeros((S, E))
    if parts:
        ps = [np.zeros((S, E)) for _ in range(4)]
        lib().or_exposure(_p(ecb), _p(wecb), _p(af), _p(ab), _p(ltf_), _l(S), _l(Cn),
                          _l(E), _p(out), *[_p(x) for x in ps])
        return out, ps
    nul = C.cast(None, _d)
    lib().or_exposure(_p(ecb), _p(wecb), _p(af), _p(ab), _p(ltf_), _l(S), _l(Cn), _l(E),
                      _p(out), nul, nul, nul, nul)
    return out


# ---------------------------------------------------------------- mean PSF
def king(params):
    p = _c(params)
    mc, me, _ = p.shape
    out = np.zeros((ND, mc, me))
    lib().or_king(_p(p), _l(mc), _l(me), _p(out))
    return out


def psf_lookup_table_and_partial_integrals(front, back, costh, logE, af, ab,
                                           ecb, wecb, ltf_, expo):
    """front/back = (params[Mc,Me,6] normalised, cosths[Mc], logEs[Me])."""
    pf, cf, ef = (_c(x) for x in front)
    pb, cb, eb = (_c(x) for x in back)
    costh, logE, af, ab, ecb, wecb, ltf_, expo = map(
        _c, (costh, logE, af, ab, ecb, wecb, ltf_, expo))
    S, E = expo.shape
    upsf = np.zeros((S, ND, E))
    parint = np.zeros((S, ND, E))
    rc = lib().or_psf_lookup_table_and_partial_integrals(
        _p(pf), _p(cf), _l(len(cf)), _p(ef), _l(len(ef)),
        _p(pb), _p(cb), _l(len(cb)), _p(eb), _l(len(eb)),
        _p(costh), _l(len(costh)), _p(logE), _l(E), _p(af), _p(ab), _p(ecb),
        _p(wecb), _p(ltf_), _p(expo), _l(S), _p(upsf), _p(parint))
    if rc:
        raise ValueError("psf lookup: sample outside the IRF grid")
    return upsf, parint


def bilerp(costh, logE, pc, pe, kings):
    """psf/psf.bilerp.cxx:6-52 -> [C, 401, E]."""
    costh, logE, pc, pe, k = map(_c, (costh, logE, pc, pe, kings))
    out = np.zeros((len(costh), ND, len(logE)))
    rc = lib().or_bilerp(_p(costh), _l(len(costh)), _p(logE), _l(len(logE)), _p(pc),
                         _l(len(pc)), _p(pe), _l(len(pe)), _p(k), _p(out))
    if rc:
        raise ValueError("bilerp: sample outside the IRF grid")
    return out


def corrected_exposure_psf(obs_psf, aeff, ecb, wecb, ltf_):
    """psf/psf.corrected_exposure_psf.cxx:12-120 -> [S, 401, E]."""
    obs_psf, aeff, ecb, wecb, ltf_ = map(_c, (obs_psf, aeff, ecb, wecb, ltf_))
    C_, D, E = obs_psf.shape
    S = ecb.shape[0]
    out = np.zeros((S, D, E))
    lib().or_corrected_exposure_psf(_p(obs_psf), _p(aeff), _p(ecb), _p(wecb), _p(ltf_), _l(S),
                                    _l(C_), _l(D), _l(E), _p(out))
    return out


def mean_psf(front, back, expo):
    """psf/psf.mean_psf.cxx:6-33."""
    front, back, expo = map(_c, (front, back, expo))
    S, D, E = front.shape
    out = np.zeros_like(front)
    lib().or_mean_psf(_p(front), _p(back), _p(expo), _l(S), _l(D), _l(E), _p(out))
    return out


def normalize(upsf, total):
    """psf/psf.normalize.cxx:7-24 (returns a normalised copy)."""
    u = _c(upsf).copy()
    t = _c(total)
    S, D, E = u.shape
    lib().or_normalize(_p(u), _p(t), _l(S), _l(D), _l(E))
    return u


def partial_total_integral(upsf):
    u = _c(upsf)
    S, D, E = u.shape
    par, tot = np.zeros_like(u), np.zeros((S, E))
    lib().or_partial_total_integral(_p(u), _l(S), _l(D), _l(E), _p(par), _p(tot))
    return par, tot


# ---------------------------------------------------------------- WCS
def make_wcs(naxis, crpix, crval, cdelt, crota2=0.0, proj="CAR") -> Wcs:
    w = Wcs()
    w.naxis1, w.naxis2 = int(naxis[0]), int(naxis[1])
    w.crpix1, w.crpix2 = float(crpix[0]), float(crpix[1])
    w.crval1, w.crval2 = float(crval[0]), float(crval[1])
    w.cdelt1, w.cdelt2 = float(cdelt[0]), float(cdelt[1])
    w.crota2 = float(crota2)
    w.proj = PROJ[str(proj)]
    lib().or_wcs_setup(C.byref(w))
    return w


def pix2sph(w, p1, p2):
    a, b = C.c_double(), C.c_double()
    rc = lib().or_pix2sph(C.byref(w), float(p1), float(p2), C.byref(a), C.byref(b))
    return (a.value, b.value) if rc == 0 else (np.nan, np.nan)


def sph2pix(w, lon, lat):
    a, b = C.c_double(), C.c_double()
    lib().or_sph2pix(C.byref(w), float(lon), float(lat), C.byref(a), C.byref(b))
    return a.value, b.value


def pix2dir(w, p1, p2):
    d = np.zeros(3)
    lib().or_pix2dir(C.byref(w), float(p1), float(p2), _p(d))
    return d


def sph2dir(lon, lat):
    d = np.zeros(3)
    lib().or_sph2dir(float(lon), float(lat), _p(d))
    return d


# ---------------------------------------------------------------- model map
def square_ptswts(Nh, Nw, w, rule):
    rule = _c(rule)
    N = rule.shape[0]
    out = np.zeros((Nh, Nw, 4 * N))
    lib().or_square_ptswts(_l(Nh), _l(Nw), C.byref(w), _p(rule), _l(N), _p(out))
    return out


def naive_ptswts(Nh, Nw, w, npix=64):
    out = np.zeros((Nh, Nw, 4 * npix * npix))
    lib().or_naive_ptswts(_l(Nh), _l(Nw), C.byref(w), _l(npix), _p(out))
    return out


def cubature_convolve_map_psf(Nh, Nw, src_sph, lut, w, ptswts, searchsep=False):
    """Raw cubature map [S, H, W, E] with precomputed points [H, W, 4N]."""
    src, lut, pts = _c(src_sph), _c(lut), _c(ptswts)
    S, E = src.shape[0], lut.shape[2]
    N = pts.shape[2] // 4
    out = np.zeros((S, Nh, Nw, E))
    lib().or_cubature_convolve_map_psf(_l(Nh), _l(Nw), _p(src), _l(S), _p(lut), _l(E),
                                       C.byref(w), _p(pts), _l(N),
                                       C.c_int(1 if searchsep else 0), _p(out))
    return out


def solid_angle(Nh, Nw, w):
    out = np.zeros((Nh, Nw))
    lib().or_solid_angle(_l(Nh), _l(Nw), C.byref(w), _p(out))
    return out


def psf_boundary_radius(Nh, Nw, src_sph, w):
    src = _c(src_sph)
    S = src.shape[0]
    rad, fov = np.zeros(S), np.zeros(S, np.uint8)
    lib().or_psf_boundary_radius(_l(Nh), _l(Nw), _p(src), _l(S), C.byref(w),
                                 _p(rad), _p(fov, _u8))
    return rad, fov


def map_correction_factor(invmi, radius, in_fov, upsf, parint, exact=False):
    """As coded (exact=False) or the FermiTools-consistent exact-radius variant."""
    invmi, radius, upsf, parint = map(_c, (invmi, radius, upsf, parint))
    fov = _c(in_fov, np.uint8)
    S, D, E = upsf.shape
    out = np.zeros((S, E))
    lib().or_map_correction_factor_mode(_p(invmi), _p(radius), _p(fov, _u8), _p(upsf),
                                        _p(parint), _l(S), _l(D), _l(E),
                                        C.c_int(1 if exact else 0), _p(out))
    return out


def psf_radius_integral(radius, upsf, parint, exact=False):
    """PSF integral [S, E] to each source's radius (uncapped)."""
    radius, upsf, parint = map(_c, (radius, upsf, parint))
    S, D, E = upsf.shape
    out = np.zeros((S, E))
    lib().or_psf_radius_integral(_p(radius), _p(upsf), _p(parint), _l(S), _l(D), _l(E),
                                 C.c_int(1 if exact else 0), _p(out))
    return out


def point_src_model_map_wcs(Nh, Nw, w, src_sph, expo, upsf, parint, rule,
                            details=False):
    """Full model map [S, H, W, E] (reference layout) in FP64."""
    src, expo, upsf, parint, rule = map(_c, (src_sph, expo, upsf, parint, rule))
    S, E = expo.shape
    N = rule.shape[0]
    out = np.zeros((S, Nh, Nw, E))
    rad, fov = np.zeros(S), np.zeros(S, np.uint8)
    invmi, cf = np.zeros((S, E)), np.zeros((S, E))
    lib().or_point_src_model_map_wcs(_l(Nh), _l(Nw), C.byref(w), _p(src), _l(S),
                                     _p(expo), _p(upsf), _p(parint), _l(E), _p(rule),
                                     _l(N), _p(out), _p(rad), _p(fov, _u8), _p(invmi),
                                     _p(cf))
    if details:
        return out, dict(radius=rad, in_fov=fov, inv_mapinteg=invmi, cf=cf)
    return out


def num_threads() -> int:
    return lib().or_num_threads()


# ---------------------------------------------------------------- diffuse sources
class DiffuseGeom(C.Structure):
    _fields_ = [("a1", C.c_long), ("a2", C.c_long), ("n1", C.c_long), ("n2", C.c_long),
                ("m1", C.c_long), ("m2", C.c_long), ("step1", C.c_double),
                ("step2", C.c_double), ("grid", Wcs), ("wrap", C.c_int), ("frame", C.c_int),
                ("nk", C.c_long), ("kc1", C.c_double), ("kc2", C.c_double),
                ("fft_p", C.c_long), ("fft_q", C.c_long), ("period", C.c_double),
                ("vlo1", C.c_double), ("vhi1", C.c_double), ("vlo2", C.c_double),
                ("vhi2", C.c_double)]


def cel2gal(ra, dec):
    """J2000 celestial -> galactic (degrees), or_cel2gal."""
    l, b = C.c_double(), C.c_double()
    lib().or_cel2gal(C.c_double(ra), C.c_double(dec), C.byref(l), C.byref(b))
    return l.value, b.value


def gal2cel(l, b):
    ra, dec = C.c_double(), C.c_double()
    lib().or_gal2cel(C.c_double(l), C.c_double(b), C.byref(ra), C.byref(dec))
    return ra.value, dec.value


def diffuse_geometry(roi, roi_gal, cube, cube_gal, target, margin=0.0) -> DiffuseGeom:
    """or_diffuse_geometry: crop box, resampled grid, kernel and transform sizes."""
    g = DiffuseGeom()
    rc = lib().or_diffuse_geometry(C.byref(roi), C.c_int(int(roi_gal)), C.byref(cube),
                                   C.c_int(int(cube_gal)), C.c_double(target),
                                   C.c_double(margin), C.byref(g))
    if rc:
        raise ValueError(f"or_diffuse_geometry failed ({rc})")
    return g


def diffuse_resample(cube_data, cube, g, E):
    """or_diffuse_resample (math/trilerp.hpp:9-56 over the crop): [E, m1, m2]."""
    d = _c(cube_data, np.float32)
    out = np.zeros((E, g.m1, g.m2))
    lib().or_diffuse_resample(_p(d, _f), _l(cube.naxis1), _l(cube.naxis2), _l(d.shape[0]),
                              C.byref(g), _l(E), _p(out))
    return out


def psf_make_kernel(nk, grid, kc1, kc2, upsf):
    """or_psf_make_kernel (psf/psf.make_kernel.cpp:8-90): [E, nk, nk]."""
    u = _c(upsf)
    E = u.shape[1]
    out = np.zeros((E, nk, nk))
    lib().or_psf_make_kernel(_l(nk), C.byref(grid), C.c_double(kc1), C.c_double(kc2), _p(u),
                             _l(E), _p(out))
    return out


def convolve_same(img, ker, c=None):
    """or_convolve_same: direct convolution, rows/cols c .. c + m - 1 of the full one."""
    im, k = _c(img), _c(ker)
    E, m1, m2 = im.shape
    nk = k.shape[1]
    c = (nk - 1) // 2 if c is None else c
    out = np.zeros_like(im)
    lib().or_convolve_same(_p(im), _l(E), _l(m1), _l(m2), _p(k), _l(nk), _l(c), _p(out))
    return out


def diffuse_sample(conv, g, roi, omega, expo):
    """or_diffuse_sample: bilinear samples x solid angle x exposure, [E, H, W]."""
    cv, om, ex = _c(conv), _c(omega), _c(expo)
    E = ex.shape[0]
    out = np.zeros((E, roi.naxis1, roi.naxis2))
    lib().or_diffuse_sample(_p(cv), C.byref(g), C.byref(roi), _p(om), _p(ex), _l(E), _p(out))
    return out


def diffuse_src_model_map_wcs(roi, roi_gal, cube, cube_gal, cube_data, upsf, expo, target,
                              margin=0.0):
    """or_diffuse_src_model_map_wcs -> [E, roi.naxis1, roi.naxis2]."""
    d, u, ex = _c(cube_data, np.float32), _c(upsf), _c(expo)
    E = ex.shape[0]
    out = np.zeros((E, roi.naxis1, roi.naxis2))
    rc = lib().or_diffuse_src_model_map_wcs(
        C.byref(roi), C.c_int(int(roi_gal)), C.byref(cube), C.c_int(int(cube_gal)), _p(d, _f),
        _l(d.shape[0]), _p(u), _p(ex), _l(E), C.c_double(target), C.c_double(margin), _p(out))
    if rc:
        raise ValueError(f"or_diffuse_src_model_map_wcs failed ({rc})")
    return out
