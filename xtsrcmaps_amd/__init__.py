"""xtsrcmaps_amd -- MI355X (gfx950) point-source model-map path.

Python host mirror of the reference's C++ interface for this path
(``Exposure::``, ``Psf::``, ``ModelMap::`` in Areustle/xtsrcmaps), calling the
C ABI of ``libxtsrcmaps_amd.so`` (declared in ``include/xtsrcmaps_amd.h``).
Arrays are numpy, row-major, in the reference's layouts.  Every call runs on
the GPU; there is no CPU fallback and a missing library raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import capi
from .capi import (XtsError, Wcs, PROJ_CAR, PROJ_AIT, LAYOUT_SHWE, LAYOUT_SEHW,
                   OUT_F64, OUT_F32, COMPUTE_F32, COMPUTE_F64, CF_EXACT_RADIUS, SURROGATE,
                   SURROGATE_NO_NET, SURROGATE_FP8, PREC_FP64, PREC_BF16X3)

ND = 401

__all__ = [
    "Context", "Wcs", "XtsError", "separations", "costhetas",
    "livetime_efficiency_factors", "PROJ_CAR", "PROJ_AIT", "LAYOUT_SHWE",
    "LAYOUT_SEHW", "OUT_F64", "OUT_F32", "COMPUTE_F32", "COMPUTE_F64", "CF_EXACT_RADIUS",
    "SURROGATE", "SURROGATE_NO_NET", "SURROGATE_FP8",
    "PREC_FP64", "PREC_BF16X3", "ND", "write_src_model", "read_irf_pars", "read_expcube",
]


def separations() -> np.ndarray:
    """Psf::separations (psf/psf.hxx:24-31): 0, then 1e-4 .. 70 deg log-spaced."""
    x = np.log(70.0 / 1e-4) / 399.0
    out = np.zeros(ND)
    out[1:] = 1e-4 * np.exp(np.arange(400) * x)
    return out


def costhetas(nbins: int, cosmin: float, thetabin: bool = True) -> np.ndarray:
    """Exposure::costhetas (exposure/exposure.cxx:29-40)."""
    out = np.zeros(nbins)
    capi.check(capi.lib().xts_costhetas(nbins, float(cosmin), int(bool(thetabin)),
                                        capi.dptr(out)))
    return out


def livetime_efficiency_factors(logEs, p0, p1) -> np.ndarray:
    """Irf::livetime_efficiency_factors (irf/irf.cxx:70-109) -> [2, E]."""
    logEs = capi.f64(logEs)
    p0 = np.ascontiguousarray(p0, np.float32)
    p1 = np.ascontiguousarray(p1, np.float32)
    out = np.zeros((2, len(logEs)))
    capi.check(capi.lib().xts_livetime_efficiency_factors(
        capi.dptr(logEs), len(logEs), capi.fptr(p0), capi.fptr(p1), capi.dptr(out)))
    return out


def read_mapcube(path):
    """The primary image of a MapCube FITS file (xts_read_mapcube): (wcs,
    data float32 [planes, naxis2, naxis1], energies float64 or None)."""
    w, planes, ne = Wcs(), C.c_int64(), C.c_int64()
    p = str(path).encode()
    capi.check(capi.lib().xts_read_mapcube(p, C.byref(w), C.byref(planes), None, None,
                                           C.byref(ne)))
    data = np.zeros((planes.value, w.naxis2, w.naxis1), np.float32)
    en = np.zeros(max(ne.value, 1))
    capi.check(capi.lib().xts_read_mapcube(p, C.byref(w), C.byref(planes), capi.fptr(data),
                                           capi.dptr(en), C.byref(ne)))
    return w, data, (en[:ne.value] if ne.value else None)


def diffuse_geometry(roi: Wcs, cube_wcs: Wcs, target_pix_size: float,
                     margin_deg: float = 0.0) -> dict:
    """The crop, resampled grid, kernel and transform sizes of a diffuse map
    (xts_diffuse_geometry; host only)."""
    g = capi.DiffuseGeom()
    capi.check(capi.lib().xts_diffuse_geometry(C.byref(roi), C.byref(cube_wcs),
                                               float(target_pix_size), float(margin_deg),
                                               C.byref(g)))
    d = {n: getattr(g, n) for n, _ in capi.DiffuseGeom._fields_ if n != "grid"}
    gw = Wcs()
    C.memmove(C.byref(gw), C.byref(g.grid), C.sizeof(Wcs))
    d["grid"] = gw
    return d


def write_src_model(path, model_map, src_names, layout: int = LAYOUT_SHWE,
                    float32: bool = False, wcs: Wcs | None = None) -> None:
    """fits::write_src_model (fits/fits.write_src_model.cxx:10-66): one IMAGE
    HDU per source, EXTNAME = its name, axes (H, W, E) in FITS order, with
    CHECKSUM/DATASUM.  model_map: [S, H, W, E] (LAYOUT_SHWE) or [S, E, H, W]
    (LAYOUT_SEHW); float32 writes BITPIX -32.  Host only."""
    m = capi.f64(model_map)
    if m.ndim != 4:
        raise ValueError("model_map must have 4 axes")
    S = m.shape[0]
    if layout == LAYOUT_SEHW:
        E, H, W = m.shape[1:]
    else:
        H, W, E = m.shape[1:]
    if len(src_names) != S:
        raise ValueError(f"{len(src_names)} names for {S} sources")
    names = (C.c_char_p * S)(*[str(n).encode() for n in src_names])
    flags = int(layout) | (OUT_F32 if float32 else 0)
    capi.check(capi.lib().xts_write_src_model(os.fsencode(path), capi.dptr(m), S, H, W, E,
                                              names, flags,
                                              C.byref(wcs) if wcs is not None else None))


def read_irf_pars(path, table: str):
    """fits::read_irf_pars (fits/fits.read_irf_pars.cxx:12-126) -> (extents,
    offsets, rowdata [nrows, row width] float32).  Host only."""
    L = capi.lib()
    nc, nr, rw = C.c_int64(), C.c_int64(), C.c_int64()
    capi.check(L.xts_read_irf_pars(os.fsencode(path), table.encode(), C.byref(nc), None, 0,
                                   C.byref(nr), C.byref(rw), None))
    ext = (C.c_int64 * max(nc.value, 1))()
    rows = np.zeros((nr.value, rw.value), np.float32)
    capi.check(L.xts_read_irf_pars(os.fsencode(path), table.encode(), C.byref(nc), ext,
                                   nc.value, C.byref(nr), C.byref(rw),
                                   rows.ctypes.data_as(C.POINTER(C.c_float))))
    extents = np.array(ext[:nc.value], np.int64)
    offsets = np.concatenate([[0], np.cumsum(extents)[:-1]]).astype(np.int64)
    return extents, offsets, rows


def read_expcube(path, table: str = "EXPOSURE") -> dict:
    """fits::read_expcube (fits/fits.read_expcube.cxx:8-116): the keywords and
    cosbins [12 nside^2, nbrbins], ra, dec [12 nside^2] as float32."""
    L = capi.lib()
    info = capi.ExpcubeInfo()
    capi.check(L.xts_read_expcube(os.fsencode(path), table.encode(), C.byref(info), None, None,
                                  None))
    npix = 12 * info.nside * info.nside
    cos = np.zeros((npix, info.nbrbins), np.float32)
    ra = np.zeros(npix, np.float32)
    dec = np.zeros(npix, np.float32)
    fp = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))
    capi.check(L.xts_read_expcube(os.fsencode(path), table.encode(), C.byref(info), fp(cos),
                                  fp(ra), fp(dec)))
    return dict(cosbins=cos, ra=ra, dec=dec, nside=info.nside, nbrbins=info.nbrbins,
                cosmin=info.cosmin, thetabin=bool(info.thetabin),
                ordering=info.ordering.decode(), coordsys=info.coordsys.decode())


class Context:
    """One device context (stream + device arena) of the C ABI."""

    def __init__(self, device: int = 0):
        self._h = C.c_void_p()
        capi.check(capi.lib().xts_ctx_create(int(device), C.byref(self._h)))
        self.device = device

    def close(self):
        if self._h:
            capi.lib().xts_ctx_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def handle(self):
        return self._h

    def set_stream(self, stream_ptr: int | None):
        capi.check(capi.lib().xts_ctx_set_stream(self._h, C.c_void_p(stream_ptr or 0)))

    def synchronize(self):
        capi.check(capi.lib().xts_ctx_synchronize(self._h))

    def get_tuning(self, key: str) -> int:
        """The last batched model map's shape: "batch_sources", "rowweight_blocks", "batches"."""
        v = C.c_int64()
        capi.check(capi.lib().xts_ctx_get_tuning(self._h, key.encode(), C.byref(v)))
        return int(v.value)

    def set_tuning(self, key: str, value: int):
        """Launch-shape override ("tiles_per_block"; 0 = default)."""
        capi.check(capi.lib().xts_ctx_set_tuning(self._h, key.encode(), int(value)))

    def enable_timing(self, on: bool = True):
        capi.check(capi.lib().xts_ctx_enable_timing(self._h, int(bool(on))))

    def kernel_times(self, name: str):
        """(total ms, launches) of a timed stage since the last query."""
        ms, n = C.c_double(), C.c_int64()
        capi.check(capi.lib().xts_ctx_kernel_times(self._h, name.encode(), C.byref(ms),
                                                   C.byref(n)))
        return ms.value, n.value

    # ------------------------------------------------------------ (a) exposure
    def aeff_value(self, costh, logE, grid_costh, grid_logE, params, min_costheta):
        """Irf::aeff_value (irf/irf.aeff_value.cxx:11-55) -> [C, E] cm^2."""
        costh, logE = capi.f64(costh), capi.f64(logE)
        g = capi.aeff_grid(grid_costh, grid_logE, params, min_costheta)
        out = np.zeros((len(costh), len(logE)))
        capi.check(capi.lib().xts_aeff_value(self._h, capi.dptr(costh), len(costh),
                                             capi.dptr(logE), len(logE), C.byref(g[0]),
                                             capi.dptr(out)))
        return out

    def src_cosbins(self, src_sph, cube, nside):
        """Exposure::src_cosbins (exposure/exposure.cxx:42-63) -> [S, C]."""
        src, cube = capi.f64(src_sph), capi.f64(cube)
        S, Cn = src.shape[0], cube.shape[1]
        out = np.zeros((S, Cn))
        capi.check(capi.lib().xts_src_cosbins(self._h, capi.dptr(src), S, capi.dptr(cube),
                                              int(nside), Cn, capi.dptr(out)))
        return out

    def contract(self, A, B, precision=PREC_FP64):
        """Exposure::contract (exposure/exposure.contract.cpp:13-42)."""
        A, B = capi.f64(A), capi.f64(B)
        out = np.zeros((A.shape[0], B.shape[1]))
        capi.check(capi.lib().xts_contract(self._h, capi.dptr(A), capi.dptr(B), A.shape[0],
                                           A.shape[1], B.shape[1], capi.dptr(out),
                                           int(precision)))
        return out

    def exposure(self, ecb, wecb, aeff_front, aeff_back, front_ltf, precision=PREC_FP64):
        """Exposure::exposure (exposure/exposure.cxx:69-173) -> [S, E]."""
        ecb, wecb, af, ab, ltf = map(capi.f64, (ecb, wecb, aeff_front, aeff_back, front_ltf))
        S, Cn = ecb.shape
        E = af.shape[1]
        out = np.zeros((S, E))
        capi.check(capi.lib().xts_exposure(self._h, capi.dptr(ecb), capi.dptr(wecb),
                                           capi.dptr(af), capi.dptr(ab), capi.dptr(ltf), S, Cn,
                                           E, capi.dptr(out), int(precision)))
        return out

    # ------------------------------------------------------------ (b) mean PSF
    def normalize_rpsf(self, params, logEs, scale3):
        """irf_private::normalize_rpsf (irf/irf.normalize_rpsf.cpp:11-61)."""
        p = capi.f64(params).copy()
        logEs = capi.f64(logEs)
        Mc, Me, _ = p.shape
        capi.check(capi.lib().xts_normalize_rpsf(self._h, capi.dptr(p), capi.dptr(logEs), Mc, Me,
                                                 C.c_float(scale3[0]), C.c_float(scale3[1]),
                                                 C.c_float(scale3[2])))
        return p

    def king(self, params, cosths, logEs):
        """Psf::king (psf/psf.king.cxx:9-43) -> [401, Mc, Me]."""
        g = capi.psf_grid(params, cosths, logEs)
        Mc, Me = g[0].Mc, g[0].Me
        out = np.zeros((ND, Mc, Me))
        capi.check(capi.lib().xts_king(self._h, C.byref(g[0]), capi.dptr(out)))
        return out

    def psf_lookup_table_and_partial_integrals(self, front, back, costh, logE, aeff_front,
                                               aeff_back, ecb, wecb, front_ltf, exposure,
                                               precision=PREC_FP64):
        """Psf::psf_lookup_table_and_partial_integrals -> (uPsf, partial) [S, 401, E].

        front/back are (params [Mc, Me, 6] normalised, cosths [Mc], logEs [Me])."""
        gf, gb = capi.psf_grid(*front), capi.psf_grid(*back)
        costh, logE, af, ab, ecb, wecb, ltf, expo = map(
            capi.f64, (costh, logE, aeff_front, aeff_back, ecb, wecb, front_ltf, exposure))
        S, E = expo.shape
        upsf = np.zeros((S, ND, E))
        parint = np.zeros((S, ND, E))
        capi.check(capi.lib().xts_psf_lookup_table_and_partial_integrals(
            self._h, C.byref(gf[0]), C.byref(gb[0]), capi.dptr(costh), len(costh),
            capi.dptr(logE), E, capi.dptr(af), capi.dptr(ab), capi.dptr(ecb), capi.dptr(wecb),
            capi.dptr(ltf), capi.dptr(expo), S, capi.dptr(upsf), capi.dptr(parint),
            int(precision)))
        return upsf, parint

    def bilerp(self, costh, logE, cosths, logEs, kings):
        """Psf::bilerp (psf/psf.bilerp.cxx:6-52): kings [401, Mc, Me] -> [C, 401, E]."""
        costh, logE, k = capi.f64(costh), capi.f64(logE), capi.f64(kings)
        g = capi.psf_grid(np.zeros((len(cosths), len(logEs), 6)), cosths, logEs)
        out = np.zeros((len(costh), ND, len(logE)))
        capi.check(capi.lib().xts_bilerp(self._h, capi.dptr(costh), len(costh), capi.dptr(logE),
                                         len(logE), C.byref(g[0]), capi.dptr(k),
                                         capi.dptr(out)))
        return out

    def corrected_exposure_psf(self, obs_psf, obs_aeff, ecb, wecb, front_ltf):
        """Psf::corrected_exposure_psf (psf/psf.corrected_exposure_psf.cxx:12-120)."""
        obs_psf, obs_aeff, ecb, wecb, ltf = map(capi.f64, (obs_psf, obs_aeff, ecb, wecb,
                                                           front_ltf))
        Cn, _, E = obs_psf.shape
        S = ecb.shape[0]
        out = np.zeros((S, ND, E))
        capi.check(capi.lib().xts_corrected_exposure_psf(
            self._h, capi.dptr(obs_psf), capi.dptr(obs_aeff), capi.dptr(ecb), capi.dptr(wecb),
            capi.dptr(ltf), S, Cn, E, capi.dptr(out)))
        return out

    def mean_psf(self, front, back, exposure):
        """Psf::mean_psf (psf/psf.mean_psf.cxx:6-33)."""
        front, back, expo = map(capi.f64, (front, back, exposure))
        S, _, E = front.shape
        out = np.zeros_like(front)
        capi.check(capi.lib().xts_mean_psf(self._h, capi.dptr(front), capi.dptr(back),
                                           capi.dptr(expo), S, E, capi.dptr(out)))
        return out

    def normalize(self, upsf, total):
        """Psf::normalize (psf/psf.normalize.cxx:7-24); returns the normalised table."""
        u = capi.f64(upsf).copy()
        t = capi.f64(total)
        S, _, E = u.shape
        capi.check(capi.lib().xts_normalize(self._h, capi.dptr(u), capi.dptr(t), S, E))
        return u

    def partial_total_integral(self, upsf):
        """Psf::partial_total_integral + Psf::normalize -> (uPsf_normed, partial, total)."""
        u = capi.f64(upsf).copy()
        S, _, E = u.shape
        par, tot = np.zeros_like(u), np.zeros((S, E))
        capi.check(capi.lib().xts_partial_total_integral(self._h, capi.dptr(u), S, E,
                                                         capi.dptr(par), capi.dptr(tot)))
        return u, par, tot

    # ------------------------------------------------------------ (c) model map
    def point_src_model_map_wcs(self, wcs: Wcs, src_sph, exposure, upsf, parint, rule=None,
                                flags=LAYOUT_SHWE | OUT_F64 | COMPUTE_F32):
        """ModelMap::point_src_model_map_wcs -> [S, H, W, E] (or [S, E, H, W])."""
        src, expo, upsf, parint = map(capi.f64, (src_sph, exposure, upsf, parint))
        S, E = expo.shape
        out = capi.alloc_map(S, wcs, E, flags)
        rp, N = capi.rule_ptr(rule)
        capi.check(capi.lib().xts_point_src_model_map_wcs(
            self._h, C.byref(wcs), capi.dptr(src), S, capi.dptr(expo), capi.dptr(upsf),
            capi.dptr(parint), E, rp, N, int(flags), out.ctypes.data_as(C.c_void_p)))
        return out

    def cubature_convolve_map_psf(self, wcs: Wcs, src_sph, upsf, rule=None,
                                  flags=LAYOUT_SHWE | OUT_F64 | COMPUTE_F32):
        """cubature_convolve_map_psf (cubature.implementation.cxx:21-55)."""
        src, upsf = capi.f64(src_sph), capi.f64(upsf)
        S, _, E = upsf.shape
        out = capi.alloc_map(S, wcs, E, flags)
        rp, N = capi.rule_ptr(rule)
        capi.check(capi.lib().xts_cubature_convolve_map_psf(
            self._h, C.byref(wcs), capi.dptr(src), S, capi.dptr(upsf), E, rp, N, int(flags),
            out.ctypes.data_as(C.c_void_p)))
        return out

    def solid_angle(self, wcs: Wcs):
        """Math::solid_angle (math/solid_angle.hpp:42-79) -> [H, W] sr."""
        out = np.zeros((wcs.naxis1, wcs.naxis2))
        capi.check(capi.lib().xts_solid_angle(self._h, C.byref(wcs), capi.dptr(out)))
        return out

    # ------------------------------------------------------ diffuse sources
    def diffuse_src_model_map_wcs(self, roi: Wcs, cube_wcs: Wcs, cube, upsf, exposure,
                                  target_pix_size, margin_deg=0.0):
        """ModelMap::diffuse_src_model_map_wcs
        (model_map/model_map.diffuse_src_model_map_wcs.cxx:8-97) on host
        arrays: cube [planes, cube.naxis2, cube.naxis1] float32 (FITS order),
        upsf [401, E], exposure [E] -> [E, roi.naxis1, roi.naxis2] float64."""
        cb = np.ascontiguousarray(cube, dtype=np.float32)
        u, ex = capi.f64(upsf), capi.f64(exposure)
        if cb.ndim != 3 or cb.shape[1:] != (cube_wcs.naxis2, cube_wcs.naxis1):
            raise capi.XtsError(1, f"cube shape {cb.shape} does not match its WCS")
        E = ex.shape[0]
        if u.shape != (401, E):
            raise capi.XtsError(1, f"upsf shape {u.shape}, expected (401, {E})")
        out = np.zeros((E, roi.naxis1, roi.naxis2))
        capi.check(capi.lib().xts_diffuse_src_model_map_wcs(
            self._h, C.byref(roi), C.byref(cube_wcs), capi.fptr(cb), cb.shape[0], capi.dptr(u),
            capi.dptr(ex), E, float(target_pix_size), float(margin_deg), capi.dptr(out)))
        return out

    def diffuse_src_model_map_wcs_dev(self, roi: Wcs, cube_wcs: Wcs, cube, upsf, exposure,
                                      target_pix_size, out, margin_deg=0.0, flags=0):
        """The diffuse model map on torch CUDA tensors (async on the context
        stream): out [E, roi.naxis1, roi.naxis2] float64, or float32 with
        OUT_F32."""
        E = exposure.shape[0]
        dt = "float32" if flags & OUT_F32 else "float64"
        p = capi.devs([("cube", cube, "float32", (cube.shape[0], cube_wcs.naxis2, cube_wcs.naxis1)),
                       ("upsf", upsf, "float64", (401, E)), ("exposure", exposure, "float64", (E,)),
                       ("out", out, dt, (E, roi.naxis1, roi.naxis2))])
        capi.check(capi.lib().xts_diffuse_src_model_map_wcs_dev(
            self._h, C.byref(roi), C.byref(cube_wcs), p[0], cube.shape[0], p[1], p[2], E,
            float(target_pix_size), float(margin_deg), p[3], int(flags)))

    def psf_make_kernel(self, nk, grid: Wcs, kc1, kc2, upsf):
        """Psf::make_kernel (psf/psf.make_kernel.cpp:8-90): upsf [401, E] ->
        [E, nk, nk], centred at grid pixel (kc1, kc2), sum 1 per energy."""
        u = capi.f64(upsf)
        E = u.shape[1]
        out = np.zeros((E, nk, nk))
        capi.check(capi.lib().xts_psf_make_kernel(self._h, int(nk), C.byref(grid), float(kc1),
                                                  float(kc2), capi.dptr(u), E, capi.dptr(out)))
        return out

    def fft_convolve_map_psf(self, image, kernel):
        """FFT::convolve_map_psf (model_map/mm_fft/fft.convolve_map_psf.cpp:8-39):
        image [C, H, W], kernel [C, KH, KW] -> [C, H + KH - 1, W + KW - 1]."""
        im, k = capi.f64(image), capi.f64(kernel)
        Cn, H, W = im.shape
        _, KH, KW = k.shape
        out = np.zeros((Cn, H + KH - 1, W + KW - 1))
        capi.check(capi.lib().xts_fft_convolve_map_psf(self._h, capi.dptr(im), Cn, H, W,
                                                       capi.dptr(k), KH, KW, capi.dptr(out)))
        return out

    def psf_boundary_radius(self, wcs: Wcs, src_sph):
        """ModelMap::psf_boundary_radius -> (radius [S] deg, in_fov [S] bool)."""
        src = capi.f64(src_sph)
        S = src.shape[0]
        rad, fov = np.zeros(S), np.zeros(S, np.uint8)
        capi.check(capi.lib().xts_psf_boundary_radius(self._h, C.byref(wcs), capi.dptr(src), S,
                                                      capi.dptr(rad), capi.u8ptr(fov)))
        return rad, fov.astype(bool)

    def map_integral(self, wcs: Wcs, model_map, src_sph, radius, in_fov):
        """ModelMap::map_integral -> inverse map integral [S, E]."""
        m, src, rad = capi.f64(model_map), capi.f64(src_sph), capi.f64(radius)
        fov = np.ascontiguousarray(in_fov, np.uint8)
        S, E = m.shape[0], m.shape[3]
        out = np.zeros((S, E))
        capi.check(capi.lib().xts_map_integral(self._h, C.byref(wcs), capi.dptr(m),
                                               capi.dptr(src), S, E, capi.dptr(rad),
                                               capi.u8ptr(fov), capi.dptr(out)))
        return out

    def map_correction_factor(self, inv_mapinteg, radius, in_fov, upsf, parint):
        """ModelMap::map_correction_factor -> [S, E]."""
        inv, rad, upsf, parint = map(capi.f64, (inv_mapinteg, radius, upsf, parint))
        fov = np.ascontiguousarray(in_fov, np.uint8)
        S, _, E = upsf.shape
        out = np.zeros((S, E))
        capi.check(capi.lib().xts_map_correction_factor(self._h, capi.dptr(inv), capi.dptr(rad),
                                                        capi.u8ptr(fov), capi.dptr(upsf),
                                                        capi.dptr(parint), S, E, capi.dptr(out)))
        return out

    def scale_map_by_solid_angle(self, model_map, omega):
        """ModelMap::scale_map_by_solid_angle, in place on a double [S,H,W,E] map."""
        return self._scale(model_map, omega, capi.SCALE_PIXEL)

    def scale_map_by_exposure(self, model_map, exposure):
        """ModelMap::scale_map_by_exposure, in place."""
        return self._scale(model_map, exposure, capi.SCALE_SRC_ENERGY)

    def scale_map_by_correction_factors(self, model_map, factor):
        """ModelMap::scale_map_by_correction_factors, in place."""
        return self._scale(model_map, factor, capi.SCALE_SRC_ENERGY)

    def _scale(self, m, f, kind):
        if not (isinstance(m, np.ndarray) and m.dtype == np.float64 and m.flags.c_contiguous):
            raise XtsError(1, "model map must be a C-contiguous float64 array")
        f = capi.f64(f)
        S, H, W, E = m.shape
        capi.check(capi.lib().xts_scale_map(self._h, capi.dptr(m), S, H, W, E, capi.dptr(f),
                                            int(kind)))
        return m

    def pix2sph(self, wcs: Wcs, pix):
        pix = capi.f64(pix).reshape(-1, 2)
        out = np.zeros_like(pix)
        capi.check(capi.lib().xts_pix2sph(self._h, C.byref(wcs), capi.dptr(pix), pix.shape[0],
                                          capi.dptr(out)))
        return out

    def sph2pix(self, wcs: Wcs, sph):
        sph = capi.f64(sph).reshape(-1, 2)
        out = np.zeros_like(sph)
        capi.check(capi.lib().xts_sph2pix(self._h, C.byref(wcs), capi.dptr(sph), sph.shape[0],
                                          capi.dptr(out)))
        return out

    # ---------------------------------------------------- batched model maps
    def model_map_footprints(self, wcs_list, src_begin, src_sph, parint, containment=0.995):
        """Footprint boxes of every source (xts_model_map_footprints): returns
        (boxes [S] capi.Footprint array, radius [S] deg, total output elements)."""
        n = len(wcs_list)
        begin = np.ascontiguousarray(src_begin, np.int64)
        S = int(begin[-1])
        E = parint.shape[2]
        p = capi.devs([("src_sph", src_sph, "float64", (S, 2)),
                       ("parint", parint, "float64", (S, 401, E))])
        W = (capi.Wcs * n)(*wcs_list)
        boxes = (capi.Footprint * S)()
        radius = np.zeros(S)
        total = C.c_int64()
        capi.check(capi.lib().xts_model_map_footprints(
            self._h, n, W, begin.ctypes.data_as(C.POINTER(C.c_int64)), p[0], p[1], E,
            float(containment), boxes, capi.dptr(radius), C.byref(total)))
        return boxes, radius, int(total.value)

    def point_src_model_maps_batched_dev(self, wcs_list, src_begin, src_sph, exposure, upsf,
                                         parint, boxes, out, rule=None,
                                         flags=LAYOUT_SEHW | OUT_F32 | COMPUTE_F32):
        """Batched model maps of several images into the footprint blocks of
        out (float32 device tensor of the footprints' total size)."""
        n = len(wcs_list)
        begin = np.ascontiguousarray(src_begin, np.int64)
        S, E = exposure.shape
        if int(begin[-1]) != S or len(boxes) != S:
            raise capi.XtsError(1, "src_begin, boxes and the source arrays disagree")
        p = capi.devs([("src_sph", src_sph, "float64", (S, 2)),
                       ("exposure", exposure, "float64", (S, E)),
                       ("upsf", upsf, "float64", (S, 401, E)),
                       ("parint", parint, "float64", (S, 401, E)),
                       ("out", out, "float32", None, out.numel())])
        W = (capi.Wcs * n)(*wcs_list)
        rp, N = capi.rule_ptr(rule)
        capi.check(capi.lib().xts_point_src_model_maps_batched_dev(
            self._h, n, W, begin.ctypes.data_as(C.POINTER(C.c_int64)), p[0], p[1], p[2], p[3],
            E, rp, N, int(flags), boxes, p[4], out.numel()))

    # ------------------------------------------------------- learned surrogate
    def surrogate_samples_dev(self, wcs: Wcs, rule_dev, seed, n_candidates, out, count):
        """Training samples of the image (xts_surrogate_samples_dev): rows of
        8 floats into out [capacity, 8]; count: device int64 [1], zero before
        the first call, advanced by the number of samples found."""
        N = rule_dev.shape[0]
        p = capi.devs([("rule", rule_dev, "float64", (N, 3)),
                       ("out", out, "float32", (out.shape[0], 8)),
                       ("count", count, "int64", (1,))])
        capi.check(capi.lib().xts_surrogate_samples_dev(
            self._h, C.byref(wcs), C.cast(p[0], capi._d), N, int(seed) & (2**64 - 1),
            int(n_candidates), out.shape[0], p[1], p[2]))

    @staticmethod
    def _mlp_specs(params, norm, x):
        return [("params", params, "float32", (capi.MLP_PARAMS,)),
                ("norm", norm, "float32", (capi.MLP_NORM,)),
                ("x", x, "float32", (x.shape[0], x.shape[1]))]

    def mlp_forward_dev(self, params, norm, x, y, bf16=False, fp8=False):
        """FP32 forward (or BF16 / FP8 matrix-core inference) of rows x [B, >=4] -> y [B]."""
        if x.dim() != 2 or x.shape[1] < 4:
            raise capi.XtsError(1, "x: expected [B, >= 4]")
        B = x.shape[0]
        p = capi.devs(self._mlp_specs(params, norm, x) + [("y", y, "float32", (B,))])
        fn = (capi.lib().xts_mlp_infer_fp8_dev if fp8 else
              capi.lib().xts_mlp_infer_bf16_dev if bf16 else capi.lib().xts_mlp_forward_dev)
        capi.check(fn(self._h, p[0], p[1], p[2], x.shape[1], B, p[3]))

    def mlp_grad_dev(self, params, norm, x, y_target, grad, loss):
        """Mean-squared-error gradient over rows x [B, >=4] with targets
        y_target [B] -> grad [8705], loss [1]."""
        if x.dim() != 2 or x.shape[1] < 4:
            raise capi.XtsError(1, "x: expected [B, >= 4]")
        B = x.shape[0]
        p = capi.devs(self._mlp_specs(params, norm, x) +
                      [("y_target", y_target, "float32", (B,)),
                       ("grad", grad, "float32", (capi.MLP_PARAMS,)),
                       ("loss", loss, "float32", (1,))])
        capi.check(capi.lib().xts_mlp_grad_dev(self._h, p[0], p[1], p[2], x.shape[1], p[3], 1,
                                               B, p[4], p[5]))

    def mlp_adam_dev(self, params, grad, m, v, step, lr, beta1=0.9, beta2=0.999, eps=1e-8):
        sh = (capi.MLP_PARAMS,)
        p = capi.devs([("params", params, "float32", sh), ("grad", grad, "float32", sh),
                       ("m", m, "float32", sh), ("v", v, "float32", sh)])
        capi.check(capi.lib().xts_mlp_adam_dev(self._h, *p, int(step), float(lr), float(beta1),
                                               float(beta2), float(eps)))

    def surrogate_train_dev(self, params, norm, samples, steps, batch, lr, seed, loss_hist=None):
        """Adam training on the rows of samples [n, 8] (features 0..3, target
        4), in place on params; loss_hist [steps] (optional) per step."""
        n = samples.shape[0]
        specs = [("params", params, "float32", (capi.MLP_PARAMS,)),
                 ("norm", norm, "float32", (capi.MLP_NORM,)),
                 ("samples", samples, "float32", (n, 8))]
        if loss_hist is not None:
            specs.append(("loss_hist", loss_hist, "float32", (int(steps),)))
        p = capi.devs(specs)
        ys = capi._vp(samples.data_ptr() + 16)   # column 4
        capi.check(capi.lib().xts_surrogate_train_dev(
            self._h, p[0], p[1], p[2], 8, ys, 8, n, int(steps), int(batch), float(lr),
            int(seed) & (2**64 - 1), p[3] if loss_hist is not None else None))

    def set_surrogate(self, params=None, norm=None):
        """The network of XTS_SURROGATE model maps (None clears it)."""
        if params is None:
            capi.check(capi.lib().xts_ctx_set_surrogate(self._h, None, None))
            return
        p = capi.devs([("params", params, "float32", (capi.MLP_PARAMS,)),
                       ("norm", norm, "float32", (capi.MLP_NORM,))])
        capi.check(capi.lib().xts_ctx_set_surrogate(self._h, p[0], p[1]))

    # ------------------------------------------------------------ device variants
    def aeff_value_dev(self, costh, logE, grid, out, range_error=None):
        """Device Irf::aeff_value: grid = (cosths [Mc], logEs [Me], params
        [Mc, Me], min_costheta) with device tensors -> out [C, E]."""
        gc, ge, gp, minc = grid
        Cn, E, Mc, Me = costh.shape[0], logE.shape[0], gc.shape[0], ge.shape[0]
        specs = [("costh", costh, "float64", (Cn,)), ("logE", logE, "float64", (E,)),
                 ("aeff cosths", gc, "float64", (Mc,)), ("aeff logEs", ge, "float64", (Me,)),
                 ("aeff params", gp, "float64", (Mc, Me)), ("out", out, "float64", (Cn, E))]
        if range_error is not None:
            specs.append(("range_error", range_error, "int32", (1,)))
        p = capi.devs(specs)
        g = capi.AeffGrid(C.cast(p[4], capi._d), C.cast(p[2], capi._d), C.cast(p[3], capi._d),
                          Mc, Me, float(minc))
        capi.check(capi.lib().xts_aeff_value_dev(
            self._h, C.cast(p[0], capi._d), Cn, C.cast(p[1], capi._d), E, C.byref(g),
            C.cast(p[5], capi._d), p[6] if range_error is not None else None))

    def src_cosbins_dev(self, src_sph, cube, nside, out):
        """Device-pointer Exposure::src_cosbins: out [S, C] from cube
        [12 nside^2, C] (async)."""
        S, Cn = src_sph.shape[0], cube.shape[1]
        p = capi.devs([("src_sph", src_sph, "float64", (S, 2)),
                       ("cube", cube, "float64", (12 * int(nside) ** 2, Cn)),
                       ("out", out, "float64", (S, Cn))])
        capi.check(capi.lib().xts_src_cosbins_dev(self._h, p[0], S, p[1], int(nside), Cn, p[2]))

    def exposure_dev(self, ecb, wecb, af, ab, ltf, out, precision=PREC_FP64):
        """Device-pointer Exposure::exposure on torch CUDA tensors (async)."""
        S, Cn = ecb.shape
        E = af.shape[1]
        p = capi.devs([("ecb", ecb, "float64", (S, Cn)), ("wecb", wecb, "float64", (S, Cn)),
                       ("aeff_front", af, "float64", (Cn, E)),
                       ("aeff_back", ab, "float64", (Cn, E)),
                       ("front_ltf", ltf, "float64", (2, E)), ("exposure", out, "float64", (S, E))])
        capi.check(capi.lib().xts_exposure_dev(self._h, *p[:5], S, Cn, E, p[5], int(precision)))

    def psf_lookup_dev(self, front, back, costh, logE, af, ab, ecb, wecb, ltf, expo, upsf,
                       parint, precision=PREC_FP64):
        """Device-pointer mean-PSF build; front/back are tuples of CUDA tensors."""
        S, E = expo.shape
        Cn = costh.shape[0]
        p = capi.devs(capi.psf_grid_specs(*front) + capi.psf_grid_specs(*back) + [
            ("costh", costh, "float64", (Cn,)), ("logE", logE, "float64", (E,)),
            ("aeff_front", af, "float64", (Cn, E)), ("aeff_back", ab, "float64", (Cn, E)),
            ("ecb", ecb, "float64", (S, Cn)), ("wecb", wecb, "float64", (S, Cn)),
            ("front_ltf", ltf, "float64", (2, E)), ("exposure", expo, "float64", (S, E)),
            ("upsf", upsf, "float64", (S, 401, E)), ("parint", parint, "float64", (S, 401, E))])
        gf = capi.psf_grid_dev(p[0:3], front[1].shape[0], front[2].shape[0])
        gb = capi.psf_grid_dev(p[3:6], back[1].shape[0], back[2].shape[0])
        q = p[6:]
        capi.check(capi.lib().xts_psf_lookup_table_and_partial_integrals_dev(
            self._h, C.byref(gf), C.byref(gb), q[0], Cn, q[1], E, q[2], q[3], q[4], q[5], q[6],
            q[7], S, q[8], q[9], int(precision)))

    @staticmethod
    def _map_specs(wcs: Wcs, src_sph, exposure, upsf, parint, out, flags):
        S, E = exposure.shape
        dt = "float32" if flags & OUT_F32 else "float64"
        return [("src_sph", src_sph, "float64", (S, 2)), ("exposure", exposure, "float64", (S, E)),
                ("upsf", upsf, "float64", (S, 401, E)), ("parint", parint, "float64", (S, 401, E)),
                ("out", out, dt, None, S * E * wcs.naxis1 * wcs.naxis2)]

    def point_src_model_map_wcs_dev(self, wcs: Wcs, src_sph, exposure, upsf, parint, out,
                                    rule=None, flags=LAYOUT_SEHW | OUT_F32 | COMPUTE_F32):
        """Device-pointer model map on torch CUDA tensors (async on the context stream)."""
        S, E = exposure.shape
        p = capi.devs(self._map_specs(wcs, src_sph, exposure, upsf, parint, out, int(flags)))
        rp, N = capi.rule_ptr(rule)
        capi.check(capi.lib().xts_point_src_model_map_wcs_dev(
            self._h, C.byref(wcs), p[0], S, p[1], p[2], p[3], E, rp, N, int(flags), p[4]))

    def point_src_summed_model_dev(self, wcs: Wcs, src_sph, flux, exposure, upsf, parint, out,
                                   rule=None, flags=0):
        """Flux-weighted summed model out [E, H, W] float32 = sum_s flux[s] x
        source s's map, on torch CUDA tensors (async on the context stream).
        flags: SURROGATE (with SURROGATE_NO_NET or SURROGATE_FP8), CF_EXACT_RADIUS."""
        S, E = exposure.shape
        H, W = wcs.naxis1, wcs.naxis2
        p = capi.devs([("src_sph", src_sph, "float64", (S, 2)),
                       ("flux", flux, "float64", (S,)),
                       ("exposure", exposure, "float64", (S, E)),
                       ("upsf", upsf, "float64", (S, 401, E)),
                       ("parint", parint, "float64", (S, 401, E)),
                       ("out", out, "float32", (E, H, W))])
        rp, N = capi.rule_ptr(rule)
        capi.check(capi.lib().xts_point_src_summed_model_dev(
            self._h, C.byref(wcs), p[0], S, p[1], p[2], p[3], p[4], E, rp, N, int(flags), p[5]))

    def point_src_model_map_rows_dev(self, wcs: Wcs, src_sph, E, rule=None,
                                     flags=LAYOUT_SEHW | OUT_F32 | COMPUTE_F32):
        """First half of point_src_model_map_wcs_dev: the work that needs only
        the image, the sources and the rule (row weights), enqueued on the
        context stream.  Keep `rule` alive and unchanged until the second half."""
        S = src_sph.shape[0]
        p = capi.devs([("src_sph", src_sph, "float64", (S, 2))])
        rp, N = capi.rule_ptr(rule)
        capi.check(capi.lib().xts_point_src_model_map_rows_dev(
            self._h, C.byref(wcs), p[0], S, int(E), rp, N, int(flags)))

    def point_src_model_map_finish_dev(self, wcs: Wcs, src_sph, exposure, upsf, parint, out,
                                       rule=None, flags=LAYOUT_SEHW | OUT_F32 | COMPUTE_F32):
        """Second half: table conversion, map integral, correction factor and
        the map, after point_src_model_map_rows_dev with the same arguments."""
        S, E = exposure.shape
        p = capi.devs(self._map_specs(wcs, src_sph, exposure, upsf, parint, out, int(flags)))
        rp, N = capi.rule_ptr(rule)
        capi.check(capi.lib().xts_point_src_model_map_finish_dev(
            self._h, C.byref(wcs), p[0], S, p[1], p[2], p[3], E, rp, N, int(flags), p[4]))
