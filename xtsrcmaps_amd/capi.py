"""ctypes declarations of the C ABI in include/xtsrcmaps_amd.h."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libxtsrcmaps_amd.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "xtsrcmaps_amd.h")

PROJ_CAR, PROJ_AIT = 0, 1
LAYOUT_SHWE, LAYOUT_SEHW = 0x0, 0x1
OUT_F64, OUT_F32 = 0x0, 0x2
COMPUTE_F32, COMPUTE_F64 = 0x0, 0x4
CF_EXACT_RADIUS = 0x8
SURROGATE, SURROGATE_NO_NET, SURROGATE_FP8 = 0x10, 0x20, 0x40
MLP_PARAMS, MLP_NORM = 8705, 10
PREC_FP64, PREC_BF16X3 = 0, 1
SCALE_PIXEL, SCALE_SRC_ENERGY = 0, 1

_d = C.POINTER(C.c_double)
_f = C.POINTER(C.c_float)
_u8 = C.POINTER(C.c_uint8)
_i64 = C.c_int64
_vp = C.c_void_p


class XtsError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"xtsrcmaps_amd error {code}: {msg}")
        self.code = code


class Wcs(C.Structure):
    """xts_wcs: the fields of the reference's Fermi::WcsConfig."""
    _fields_ = [("naxis1", C.c_int64), ("naxis2", C.c_int64),
                ("crpix1", C.c_double), ("crpix2", C.c_double),
                ("crval1", C.c_double), ("crval2", C.c_double),
                ("cdelt1", C.c_double), ("cdelt2", C.c_double),
                ("crota2", C.c_double), ("proj", C.c_int32), ("is_galactic", C.c_int32)]

    @classmethod
    def make(cls, naxis, crpix, crval, cdelt, crota2=0.0, proj="CAR", galactic=False):
        p = {"CAR": PROJ_CAR, "AIT": PROJ_AIT}[proj] if isinstance(proj, str) else int(proj)
        return cls(int(naxis[0]), int(naxis[1]), float(crpix[0]), float(crpix[1]),
                   float(crval[0]), float(crval[1]), float(cdelt[0]), float(cdelt[1]),
                   float(crota2), p, int(bool(galactic)))


class ExpcubeInfo(C.Structure):
    """xts_expcube_info: the livetime table's keywords."""
    _fields_ = [("nside", C.c_int64), ("nbrbins", C.c_int64), ("cosmin", C.c_double),
                ("thetabin", C.c_int32), ("ordering", C.c_char * 16),
                ("coordsys", C.c_char * 16)]


class PsfGrid(C.Structure):
    _fields_ = [("params", _d), ("cosths", _d), ("logEs", _d), ("Mc", C.c_int64),
                ("Me", C.c_int64)]


class Footprint(C.Structure):
    _fields_ = [("h0", C.c_int64), ("h1", C.c_int64), ("w0", C.c_int64), ("w1", C.c_int64),
                ("offset", C.c_int64)]


class DiffuseGeom(C.Structure):
    """xts_diffuse_geom: crop, resampled grid, kernel and transform sizes."""
    _fields_ = [("crop_a1", C.c_int64), ("crop_a2", C.c_int64), ("crop_n1", C.c_int64),
                ("crop_n2", C.c_int64), ("m1", C.c_int64), ("m2", C.c_int64),
                ("step1", C.c_double), ("step2", C.c_double), ("grid", Wcs),
                ("wrap", C.c_int64), ("nk", C.c_int64), ("kc1", C.c_double),
                ("kc2", C.c_double), ("fft_p", C.c_int64), ("fft_q", C.c_int64)]


class AeffGrid(C.Structure):
    _fields_ = [("params", _d), ("cosths", _d), ("logEs", _d), ("Mc", C.c_int64),
                ("Me", C.c_int64), ("min_costheta", C.c_double)]


# name -> (restype, argtypes); every entry point returns int
_SIGS = {
    "xts_ctx_create": [C.c_int, C.POINTER(_vp)],
    "xts_ctx_destroy": [_vp],
    "xts_ctx_set_stream": [_vp, _vp],
    "xts_ctx_synchronize": [_vp],
    "xts_ctx_enable_timing": [_vp, C.c_int],
    "xts_ctx_set_tuning": [_vp, C.c_char_p, _i64],
    "xts_ctx_kernel_times": [_vp, C.c_char_p, _d, C.POINTER(_i64)],
    "xts_abi_version": [],
    "xts_costhetas": [_i64, C.c_double, C.c_int32, _d],
    "xts_livetime_efficiency_factors": [_d, _i64, _f, _f, _d],
    "xts_aeff_value": [_vp, _d, _i64, _d, _i64, C.POINTER(AeffGrid), _d],
    "xts_src_cosbins": [_vp, _d, _i64, _d, _i64, _i64, _d],
    "xts_src_cosbins_dev": [_vp, _vp, _i64, _vp, _i64, _i64, _vp],
    "xts_contract": [_vp, _d, _d, _i64, _i64, _i64, _d, C.c_int32],
    "xts_exposure": [_vp, _d, _d, _d, _d, _d, _i64, _i64, _i64, _d, C.c_int32],
    "xts_exposure_dev": [_vp, _vp, _vp, _vp, _vp, _vp, _i64, _i64, _i64, _vp, C.c_int32],
    "xts_normalize_rpsf": [_vp, _d, _d, _i64, _i64, C.c_float, C.c_float, C.c_float],
    "xts_king": [_vp, C.POINTER(PsfGrid), _d],
    "xts_psf_lookup_table_and_partial_integrals": [
        _vp, C.POINTER(PsfGrid), C.POINTER(PsfGrid), _d, _i64, _d, _i64, _d, _d, _d, _d, _d,
        _d, _i64, _d, _d, C.c_int32],
    "xts_psf_lookup_table_and_partial_integrals_dev": [
        _vp, C.POINTER(PsfGrid), C.POINTER(PsfGrid), _vp, _i64, _vp, _i64, _vp, _vp, _vp, _vp,
        _vp, _vp, _i64, _vp, _vp, C.c_int32],
    "xts_partial_total_integral": [_vp, _d, _i64, _i64, _d, _d],
    "xts_bilerp": [_vp, _d, _i64, _d, _i64, C.POINTER(PsfGrid), _d, _d],
    "xts_corrected_exposure_psf": [_vp, _d, _d, _d, _d, _d, _i64, _i64, _i64, _d],
    "xts_mean_psf": [_vp, _d, _d, _d, _i64, _i64, _d],
    "xts_normalize": [_vp, _d, _d, _i64, _i64],
    "xts_point_src_model_map_wcs": [_vp, C.POINTER(Wcs), _d, _i64, _d, _d, _d, _i64, _d, _i64,
                                    C.c_int32, _vp],
    "xts_point_src_model_map_wcs_dev": [_vp, C.POINTER(Wcs), _vp, _i64, _vp, _vp, _vp, _i64,
                                        _d, _i64, C.c_int32, _vp],
    "xts_ctx_get_tuning": [_vp, C.c_char_p, C.POINTER(_i64)],
    "xts_point_src_summed_model_dev": [_vp, C.POINTER(Wcs), _vp, _i64, _vp, _vp, _vp, _vp, _i64,
                                       _d, _i64, C.c_int32, _vp],
    "xts_model_map_footprints": [_vp, _i64, C.POINTER(Wcs), C.POINTER(_i64), _vp, _vp, _i64,
                                 C.c_double, C.POINTER(Footprint), _d, C.POINTER(_i64)],
    "xts_point_src_model_maps_batched_dev": [_vp, _i64, C.POINTER(Wcs), C.POINTER(_i64), _vp,
                                             _vp, _vp, _vp, _i64, _d, _i64, C.c_int32,
                                             C.POINTER(Footprint), _vp, _i64],
    "xts_aeff_value_dev": [_vp, _d, _i64, _d, _i64, C.POINTER(AeffGrid), _d, _vp],
    "xts_surrogate_samples_dev": [_vp, C.POINTER(Wcs), _d, _i64, C.c_uint64, _i64, _i64, _vp,
                                  _vp],
    "xts_mlp_forward_dev": [_vp, _vp, _vp, _vp, _i64, _i64, _vp],
    "xts_mlp_infer_bf16_dev": [_vp, _vp, _vp, _vp, _i64, _i64, _vp],
    "xts_mlp_infer_fp8_dev": [_vp, _vp, _vp, _vp, _i64, _i64, _vp],
    "xts_mlp_grad_dev": [_vp, _vp, _vp, _vp, _i64, _vp, _i64, _i64, _vp, _vp],
    "xts_mlp_adam_dev": [_vp, _vp, _vp, _vp, _vp, _i64, C.c_float, C.c_float, C.c_float,
                         C.c_float],
    "xts_surrogate_train_dev": [_vp, _vp, _vp, _vp, _i64, _vp, _i64, _i64, _i64, _i64,
                                C.c_float, C.c_uint64, _vp],
    "xts_ctx_set_surrogate": [_vp, _vp, _vp],
    "xts_point_src_model_map_wcs_dev_noalloc": [_vp, C.POINTER(Wcs), _vp, _i64, _vp, _vp, _vp,
                                                _i64, _d, _i64, C.c_int32, _vp],
    "xts_point_src_model_map_rows_dev": [_vp, C.POINTER(Wcs), _vp, _i64, _i64, _d, _i64,
                                         C.c_int32],
    "xts_point_src_model_map_finish_dev": [_vp, C.POINTER(Wcs), _vp, _i64, _vp, _vp, _vp, _i64,
                                           _d, _i64, C.c_int32, _vp],
    "xts_cubature_convolve_map_psf": [_vp, C.POINTER(Wcs), _d, _i64, _d, _i64, _d, _i64,
                                      C.c_int32, _vp],
    "xts_solid_angle": [_vp, C.POINTER(Wcs), _d],
    "xts_psf_boundary_radius": [_vp, C.POINTER(Wcs), _d, _i64, _d, _u8],
    "xts_map_integral": [_vp, C.POINTER(Wcs), _d, _d, _i64, _i64, _d, _u8, _d],
    "xts_map_correction_factor": [_vp, _d, _d, _u8, _d, _d, _i64, _i64, _d],
    "xts_scale_map": [_vp, _d, _i64, _i64, _i64, _i64, _d, C.c_int32],
    "xts_pix2sph": [_vp, C.POINTER(Wcs), _d, _i64, _d],
    "xts_write_src_model": [C.c_char_p, _d, _i64, _i64, _i64, _i64, C.POINTER(C.c_char_p),
                            C.c_int32, C.POINTER(Wcs)],
    "xts_sph2pix": [_vp, C.POINTER(Wcs), _d, _i64, _d],
    "xts_read_irf_pars": [C.c_char_p, C.c_char_p, C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                          _i64, C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                          C.POINTER(C.c_float)],
    "xts_read_mapcube": [C.c_char_p, C.POINTER(Wcs), C.POINTER(C.c_int64), _f, _d,
                         C.POINTER(C.c_int64)],
    "xts_diffuse_geometry": [C.POINTER(Wcs), C.POINTER(Wcs), C.c_double, C.c_double,
                             C.POINTER(DiffuseGeom)],
    "xts_diffuse_src_model_map_wcs_dev": [_vp, C.POINTER(Wcs), C.POINTER(Wcs), _vp, _i64, _vp,
                                          _vp, _i64, C.c_double, C.c_double, _vp, C.c_int32],
    "xts_diffuse_src_model_map_wcs": [_vp, C.POINTER(Wcs), C.POINTER(Wcs), _f, _i64, _d, _d,
                                      _i64, C.c_double, C.c_double, _d],
    "xts_psf_make_kernel": [_vp, _i64, C.POINTER(Wcs), C.c_double, C.c_double, _d, _i64, _d],
    "xts_fft_convolve_map_psf": [_vp, _d, _i64, _i64, _i64, _d, _i64, _i64, _d],
    "xts_read_expcube": [C.c_char_p, C.c_char_p, C.POINTER(ExpcubeInfo), C.POINTER(C.c_float),
                         C.POINTER(C.c_float), C.POINTER(C.c_float)],
}

_lib = None


def lib():
    """Load the in-tree library; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: run `python -m xtsrcmaps_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, args in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = C.c_int
            fn.argtypes = args
        L.xts_last_error.restype = C.c_char_p
        L.xts_last_error.argtypes = []
        _lib = L
    return _lib


def declared_symbols() -> list[str]:
    """Entry points declared in include/xtsrcmaps_amd.h."""
    import re
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(xts_[a-z0-9_]+)\s*\(", text)))


def check(rc: int) -> None:
    if rc != 0:
        raise XtsError(rc, lib().xts_last_error().decode())


def f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def dptr(a: np.ndarray):
    return a.ctypes.data_as(_d)


def fptr(a: np.ndarray):
    return a.ctypes.data_as(_f)


def u8ptr(a: np.ndarray):
    return a.ctypes.data_as(_u8)


def tptr(t) -> _vp:
    """Device pointer of a contiguous torch CUDA tensor."""
    if not t.is_cuda or not t.is_contiguous():
        raise XtsError(1, "device entry points need contiguous CUDA tensors")
    return _vp(t.data_ptr())


def _meta(name, t, dtype, shape=None, numel=None):
    got = str(t.dtype).replace("torch.", "")
    if got != dtype:
        raise XtsError(1, f"{name}: dtype {got}, expected {dtype}")
    if shape is not None and tuple(t.shape) != tuple(int(x) for x in shape):
        raise XtsError(1, f"{name}: shape {tuple(t.shape)}, expected {tuple(shape)}")
    if numel is not None and t.numel() != int(numel):
        raise XtsError(1, f"{name}: {t.numel()} elements, expected {int(numel)}")
    if not t.is_contiguous():
        raise XtsError(1, f"{name}: not contiguous")


def devs(specs, device=None) -> list:
    """Device pointers of torch tensors checked against what an entry point
    will read or write through them.  specs: (name, tensor, dtype, shape) or
    (name, tensor, dtype, None, numel).  Every tensor's dtype, shape or
    element count and contiguity are checked first, then that all are on one
    GPU (device, if given, or the first tensor's).  A mismatch raises
    XtsError before any launch."""
    for sp in specs:
        _meta(*sp)
    if device is None and specs:
        device = specs[0][1].device
    for sp in specs:
        name, t = sp[0], sp[1]
        if not t.is_cuda:
            raise XtsError(1, f"{name}: device entry points need CUDA tensors")
        if t.device != device:
            raise XtsError(1, f"{name}: on {t.device}, the other arguments on {device}")
    return [_vp(sp[1].data_ptr()) for sp in specs]


_keep = []


def psf_grid(params, cosths, logEs):
    p, c, e = f64(params), f64(cosths), f64(logEs)
    g = PsfGrid(dptr(p), dptr(c), dptr(e), len(c), len(e))
    return g, (p, c, e)


def psf_grid_specs(params, cosths, logEs):
    """Argument specs (for devs) of a device-resident King PSF grid."""
    Mc, Me = cosths.shape[0], logEs.shape[0]
    return [("psf params", params, "float64", (Mc, Me, 6)),
            ("psf cosths", cosths, "float64", (Mc,)), ("psf logEs", logEs, "float64", (Me,))]


def psf_grid_dev(ptrs, Mc, Me):
    """PsfGrid of three device pointers (params, cosths, logEs)."""
    return PsfGrid(C.cast(ptrs[0], _d), C.cast(ptrs[1], _d), C.cast(ptrs[2], _d), Mc, Me)


def aeff_grid(cosths, logEs, params, min_costheta):
    p, c, e = f64(params), f64(cosths), f64(logEs)
    g = AeffGrid(dptr(p), dptr(c), dptr(e), len(c), len(e), float(min_costheta))
    return g, (p, c, e)


def rule_ptr(rule):
    if rule is None:
        return C.cast(None, _d), 0
    r = f64(rule).reshape(-1, 3)
    _keep.append(r)
    if len(_keep) > 16:
        del _keep[:8]
    return dptr(r), r.shape[0]


def alloc_map(S, wcs, E, flags):
    dt = np.float32 if flags & OUT_F32 else np.float64
    H, W = wcs.naxis1, wcs.naxis2
    shape = (S, E, H, W) if flags & LAYOUT_SEHW else (S, H, W, E)
    return np.zeros(shape, dtype=dt)
