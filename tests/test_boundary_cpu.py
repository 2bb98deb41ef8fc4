"""CPU checks of the drop-in boundary: the reference's own call lines into
the path (xtsrcmaps/main.cxx:14-31) compile against this library's headers,
and the device-pointer Python wrappers refuse arguments whose dtype, shape,
size or residence does not match what the entry point reads or writes,
before any launch (no device is needed: the checks run first)."""
import os
import subprocess

import numpy as np
import pytest
import torch

import xtsrcmaps_amd as xa
from xtsrcmaps_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CXX = "/opt/rocm/llvm/bin/clang++"


def test_reference_main_calls_compile():
    src = os.path.join(ROOT, "tests", "cpp_syntax", "main_calls.cpp")
    r = subprocess.run([CXX, "-std=c++23", "-fsyntax-only", "-Wall", "-I",
                        os.path.join(ROOT, "include"), "-I",
                        os.path.join(ROOT, "xtsrcmaps_amd", "include"), src],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_three_argument_compute_exposure_is_gone():
    """compute_exposure takes the configuration first, as main.cxx:28 calls it."""
    code = ('#include "xtsrcmaps/exposure/exposure.hxx"\n'
            "void f(Fermi::Obs::XtObs const& o, "
            "Fermi::Source::SourceData<Fermi::Source::PointSource> const& s, "
            "Fermi::Irf::XtIrf const& i) { (void)Fermi::Exposure::compute_exposure(o, s, i); }\n")
    r = subprocess.run([CXX, "-std=c++23", "-fsyntax-only", "-x", "c++", "-", "-I",
                        os.path.join(ROOT, "include"), "-I",
                        os.path.join(ROOT, "xtsrcmaps_amd", "include")],
                       input=code, capture_output=True, text=True)
    assert r.returncode != 0


def _ctx():
    """A Context object whose handle is never reached: every case below must
    be refused by the argument checks."""
    c = object.__new__(xa.Context)
    c._h = None
    return c


def _z(*shape, dt=torch.float64):
    return torch.zeros(shape, dtype=dt)


@pytest.mark.parametrize("case", ["upsf_f32", "src_shape", "out_dtype", "out_size", "cpu"])
def test_model_map_dev_refuses_mismatched_arguments(case):
    S, E, n = 3, 5, 8
    wcs = xa.Wcs.make((n, n), (4.5, 4.5), (180.0, 0.0), (-0.1, 0.1), 0.0, "CAR")
    a = dict(src=_z(S, 2), expo=_z(S, E), upsf=_z(S, 401, E), parint=_z(S, 401, E),
             out=_z(S, E, n, n, dt=torch.float32))
    if case == "upsf_f32":
        a["upsf"] = _z(S, 401, E, dt=torch.float32)
    elif case == "src_shape":
        a["src"] = _z(S, 3)
    elif case == "out_dtype":
        a["out"] = _z(S, E, n, n)
    elif case == "out_size":
        a["out"] = _z(S, E, n, n - 1, dt=torch.float32)
    with pytest.raises(capi.XtsError) as ei:
        _ctx().point_src_model_map_wcs_dev(wcs, a["src"], a["expo"], a["upsf"], a["parint"],
                                           a["out"])
    msg = str(ei.value)
    expect = {"upsf_f32": "upsf: dtype", "src_shape": "src_sph: shape",
              "out_dtype": "out: dtype", "out_size": "out:", "cpu": "CUDA"}[case]
    assert expect in msg, msg


def test_exposure_and_psf_dev_refuse_mismatched_arguments():
    S, C, E = 4, 40, 7
    with pytest.raises(capi.XtsError, match="aeff_back: shape"):
        _ctx().exposure_dev(_z(S, C), _z(S, C), _z(C, E), _z(C, E + 1), _z(2, E), _z(S, E))
    with pytest.raises(capi.XtsError, match="exposure: dtype"):
        _ctx().exposure_dev(_z(S, C), _z(S, C), _z(C, E), _z(C, E), _z(2, E),
                            _z(S, E, dt=torch.float32))
    grid = (_z(10, 25, 6), _z(10), _z(25))
    bad = (_z(10, 25, 5), _z(10), _z(25))
    with pytest.raises(capi.XtsError, match="psf params: shape"):
        _ctx().psf_lookup_dev(grid, bad, _z(C), _z(E), _z(C, E), _z(C, E), _z(S, C), _z(S, C),
                              _z(2, E), _z(S, E), _z(S, 401, E), _z(S, 401, E))
    with pytest.raises(capi.XtsError, match="parint: shape"):
        _ctx().psf_lookup_dev(grid, grid, _z(C), _z(E), _z(C, E), _z(C, E), _z(S, C), _z(S, C),
                              _z(2, E), _z(S, E), _z(S, 401, E), _z(S, 400, E))
    with pytest.raises(capi.XtsError, match="cube: shape"):
        _ctx().src_cosbins_dev(_z(S, 2), _z(12 * 16 * 16, C), 32, _z(S, C))
