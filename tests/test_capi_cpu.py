"""CPU-side checks of the product: the C-ABI library loads and exports every
entry point the header declares; host-only entry points agree with the
oracle; argument validation fails loudly.  No device compute here."""
import ctypes as C

import numpy as np
import pytest

from xtsrcmaps_amd import capi
import xtsrcmaps_amd as xa


def test_library_exports_every_declared_symbol():
    lib = capi.lib()
    names = capi.declared_symbols()
    assert len(names) >= 25
    for n in names:
        assert hasattr(lib, n), n
    assert lib.xts_abi_version() == 1


def test_ctypes_signatures_cover_the_header():
    declared = set(capi.declared_symbols())
    assert declared - {"xts_last_error"} == set(capi._SIGS)


def test_costhetas_and_ltf_match_oracle(orc):
    np.testing.assert_allclose(xa.costhetas(40, 0.0, True), orc.costhetas(40, 0.0, True),
                               rtol=0, atol=0)
    logE = np.linspace(2, 6, 37)
    p0 = np.array([-1.2, 3.9, -0.5, 2.3, -0.07, 3.5], np.float32)
    p1 = np.array([1.9, -4.2, 0.47, 2.3, 0.066, 3.5], np.float32)
    np.testing.assert_array_equal(xa.livetime_efficiency_factors(logE, p0, p1),
                                  orc.ltf(logE, p0, p1))


def test_separations_match_oracle(orc):
    np.testing.assert_allclose(xa.separations(), orc.separations(), rtol=1e-15)


def test_context_without_device_fails_loudly():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a device is present")
    with pytest.raises(capi.XtsError):
        xa.Context(0)


def test_null_pointer_is_rejected():
    rc = capi.lib().xts_costhetas(40, 0.0, 1, C.cast(None, C.POINTER(C.c_double)))
    assert rc == 1
    assert b"bad arguments" in capi.lib().xts_last_error()
