"""asx_estimate_error (HIP) vs the oracle's AsterX_EstimateError on a seeded
curved-space state: every regrid_method over cell-, vertex- and
edge-centred grid functions at once.  Both sides evaluate the same
expressions in IEEE FP64 without contraction; bar 1e-15 relative.  Also the
C-ABI refusals and the local_estimate_error switch."""
import ctypes as C

import numpy as np
import pytest

import oracle
from asterx_amd import abi
from states import make_state, rel_err

pytestmark = pytest.mark.gpu

NAMES = ["rho", "press", "velx", "alp", "Avec_x"]


def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch


def _slots(names):
    return (C.c_int32 * max(1, len(names)))(*[abi.SLOT[n] for n in names])


@pytest.mark.parametrize("method", list(abi.REGRID))
def test_regrid_parity(method):
    torch = _torch()
    from asterx_amd import lib
    h = make_state(curved=True)
    h["regrid_error"] = -1.0
    d = h.to("cuda")
    par = abi.params(regrid_method=method, epsilon_err=0.02)
    s = _slots(NAMES)
    abi.check("asx_estimate_error", lib.lib().asx_estimate_error(
        C.byref(d.struct()), C.byref(par), len(NAMES), s, None))
    torch.cuda.synchronize()
    abi.check("orc_estimate_error", oracle.lib().orc_estimate_error(
        C.byref(h.struct()), C.byref(par), len(NAMES), s))
    assert rel_err(d.numpy("regrid_error"), h.numpy("regrid_error")) <= 1e-15
    assert np.all(d.numpy("regrid_error")[h.interior("regrid_error")] >= 0)


def test_regrid_refusals_and_switch():
    torch = _torch()
    from asterx_amd import lib
    L = lib.lib()
    h = make_state(curved=True)
    h["regrid_error"] = -1.0
    d = h.to("cuda")
    ps = d.struct()
    assert L.asx_estimate_error(C.byref(ps), C.byref(abi.params()), 10, _slots(["rho"] * 10),
                                None) == 7
    bad = (C.c_int32 * 1)(abi.NUM_GF)
    assert L.asx_estimate_error(C.byref(ps), C.byref(abi.params()), 1, bad, None) == 3
    par = abi.params()
    par.regrid_method = 9
    assert L.asx_estimate_error(C.byref(ps), C.byref(par), 1, _slots(["rho"]), None) == 3
    assert L.asx_estimate_error(C.byref(ps), C.byref(abi.params(local_estimate_error=0)), 1,
                                _slots(["rho"]), None) == 0
    torch.cuda.synchronize()
    assert np.all(d.numpy("regrid_error") == -1.0)
    abi.check("asx_estimate_error", L.asx_estimate_error(C.byref(ps), C.byref(abi.params()), 0,
                                                         None, None))
    torch.cuda.synchronize()
    assert np.all(d.numpy("regrid_error")[h.interior("regrid_error")] == 0.0)
