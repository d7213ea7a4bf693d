"""asx_con2prim_group_noalloc against asx_con2prim_group.

The allocation-free entry point runs the same kernels with the queue, its
count and the counters in the library's fixed device array, so on the same
input it must give the same primitives, conserved variables, masks, flags
and counters, bit for bit.  Two calls in a row, the second on a smaller
patch, check that what the first call leaves in the shared array does not
reach the second.  Patches above the array's 2^24 interior zones are refused
before any device work.
"""
import ctypes as C

import numpy as np
import pytest

from asterx_amd import abi
from states import make_state

pytestmark = pytest.mark.gpu

OUT = ["rho", "velx", "vely", "velz", "eps", "press", "temperature", "entropy", "Ye",
       "Bvecx", "Bvecy", "Bvecz", "con2prim_flag", "aster_mask_cc"] + abi.CONS + \
    abi.DB + abi.SAVED + abi.ZS


def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch


def _poison(p):
    """Two zones below the density cut (atmosphere) and one NaN tau (the
    primary solver fails it, so it goes through the queue)."""
    dens = p.numpy("dens").copy()
    dens[5, 6, 7] = 1e-14
    dens[6, 7, 5] = 2e-13
    tau = p.numpy("tau").copy()
    tau[7, 5, 6] = np.nan
    p["dens"] = dens
    p["tau"] = tau


def _run_pair(n, seed):
    torch = _torch()
    from asterx_amd import lib
    L = lib.lib()
    h = make_state(n=n, curved=True, seed=seed)
    _poison(h)
    a, b = h.to("cuda"), h.to("cuda")
    par = abi.params(c2p_prime="Noble", c2p_second="Palenzuela", gl_gamma=5.0 / 3.0,
                     rho_abs_min=1e-12)
    ca, cb = abi.asx_counters(), abi.asx_counters()
    abi.check("asx_con2prim_group", L.asx_con2prim_group(
        C.byref(a.struct()), C.byref(par), C.byref(ca), None))
    abi.check("asx_con2prim_group_noalloc", L.asx_con2prim_group_noalloc(
        C.byref(b.struct()), C.byref(par), C.byref(cb), None))
    torch.cuda.synchronize()
    for name in OUT:
        x, y = a.numpy(name), b.numpy(name)
        assert x.tobytes() == y.tobytes(), name
    assert list(ca.c2p_flags) == list(cb.c2p_flags)
    assert list(ca.c2p_iters) == list(cb.c2p_iters)
    fl = a.numpy("con2prim_flag")[h.interior("rho")]
    assert (fl != 1).sum() >= 3, np.unique(fl)
    return ca


def test_noalloc_matches_allocating_entry_point_twice():
    c1 = _run_pair((20, 18, 16), seed=5)
    c2 = _run_pair((12, 10, 8), seed=6)
    assert sum(c1.c2p_flags) == 20 * 18 * 16
    assert sum(c2.c2p_flags) == 12 * 10 * 8


def test_noalloc_refuses_patches_above_its_array():
    _torch()
    from asterx_amd import lib
    pt = abi.asx_patch()
    pt.n[0], pt.n[1], pt.n[2] = 257, 256, 256
    pt.ng = 3
    for d in range(3):
        pt.dx[d] = 1.0 / 256
    par = abi.params()
    st = lib.lib().asx_con2prim_group_noalloc(C.byref(pt), C.byref(par), None, None)
    assert abi.STATUS[st] == "ASX_ERR_SIZE"
