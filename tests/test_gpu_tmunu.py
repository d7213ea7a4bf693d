"""asx_tmunu (HIP) vs the oracle's AsterX_Tmunu on a seeded curved-space
state: both interpolation orders, v_vec and z_vec velocities, eT* holding
seeded values beforehand (the entry adds into them).  Bar: 1e-12 relative in
max norm per component (the device evaluates one of the three equal axis
orders of the order-4 stencil, the reference averages all three).  Also the
C-ABI refusals and the update_tmunu switch, a 128^3 patch compared with
the oracle on sampled vertices, and the driver's post-step tensor after two
RK2 steps of the Alfven wave on both backends."""
import ctypes as C

import numpy as np
import pytest

import oracle
from asterx_amd import abi
from states import make_state, rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch


def _lib():
    from asterx_amd import lib
    return lib.lib()


def _with_eT(h, seed=11):
    rng = np.random.default_rng(seed)
    for n in abi.TMUNU:
        h[n] = rng.uniform(-1.0, 1.0, h.shape(n))
    return h


def _variants():
    from asterx_amd import lib
    return range(lib.lib().asx_variant_count(abi.KCLASS["tmunu"]))


@pytest.mark.parametrize("order", [2, 4])
@pytest.mark.parametrize("rec", ["v_vec", "z_vec"])
@pytest.mark.parametrize("variant", list(range(9)))
def test_tmunu_parity(order, rec, variant):
    torch = _torch()
    if variant not in _variants():
        pytest.skip("variant not built")
    h = _with_eT(make_state(curved=True))
    d = h.to("cuda")
    par = abi.params(tmunu_interp_order=order, recon_type=rec)
    L = _lib()
    abi.check("asx_set_variant", L.asx_set_variant(abi.KCLASS["tmunu"], variant))
    try:
        abi.check("asx_tmunu", L.asx_tmunu(C.byref(d.struct()), C.byref(par), None))
        torch.cuda.synchronize()
    finally:
        abi.check("asx_set_variant", L.asx_set_variant(abi.KCLASS["tmunu"], 0))
    before = {n: h.numpy(n).copy() for n in abi.TMUNU}
    abi.check("orc_tmunu", oracle.lib().orc_tmunu(C.byref(h.struct()), C.byref(par)))
    bad = {}
    for n in abi.TMUNU:
        e = rel_err(d.numpy(n), h.numpy(n))
        if not e <= TOL:
            bad[n] = e
        # the added tensor itself, not only the sum, agrees
        sl = h.interior(n)
        inc_d = d.numpy(n)[sl] - before[n][sl]
        inc_h = h.numpy(n)[sl] - before[n][sl]
        assert np.max(np.abs(inc_d - inc_h)) <= 1e-12 * np.max(np.abs(inc_h)), n
    assert not bad, bad


def test_tmunu_refusals_and_switch():
    torch = _torch()
    h = _with_eT(make_state(curved=True))
    d = h.to("cuda")
    lib = _lib()
    ref = {n: d.numpy(n).copy() for n in abi.TMUNU}
    assert lib.asx_tmunu(C.byref(d.struct()), C.byref(abi.params(tmunu_interp_order=3)),
                         None) == 3
    assert lib.asx_tmunu(C.byref(d.struct()), C.byref(abi.params(update_tmunu=0)), None) == 0
    s = d.struct()
    s.gf[abi.SLOT["eTyz"]].data = None
    assert lib.asx_tmunu(C.byref(s), C.byref(abi.params()), None) == 1
    torch.cuda.synchronize()
    for n in abi.TMUNU:
        assert np.array_equal(d.numpy(n), ref[n]), n


def test_tmunu_large_patch_sampled():
    """128^3 cells: every interior vertex computed on the device; the oracle
    evaluated on a 16^3-cell window cut out of the same data at a corner and
    at the centre (the window carries its own ghosts, so its vertices are the
    patch's vertices)."""
    torch = _torch()
    from asterx_amd.patch import Patch
    n, ng = 128, 2
    rng = np.random.default_rng(5)
    big = Patch((n,) * 3, ng, (-1.0, -1.0, -1.0), (2.0 / n,) * 3, device="cuda")
    names = abi.ADM[:10] + ["rho", "eps", "press", "velx", "vely", "velz",
                            "Bvecx", "Bvecy", "Bvecz"] + abi.TMUNU
    host = {}
    for nm in names:
        shp = big.shape(nm)
        if nm == "alp":
            a = 0.9 + 0.05 * rng.random(shp)
        elif nm in ("gxx", "gyy", "gzz"):
            a = 1.0 + 0.1 * rng.random(shp)
        elif nm in ("velx", "vely", "velz"):
            a = 0.3 * (rng.random(shp) - 0.5)
        elif nm in ("rho", "eps", "press"):
            a = 0.5 + rng.random(shp)
        else:
            a = 0.05 * (rng.random(shp) - 0.5)
        host[nm] = a
        big[nm] = a
    par = abi.params()
    abi.check("asx_tmunu", _lib().asx_tmunu(C.byref(big.struct()), C.byref(par), None))
    torch.cuda.synchronize()
    m = 16
    for c0 in (0, n // 2 - m // 2, n - m):
        win = Patch((m,) * 3, ng, (0, 0, 0), (2.0 / n,) * 3)
        for nm in names:
            cv = abi.centering(nm) & 1
            e = m + 2 * ng + cv
            win[nm] = host[nm][c0:c0 + e, c0:c0 + e, c0:c0 + e]
        abi.check("orc_tmunu", oracle.lib().orc_tmunu(C.byref(win.struct()), C.byref(par)))
        for nm in abi.TMUNU:
            e = m + 2 * ng + 1
            got = big.numpy(nm)[c0:c0 + e, c0:c0 + e, c0:c0 + e][win.interior(nm)]
            want = win.numpy(nm)[win.interior(nm)]
            assert rel_err(got, want) <= TOL, (c0, nm)


def test_evolution_with_tmunu_matches_oracle():
    """Two RK2 steps of the periodic Alfven wave with the stress-energy
    tensor set after every con2prim, on the device and on the oracle: the
    tensors agree to the evolution's own agreement (1e-10 relative)."""
    torch = _torch()
    import os
    import sys
    sys.path.insert(0, os.path.dirname(__file__))
    import tsv_cases
    from asterx_amd import lib
    from asterx_amd.driver import Evolution
    c = tsv_cases.CASES["Alfven_wave"]
    par = tsv_cases.case_params(c)
    out = []
    for dev, be in ((None, oracle.backend()), ("cuda", lib.backend())):
        p = tsv_cases.make_patch(c, device=dev)
        ev = Evolution(be, p, par, bc="periodic", tmunu=True)
        for _ in range(2):
            ev.rk2_step(0.2 * min(p.dx))
        if dev:
            torch.cuda.synchronize()
        out.append({n: p.numpy(n).copy() for n in abi.TMUNU})
    for n in abi.TMUNU:
        assert rel_err(out[1][n], out[0][n]) <= 1e-10, (n, rel_err(out[1][n], out[0][n]))
