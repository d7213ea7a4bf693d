"""Debug poison (asx_set_poison; CarpetX::poison_undefined_values = yes,
FishboneMoncriefIDX/par/FMdisk_unigrid.par:26).  Every poisoned entry
point, run on a seeded curved state with poison on, must write its whole
declared region (no ASX_ERR_POISON) and give results bit-identical to the
same call without poison -- so no kernel reads an output point it has not
written in that call (a poisoned value read would propagate as NaN).
Points outside the declared region keep the pattern.  asx_poison_check
counts planted patterns exactly (the same count kernel decides
ASX_ERR_POISON after a poisoned call)."""
import ctypes as C

import numpy as np
import pytest

from asterx_amd import abi
from states import make_state

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch


def names(a, b):
    i, j = abi.SLOTS.index(a), abi.SLOTS.index(b)
    return abi.SLOTS[i:j + 1]


def face_region(p, name, d, halo):
    """Slice of the face box of direction d with `halo` layers on the
    transverse axes."""
    c = abi.centering(name)
    sl = []
    for ax in (2, 1, 0):
        h = 0 if ax == d else halo
        sl.append(slice(p.ng - h, p.ng + p.n[ax] + ((c >> ax) & 1) + h))
    return tuple(sl)


def _outputs(p, case):
    """{slot name: region slice} the entry point of `case` declares written."""
    out = {}
    if case == "fluxes":
        for d, f in enumerate("xyz"):
            for n in names("f%sdens" % f, "f%sBz" % f):
                out[n] = face_region(p, n, d, 1)
        for d, pair in enumerate([("vtilde_y_xface", "vtilde_z_xface"),
                                  ("vtilde_x_yface", "vtilde_z_yface"),
                                  ("vtilde_x_zface", "vtilde_y_zface")]):
            for n in pair + ("amax_%sface" % "xyz"[d], "amin_%sface" % "xyz"[d]):
                out[n] = face_region(p, n, d, 1)
        for n in names("Fx", "G"):
            out[n] = p.interior(n, halo=2)
    elif case == "source_terms":
        out = {n: p.interior(n) for n in names("densrhs", "DYe_rhs")}
    elif case == "rhs":
        out = {n: p.interior(n) for n in names("Avec_x_rhs", "Psi_rhs")}
    elif case == "bstag_from_A":
        out = {n: p.interior(n) for n in names("dBx_stag", "dBz_stag")}
    elif case == "dB_from_bstag":
        out = {n: p.interior(n) for n in names("dBx", "dBz")}
    elif case == "B_from_dB":
        out = {n: p.interior(n) for n in names("Bvecx", "Bvecz")}
    elif case == "prim2con":
        out = {n: p.interior(n) for n in names("dens", "dBz")}
    elif case == "estimate_error":
        out = {"regrid_error": p.interior("regrid_error")}
    return out


def _call(L, case, d, par, stream=None):
    ps = d.struct()
    if case == "fluxes":
        return L.asx_fluxes(C.byref(ps), C.byref(par), 1, None, stream)
    if case == "source_terms":
        return L.asx_source_terms(C.byref(ps), C.byref(par), stream)
    if case == "rhs":
        return L.asx_rhs(C.byref(ps), C.byref(par), stream)
    if case == "bstag_from_A":
        return L.asx_bstag_from_A(C.byref(ps), stream)
    if case == "dB_from_bstag":
        return L.asx_dB_from_bstag(C.byref(ps), stream)
    if case == "B_from_dB":
        return L.asx_B_from_dB(C.byref(ps), stream)
    if case == "prim2con":
        return L.asx_prim2con(C.byref(ps), stream)
    slots = (C.c_int32 * 2)(abi.SLOT["rho"], abi.SLOT["Avec_x"])
    return L.asx_estimate_error(C.byref(ps), C.byref(par), 2, slots, stream)


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint64)


CASES = ["fluxes", "source_terms", "rhs", "bstag_from_A", "dB_from_bstag", "B_from_dB",
         "prim2con", "estimate_error"]


@pytest.mark.parametrize("case", CASES)
def test_poisoned_call_writes_its_region_and_matches(case):
    torch = _torch()
    from asterx_amd import lib
    L = lib.lib()
    h = make_state(curved=True)
    h.alloc("regrid_error", 0.0)
    par = abi.params(reconstruction_method="wenoz", flux_type="HLLE", gl_gamma=5.0 / 3.0,
                     use_uct=1 if case in ("fluxes", "rhs") else 0,
                     local_spatial_order=4, regrid_method="first gradient")
    runs = []
    for poison in (0, 1):
        d = h.to("cuda")
        if case == "rhs":  # fluxes and sources first, unpoisoned, in both runs
            abi.check("asx_fluxes", _call(L, "fluxes", d, par))
            abi.check("asx_source_terms", _call(L, "source_terms", d, par))
        abi.check("asx_set_poison", L.asx_set_poison(poison))
        try:
            st = _call(L, case, d, par)
        finally:
            L.asx_set_poison(0)
        torch.cuda.synchronize()
        if st == 8:
            slot, pts = C.c_int32(), C.c_int64()
            L.asx_poison_report(C.byref(slot), C.byref(pts))
            pytest.fail("%s left %d points of %s unwritten" %
                        (case, pts.value, abi.SLOTS[slot.value]))
        abi.check(case, st)
        runs.append(d)
    outs = _outputs(h, case)
    assert outs
    for n, sl in outs.items():
        a, b = runs[0].numpy(n), runs[1].numpy(n)
        assert np.array_equal(_bits(a[sl]), _bits(b[sl])), n
        outside = np.ones(a.shape, bool)
        outside[sl] = False
        if outside.any():
            assert (_bits(b)[outside] == abi.POISON_BITS).all(), n


def test_poisoned_substep_matches():
    torch = _torch()
    from asterx_amd import lib
    from asterx_amd.patch import StateBuf
    L = lib.lib()
    h = make_state(curved=True)
    par = abi.params(reconstruction_method="mp5", flux_type="HLLE", gl_gamma=5.0 / 3.0)
    res = []
    for poison in (0, 1):
        d = h.to("cuda")
        u0 = StateBuf(d, fill=0.0)
        for name in abi.STATE:
            u0.arr[name] = d[name].clone() * 1.01
        out = StateBuf(d, fill=0.0)
        abi.check("asx_set_poison", L.asx_set_poison(poison))
        try:
            st = L.asx_substep(C.byref(d.struct()), C.byref(u0.struct()), C.byref(out.struct()),
                               0.75, 0.25, 0.0025, C.byref(par), None)
        finally:
            L.asx_set_poison(0)
        torch.cuda.synchronize()
        abi.check("asx_substep", st)
        res.append(out)
    for name in abi.STATE:
        sl = h.interior(name)
        a, b = res[0].arr[name].cpu().numpy(), res[1].arr[name].cpu().numpy()
        assert np.array_equal(_bits(a[sl]), _bits(b[sl])), name
        assert not np.isnan(b[sl]).any(), name
        outside = np.ones(a.shape, bool)
        outside[sl] = False
        assert (_bits(b)[outside] == abi.POISON_BITS).all(), name


@pytest.mark.parametrize("uct", [0, 1])
def test_poisoned_substep_bc_matches(uct):
    """asx_substep_bc fills the face data's ghost points by the boundary
    kind and (with upwind CT) reconstructs face data three rows into them:
    with the face scratch poisoned, a ghost point the fill missed would
    reach the result as NaN."""
    torch = _torch()
    from asterx_amd import lib
    from asterx_amd.patch import StateBuf
    L = lib.lib()
    h = make_state(curved=True)
    par = abi.params(reconstruction_method="wenoz", flux_type="HLLE", gl_gamma=5.0 / 3.0,
                     use_uct=uct, lorenz_damp_fac=0.2)
    res = []
    for poison in (0, 1):
        d = h.to("cuda")
        u0 = StateBuf(d, fill=0.0)
        for name in abi.STATE:
            u0.arr[name] = d[name].clone() * 1.01
        out = StateBuf(d, fill=0.0)
        abi.check("asx_set_poison", L.asx_set_poison(poison))
        try:
            st = L.asx_substep_bc(C.byref(d.struct()), C.byref(u0.struct()),
                                  C.byref(out.struct()), 0.75, 0.25, 0.0025, C.byref(par),
                                  abi.BC["copy"], None)
        finally:
            L.asx_set_poison(0)
        torch.cuda.synchronize()
        abi.check("asx_substep_bc", st)
        res.append(out)
    for name in abi.STATE:
        sl = h.interior(name)
        a, b = res[0].arr[name].cpu().numpy(), res[1].arr[name].cpu().numpy()
        assert np.array_equal(_bits(a[sl]), _bits(b[sl])), name
        assert not np.isnan(b[sl]).any(), name


def test_poison_check_and_report():
    torch = _torch()
    from asterx_amd import lib
    L = lib.lib()
    h = make_state(curved=True)
    d = h.to("cuda")
    a = d["rho"]
    flat = a.view(torch.int64).view(-1)
    pat = abi.POISON_BITS  # < 2**63: the same bits as an int64
    planted = [5, 77, 300, flat.numel() - 1]
    for q in planted:
        flat[q] = pat
    ex, ey, ez = abi.extents(d.n, d.ng, abi.centering("rho"))
    pts = C.c_int64()
    full = ((C.c_int32 * 3)(0, 0, 0), (C.c_int32 * 3)(ex, ey, ez))
    abi.check("asx_poison_check", L.asx_poison_check(C.byref(d.struct()), abi.SLOT["rho"],
                                                     full[0], full[1], C.byref(pts), None))
    assert pts.value == len(planted)
    # a box holding only the first plane of x rows 0..ey-1 at k = 0
    lo, cnt = (C.c_int32 * 3)(0, 0, 0), (C.c_int32 * 3)(ex, ey, 1)
    abi.check("asx_poison_check", L.asx_poison_check(C.byref(d.struct()), abi.SLOT["rho"],
                                                     lo, cnt, C.byref(pts), None))
    assert pts.value == sum(1 for q in planted if q < ex * ey)
    bad = (C.c_int32 * 3)(ex, ey, ez + 1)
    assert L.asx_poison_check(C.byref(d.struct()), abi.SLOT["rho"], full[0], bad,
                              C.byref(pts), None) == 7
    # a call that does no work (local_estimate_error = no) poisons nothing
    # and leaves the report clear
    h.alloc("regrid_error", 0.0)
    d = h.to("cuda")
    par = abi.params(local_estimate_error=0)
    abi.check("asx_set_poison", L.asx_set_poison(1))
    try:
        st = _call(L, "estimate_error", d, par)
    finally:
        L.asx_set_poison(0)
    assert st == 0 and (d.numpy("regrid_error") == 0.0).all()
    slot, cnt = C.c_int32(), C.c_int64()
    abi.check("asx_poison_report", L.asx_poison_report(C.byref(slot), C.byref(cnt)))
    assert slot.value == -1 and cnt.value == 0
