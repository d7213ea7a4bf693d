"""Host side of the batched-patch plan: patch sampling and cutting, and the
plan's argument checks (which return before any device call)."""
import ctypes as C

import numpy as np
import pytest

from asterx_amd import abi, batch, torus
from asterx_amd.patch import Patch


def test_sample_boxes_fit_and_cover_edges():
    src_n = (64, 64, 64)
    boxes = batch.sample_boxes(src_n, 400, seed=1)
    assert len(boxes) == 400
    seen = set()
    for n, o in boxes:
        for d in range(3):
            assert n[d] in batch.EDGES
            assert 0 <= o[d] and o[d] + n[d] <= src_n[d]
            seen.add(n[d])
    assert seen == set(batch.EDGES)
    assert boxes == batch.sample_boxes(src_n, 400, seed=1)
    with pytest.raises(ValueError):
        batch.sample_boxes((8, 8, 8), 3, seed=0, edges=(16,))


def test_cut_patch_is_a_subblock():
    n, ng = (10, 9, 8), 3
    src = Patch(n, ng, (-1.0, -2.0, -3.0), (0.1, 0.2, 0.3))
    rng = np.random.default_rng(0)
    for name in ("rho", "alp", "Avec_x", "Avec_z", "dens"):
        src.arr[name] = rng.standard_normal(src.shape(name))
    p = batch.cut_patch(src, (4, 5, 3), (2, 1, 5), ["rho", "alp", "Avec_x", "Avec_z"])
    assert p.n == (4, 5, 3) and p.ng == ng
    assert np.allclose(p.x0, (-1.0 + 0.2, -2.0 + 0.2, -3.0 + 1.5))
    # the cell-centred point (i, j, k) of the cut is source point (i+2, j+1, k+5)
    assert np.array_equal(p["rho"], src["rho"][5:5 + 9, 1:1 + 11, 2:2 + 10])
    assert np.array_equal(p["alp"], src["alp"][5:5 + 10, 1:1 + 12, 2:2 + 11])
    assert np.array_equal(p["Avec_x"], src["Avec_x"][5:5 + 10, 1:1 + 12, 2:2 + 10])
    assert np.array_equal(p["Avec_z"], src["Avec_z"][5:5 + 9, 1:1 + 12, 2:2 + 11])
    # same physical coordinates
    for cent in (0, 7, 6):
        zs, ys, xs = src.coords(cent)
        zp, yp, xp = p.coords(cent)
        assert np.allclose(xp.ravel(), xs.ravel()[2:2 + xp.size])
        assert np.allclose(zp.ravel(), zs.ravel()[5:5 + zp.size])


def _host_patches(k=2):
    out = []
    for n in ((8, 12, 8), (12, 8, 16))[:k]:
        p = Patch(n, 3, (0.0, 0.0, 0.0), (0.5, 0.5, 0.5), batch.PATCH_SLOTS)
        out.append(p)
    return out


def _create(L, patches, par, bufs=None, bc="copy", npatch=None):
    ps = (abi.asx_patch * max(1, len(patches)))()
    for i, p in enumerate(patches):
        ps[i] = p.struct()
    bs = (abi.asx_state * max(2, 2 * len(patches)))()
    keep = []
    for i, p in enumerate(patches):
        for k in range(2):
            d = bufs[2 * i + k] if bufs else {n: np.zeros(p.shape(n)) for n in abi.STATE}
            keep.append(d)
            bs[2 * i + k] = batch.state_struct(p, d)
    h = C.c_void_p()
    kind = abi.BC[bc] if isinstance(bc, str) else bc
    r = L.asx_batch_create(len(patches) if npatch is None else npatch, ps, bs, C.byref(par),
                           0.01, kind, C.byref(h))
    assert not h.value
    return r


def test_plan_argument_checks():
    from asterx_amd import lib
    L = lib.lib()
    par = torus.eos_params(torus.TorusSpec(), reconstruction_method="wenoz", flux_type="HLLE")
    ps = _host_patches()
    assert _create(L, ps, par, npatch=0) == 7  # ASX_ERR_SIZE
    assert _create(L, ps, par, bc=9) == 3  # ASX_ERR_PARAM
    bad = abi.asx_params.from_buffer_copy(par)
    bad.use_uct = 1
    assert _create(L, ps, bad) == 4  # ASX_ERR_UNSUPPORTED
    # a missing grid function
    del ps[1].arr["saved_rho"]
    assert _create(L, ps, par) == 1  # ASX_ERR_NULL
    # a state buffer whose strides do not match its centering
    ps = _host_patches()
    bufs = [{n: np.zeros(p.shape(n)) for n in abi.STATE} for p in ps for _ in range(2)]
    bs = (abi.asx_state * 4)()
    for i, p in enumerate(ps):
        for k in range(2):
            bs[2 * i + k] = batch.state_struct(p, bufs[2 * i + k])
    bs[3].gf[abi.STATE.index("Avec_x")].sj += 1
    pa = (abi.asx_patch * 2)(*[p.struct() for p in ps])
    h = C.c_void_p()
    assert L.asx_batch_create(2, pa, bs, C.byref(par), 0.01, abi.BC["copy"],
                              C.byref(h)) == 2  # ASX_ERR_LAYOUT
    assert not h.value
    # too few ghost zones for the fused substep
    p1 = Patch((8, 8, 8), 2, (0.0,) * 3, (0.5,) * 3, batch.PATCH_SLOTS)
    assert _create(L, [p1], par) == 5  # ASX_ERR_GHOSTS
    # patches disagree on carrying dB_stag
    ps = _host_patches()
    for s in abi.DBSTAG:
        ps[0].alloc(s)
    assert _create(L, ps, par) == 1
    assert L.asx_batch_create(1, None, None, C.byref(par), 0.01, 2, C.byref(h)) == 1
    assert L.asx_batch_destroy(None) == 0
    assert L.asx_batch_stage(None, 0, None) == 1
    assert L.asx_batch_step(None, 1, 1, None) == 1
    assert L.asx_batch_autotune(None, 1, None) == 1
    assert L.asx_batch_info(None, None, None) == 1
    v = C.c_int32()
    assert L.asx_batch_get_variant(None, abi.KCLASS["update"], C.byref(v)) == 1
