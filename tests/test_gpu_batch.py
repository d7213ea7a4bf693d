"""Batched patches on the GPU.

Patches of uneven sizes are cut from a torus patch (Kerr-Schild a=0.9375,
Fishbone-Moncrief disk, WENO-Z + HLLE, Noble/Palenzuela con2prim).  One
SSPRK3 step of the batch plan must agree bit for bit with the single-patch
entry points (asx_substep, asx_bc_fill_many, asx_con2prim_group) run patch by
patch in the same order -- every grid function and all three state buffers.
The single-patch path is itself pinned to the oracle (test_gpu_parity.py), so
this closes the chain.  Graph replays must agree with plain launches on every
replay, results must not depend on the launch-shape variants, and the
autotuner must leave the evolved state alone.
"""
import ctypes as C

import pytest

from asterx_amd import abi, batch, torus

pytestmark = pytest.mark.gpu

EDGES = (8, 12, 16, 20)
NPATCH = 7


def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch


_SRC = {}


def _source():
    """A 48^3 torus patch over [-20, 20]^3 (built once per session)."""
    if "src" not in _SRC:
        torch = _torch()
        from asterx_amd import lib
        from asterx_amd.patch import Patch
        be = lib.backend(None)
        n, ng, x0, dx = torus.torus_patch_geometry(n=48, L=20.0)
        spec = torus.TorusSpec()
        par = torus.eos_params(spec, reconstruction_method="wenoz", flux_type="HLLE")
        p = Patch(n, ng, x0, dx, batch.PATCH_SLOTS + abi.DBSTAG, device="cuda")
        torus.build_torus(p, be, par, spec)
        torch.cuda.synchronize()
        _SRC["src"] = (p, par, 0.2 * dx[0])
    return _SRC["src"]


def _cuts(seed=3, names=None):
    src, par, dt = _source()
    boxes = batch.sample_boxes(src.n, NPATCH, seed, edges=EDGES)
    return [batch.cut_patch(src, n, o, names) for n, o in boxes], par, dt


def _L():
    from asterx_amd import lib
    return lib.lib()


def _eq(x, y):
    """Bitwise-equal values, NaN matching NaN."""
    torch = _torch()
    return torch.equal(x, y) or bool(((x == y) | (torch.isnan(x) & torch.isnan(y))).all())


def _assert_same(pa, pb, what):
    for i, (a, b) in enumerate(zip(pa, pb)):
        assert set(a.arr) == set(b.arr)
        for name in a.arr:
            x, y = a.arr[name], b.arr[name]
            assert _eq(x, y), (what, i, a.n, name, float((x - y).abs().nan_to_num(0).max()))


def test_plan_matches_single_patch_path():
    torch = _torch()
    # with the staggered field slots: the plan writes them as
    # asx_con2prim_group does
    names = batch.PATCH_SLOTS + abi.DBSTAG
    pa, par, dt = _cuts(names=names)
    pb, _, _ = _cuts(names=names)
    L = _L()
    plan = batch.Plan(L, pa, par, dt)
    zones, nbytes = plan.info()
    assert zones == sum(p.n[0] * p.n[1] * p.n[2] for p in pa)
    assert nbytes > 0
    plan.step(1, graph=False)
    bufs_b = []
    for p in pb:
        bufs = [{n: p.arr[n] for n in abi.STATE}]
        bufs += [{n: torch.empty_like(p.arr[n]) for n in abi.STATE} for _ in range(2)]
        batch.single_step(L, p, bufs, par, dt)
        bufs_b.append(bufs)
    torch.cuda.synchronize()
    _assert_same(pa, pb, "grid functions after one step")
    for i in range(len(pa)):
        for k in (1, 2):
            ga, gb = plan.buffer(i, k), bufs_b[i][k]
            for name in abi.STATE:
                assert _eq(ga[name], gb[name]), ("buffer", k, i, name)
    # the step did something: the state moved and stayed finite
    for p in pa:
        for name in abi.CONS:
            assert bool(torch.isfinite(p.arr[name]).all()), name
    plan.close()


def test_graph_replays_match_plain_launches():
    torch = _torch()
    pa, par, dt = _cuts(seed=5)
    pb, _, _ = _cuts(seed=5)
    L = _L()
    ga = batch.Plan(L, pa, par, dt)
    gb = batch.Plan(L, pb, par, dt)
    before = pa[0].arr["dens"].clone()
    for r in range(3):
        ga.step(1, graph=True)
        gb.step(1, graph=False)
        torch.cuda.synchronize()
        _assert_same(pa, pb, "replay %d" % r)
    assert not torch.equal(before, pa[0].arr["dens"])
    ga.close()
    gb.close()


@pytest.mark.parametrize("recon", ["wenoz", "mp5"])
def test_persistent_stage_matches_plain_launches(recon):
    """One persistent cooperative launch per stage (asx_batch_step mode 2)
    against plain launches with the same launch shapes (face variant 0,
    update variant 1), bit for bit, over three steps."""
    torch = _torch()
    pa, par, dt = _cuts(seed=7)
    pb, _, _ = _cuts(seed=7)
    par = abi.asx_params.from_buffer_copy(par)
    par.reconstruction_method = abi.RECON[recon]
    L = _L()
    classes = [abi.KCLASS[k] for k in ("face_x", "face_y", "face_z", "update")]
    saved = []
    for kc in classes:
        v = C.c_int32()
        abi.check("asx_get_variant", L.asx_get_variant(kc, C.byref(v)))
        saved.append(v.value)
    try:
        for kc, v in zip(classes, (0, 0, 0, 1)):
            abi.check("asx_set_variant", L.asx_set_variant(kc, v))
        ga = batch.Plan(L, pa, par, dt)
        gb = batch.Plan(L, pb, par, dt)
        before = pa[0].arr["dens"].clone()
        for r in range(3):
            ga.step(1, persistent=True)
            gb.step(1, graph=False)
            torch.cuda.synchronize()
            _assert_same(pa, pb, "persistent step %d" % r)
        assert not torch.equal(before, pa[0].arr["dens"])
        ga.close()
        gb.close()
    finally:
        for kc, v in zip(classes, saved):
            abi.check("asx_set_variant", L.asx_set_variant(kc, v))


def test_persistent_stage_refuses_unsupported():
    """Methods the persistent stage is not built for are refused, not run."""
    _torch()
    pa, par, dt = _cuts(seed=9)
    par = abi.asx_params.from_buffer_copy(par)
    par.reconstruction_method = abi.RECON["minmod"]
    plan = batch.Plan(_L(), pa, par, dt)
    st = _L().asx_batch_step(plan.h, 1, 2, None)
    assert abi.STATUS[st] == "ASX_ERR_UNSUPPORTED"
    st = _L().asx_batch_step(plan.h, 1, 3, None)
    assert abi.STATUS[st] == "ASX_ERR_PARAM"
    plan.close()


def test_results_independent_of_variants():
    torch = _torch()
    L = _L()
    classes = [abi.KCLASS[k] for k in ("face_x", "face_y", "face_z", "update")]
    saved = []
    for kc in classes:
        v = C.c_int32()
        abi.check("asx_get_variant", L.asx_get_variant(kc, C.byref(v)))
        saved.append(v.value)
    try:
        runs = []
        for pick in (0, 1):
            for kc in classes:
                nv = L.asx_variant_count(kc)
                abi.check("asx_set_variant", L.asx_set_variant(kc, (nv - 1) if pick else 0))
            pp, par, dt = _cuts(seed=8)
            plan = batch.Plan(L, pp, par, dt)
            plan.step(1, graph=True)
            torch.cuda.synchronize()
            plan.close()
            runs.append(pp)
        _assert_same(runs[0], runs[1], "variants")
    finally:
        for kc, v in zip(classes, saved):
            abi.check("asx_set_variant", L.asx_set_variant(kc, v))


def test_autotune_leaves_state_alone():
    torch = _torch()
    L = _L()
    pa, par, dt = _cuts(seed=11)
    snap = [{n: a.clone() for n, a in p.arr.items()} for p in pa]
    plan = batch.Plan(L, pa, par, dt)
    plan.autotune(reps=1)
    torch.cuda.synchronize()
    for p, s in zip(pa, snap):
        for name, a in p.arr.items():
            assert _eq(a, s[name]), name
    for k, v in plan.variants().items():
        assert 0 <= v < L.asx_variant_count(abi.KCLASS[k])
    plan.step(1, graph=True)
    torch.cuda.synchronize()
    plan.close()
