"""The benchmark's workloads on the GPU against the CPU oracle.

* The curved-spacetime torus exactly as bench.py runs it: 192^3 cells over
  [-40, 40]^3 M, Kerr-Schild a = 0.9375, Fishbone-Moncrief disk (Gamma 4/3,
  K 1e-3, 2 % pressure perturbations, min beta 100), the pressure
  atmosphere (use_press_atmo, rho 1e-8, p 1e-10), Noble with Palenzuela
  backup.  One SSPRK3 stage -- asx_substep, SYNC of the state, B from A +
  con2prim (asx_con2prim_group), SYNC of the primitives -- for WENO-Z and MP5
  with HLLE and LxF, against orc_substep + oracle con2prim on the same
  bits; the headline variant also runs the second stage.  Bar: state to
  1e-12 relative (max norm), con2prim_flag and aster_mask_cc bit-exact,
  primitives to 1e-12 where Noble converged and to 1e-9 elsewhere.
* The batched workload: 400 patches with edges from {16, 24, 32, 48, 64}
  cut from that torus, one SSPRK3 step of the batch plan (one HIP graph),
  against the oracle's three stages on 8 of the patches (the smallest, the
  largest and 6 drawn at random).
"""
import ctypes as C

import numpy as np
import pytest

import oracle
from asterx_amd import abi, batch, torus
from asterx_amd.driver import all_slots
from asterx_amd.patch import Patch, StateBuf
from states import rel_err

pytestmark = pytest.mark.gpu

N = 192
SYNC_PRIMS = batch.SYNC_PRIMS
PRIM_OUT = ["rho", "velx", "vely", "velz", "eps", "press", "temperature", "entropy",
            "Bvecx", "Bvecy", "Bvecz"] + abi.ZS + abi.SAVED + abi.DB


def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch


_SRC = {}


def _source():
    """The 192^3 torus on the device, built once (bench.py's build)."""
    if "p" not in _SRC:
        torch = _torch()
        from asterx_amd import lib
        n, ng, x0, dx = torus.torus_patch_geometry(n=N)
        spec = torus.TorusSpec()
        par = torus.eos_params(spec, reconstruction_method="wenoz", flux_type="HLLE")
        slots = abi.ADM + abi.PRIMS + abi.CONS + abi.DB + abi.SAVED + abi.ZS + \
            ["con2prim_flag", "aster_mask_cc"] + abi.DBSTAG + abi.AVEC
        p = Patch(n, ng, x0, dx, slots, device="cuda")
        torus.build_torus(p, lib.backend(None), par, spec)
        torch.cuda.synchronize()
        for s in abi.DBSTAG:
            del p.arr[s]
        _SRC["p"], _SRC["spec"], _SRC["dt"] = p, spec, 0.2 * dx[0]
    return _SRC["p"], _SRC["spec"], _SRC["dt"]


def _gpu_stage(L, p, bufs, base, s, par, dt):
    """Stage s of SSPRK3 on the device patch p, state in bufs[base]
    (stage 0) or the previous stage's output, as bench.py's Stepper."""
    c0, c1, cf = batch.SSPRK3[s]
    kind = abi.BC["copy"]
    st = (C.c_int32 * 11)(*[abi.SLOT[n] for n in abi.STATE])
    pr = (C.c_int32 * 12)(*[abi.SLOT[n] for n in SYNC_PRIMS])
    out = batch.STAGE_OUT[s]
    ps = p.struct()
    u0 = batch.state_struct(p, bufs[0])
    os_ = batch.state_struct(p, bufs[out])
    abi.check("asx_substep", L.asx_substep(C.byref(ps), C.byref(u0), C.byref(os_), c0, c1,
                                           cf * dt, C.byref(par), None))
    for n in abi.STATE:
        p.arr[n] = bufs[out][n]
    ps = p.struct()
    abi.check("asx_bc_fill_many", L.asx_bc_fill_many(C.byref(ps), 11, st, kind, None))
    abi.check("asx_con2prim_group", L.asx_con2prim_group(C.byref(ps), C.byref(par), None, None))
    abi.check("asx_bc_fill_many", L.asx_bc_fill_many(C.byref(ps), 12, pr, kind, None))


def oracle_stage(p, bufs, s, par, dt):
    """The same stage on a host patch through the oracle (p holds every slot
    of all_slots(); bufs[k]: host state dicts)."""
    O = oracle.lib()
    c0, c1, cf = batch.SSPRK3[s]
    kind = abi.BC["copy"]
    out = batch.STAGE_OUT[s]
    ps = p.struct()
    u0 = batch.state_struct(p, bufs[0])
    os_ = batch.state_struct(p, bufs[out])
    abi.check("orc_substep", O.orc_substep(C.byref(ps), C.byref(u0), C.byref(os_), c0, c1,
                                           cf * dt, C.byref(par)))
    for n in abi.STATE:
        p.arr[n] = bufs[out][n]
    ps = p.struct()
    for n in abi.STATE:
        abi.check("orc_bc_fill", O.orc_bc_fill(C.byref(ps), abi.SLOT[n], kind))
    abi.check("orc_bstag_from_A", O.orc_bstag_from_A(C.byref(ps)))
    abi.check("orc_dB_from_bstag", O.orc_dB_from_bstag(C.byref(ps)))
    abi.check("orc_con2prim", O.orc_con2prim(C.byref(ps), C.byref(par)))
    for n in SYNC_PRIMS:
        abi.check("orc_bc_fill", O.orc_bc_fill(C.byref(ps), abi.SLOT[n], kind))


def host_copy(p):
    h = p.to(None)
    for s in all_slots():
        if s not in h:
            h.alloc(s)
    return h


def compare(d, h, what, state_tol=1e-12):
    """Device patch d against host patch h over the interior."""
    sl = h.interior("rho")
    fd = d.numpy("con2prim_flag")[sl]
    fh = h.numpy("con2prim_flag")[sl]
    assert np.array_equal(fd, fh), (what, "flags", int((fd != fh).sum()))
    assert np.array_equal(d.numpy("aster_mask_cc")[sl], h.numpy("aster_mask_cc")[sl]), \
        (what, "aster_mask_cc")
    bad = {}
    for n in abi.STATE:
        e = rel_err(d.numpy(n)[h.interior(n)], h.numpy(n)[h.interior(n)])
        if not e <= state_tol:
            bad[n] = e
    noble = fh == 1
    for n in PRIM_OUT:
        a, b = d.numpy(n)[sl], h.numpy(n)[sl]
        e1 = rel_err(a[noble], b[noble])
        e2 = rel_err(a, b)
        if not (e1 <= 1e-12 and e2 <= 1e-9):
            bad[n] = (e1, e2)
    assert not bad, (what, bad)
    return np.bincount(fh.astype(int).ravel(), minlength=7)


@pytest.mark.parametrize("recon,flux", [("wenoz", "HLLE"), ("wenoz", "LxF"),
                                        ("mp5", "HLLE"), ("mp5", "LxF")])
def test_torus_stage_matches_oracle(recon, flux):
    torch = _torch()
    from asterx_amd import lib
    L = lib.lib()
    src, spec, dt = _source()
    par = torus.eos_params(spec, reconstruction_method=recon, flux_type=flux)
    assert par.use_press_atmo == 1 and abs(par.gl_gamma - 4.0 / 3.0) < 1e-15
    d = Patch(src.n, src.ng, src.x0, src.dx, device="cuda")
    for n, a in src.arr.items():
        d.arr[n] = a.clone()
    h = host_copy(d)
    bd = [{n: d.arr[n] for n in abi.STATE}] + \
        [{n: torch.empty_like(d.arr[n]) for n in abi.STATE} for _ in range(2)]
    bh = [{n: h.arr[n] for n in abi.STATE}] + \
        [{n: np.zeros_like(h.arr[n]) for n in abi.STATE} for _ in range(2)]
    stages = 2 if (recon, flux) == ("wenoz", "HLLE") else 1
    for s in range(stages):
        _gpu_stage(L, d, bd, 0, s, par, dt)
        torch.cuda.synchronize()
        oracle_stage(h, bh, s, par, dt)
        print("torus %d^3 %s %s stage %d: flags %s" % (
            N, recon, flux, s, compare(d, h, (recon, flux, s)).tolist()), flush=True)
    # the stage moved the state
    assert not np.array_equal(h.numpy("dens")[h.interior("dens")],
                              src.arr["dens"][src.interior("dens")].cpu().numpy())
    del d, h, bd, bh
    torch.cuda.empty_cache()


def test_batch_400_patches_matches_oracle():
    torch = _torch()
    from asterx_amd import lib
    L = lib.lib()
    src, spec, dt = _source()
    par = torus.eos_params(spec, reconstruction_method="wenoz", flux_type="HLLE")
    boxes = batch.sample_boxes(src.n, 400, seed=2024)
    patches = [batch.cut_patch(src, n, o) for n, o in boxes]
    zones = [p.n[0] * p.n[1] * p.n[2] for p in patches]
    rng = np.random.default_rng(9)
    pick = {int(np.argmin(zones)), int(np.argmax(zones))}
    while len(pick) < 8:
        pick.add(int(rng.integers(0, len(patches))))
    hosts = {i: host_copy(patches[i]) for i in sorted(pick)}
    torch.cuda.synchronize()
    plan = batch.Plan(L, patches, par, dt)
    plan.step(1, graph=True)
    torch.cuda.synchronize()
    for i, h in hosts.items():
        bh = [{n: h.arr[n] for n in abi.STATE}] + \
            [{n: np.zeros_like(h.arr[n]) for n in abi.STATE} for _ in range(2)]
        for s in range(3):
            oracle_stage(h, bh, s, par, dt)
        for n in abi.STATE:
            h.arr[n] = bh[0][n]
        fl = compare(patches[i], h, ("patch", i, patches[i].n), state_tol=1e-11)
        print("batch patch %d %s: flags %s" % (i, patches[i].n, fl.tolist()), flush=True)
    plan.close()
    del patches, plan
    torch.cuda.empty_cache()
