"""Batched patches: many patches of uneven sizes cut from one source patch,
and the plan that runs whole SSPRK3 steps over all of them in one set of
launches (asx_batch_* in the C ABI).

The plan replaces the driver loop that calls the scheduled functions once per
local box and RK stage (AsterX/schedule.ccl:119-232); the per-patch work is
that of asx_substep + boundary fill + asx_con2prim_group + boundary fill.
"""
import ctypes as C

import numpy as np

from . import abi
from .patch import Patch

EDGES = (16, 24, 32, 48, 64)
# SSPRK3 stages (c0, c1, c2/dt) and the state buffers they read and write
SSPRK3 = [(0.0, 1.0, 1.0), (0.75, 0.25, 0.25), (1.0 / 3.0, 2.0 / 3.0, 2.0 / 3.0)]
STAGE_IN = (0, 1, 2)
STAGE_OUT = (1, 2, 0)
SYNC_PRIMS = ["rho", "velx", "vely", "velz", "eps", "press", "temperature", "entropy",
              "Ye", "Bvecx", "Bvecy", "Bvecz"]
# grid functions a batch patch carries (the plan owns fluxes, aux, dB_stag)
PATCH_SLOTS = abi.ADM + abi.PRIMS + abi.CONS + abi.DB + abi.SAVED + abi.ZS + \
    ["con2prim_flag", "aster_mask_cc"] + abi.AVEC


def sample_boxes(src_n, npatch, seed, edges=EDGES):
    """npatch (n, o) pairs: cell counts per axis drawn from `edges`, and the
    source interior cell o at which the patch's interior starts, uniform
    over the offsets that keep the patch and its ghosts inside the source."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(npatch):
        n = tuple(int(rng.choice(edges)) for _ in range(3))
        if any(n[d] > src_n[d] for d in range(3)):
            raise ValueError("patch edge larger than the source patch")
        o = tuple(int(rng.integers(0, src_n[d] - n[d] + 1)) for d in range(3))
        out.append((n, o))
    return out


def cut_patch(src, n, o, names=None):
    """Patch of n cells whose interior starts at source interior cell o, with
    the source's ghost width, every named array (default: all the source
    holds of PATCH_SLOTS) copied from the source, ghosts included."""
    x0 = tuple(src.x0[d] + o[d] * src.dx[d] for d in range(3))
    p = Patch(n, src.ng, x0, src.dx, device=src.device)
    for name in (names or [s for s in PATCH_SLOTS if s in src]):
        ex, ey, ez = abi.extents(n, src.ng, abi.centering(name))
        a = src.arr[name][o[2]:o[2] + ez, o[1]:o[1] + ey, o[0]:o[0] + ex]
        if src.device is None:
            p.arr[name] = np.ascontiguousarray(a)
        else:
            p.arr[name] = a.contiguous().clone()
    return p


def state_struct(patch, arrays):
    """asx_state over the 11 state arrays `arrays` (dict) of patch's shape."""
    s = abi.asx_state()
    for q, name in enumerate(abi.STATE):
        ex, ey, _ = abi.extents(patch.n, patch.ng, abi.centering(name))
        s.gf[q].data = patch._ptr(arrays[name])
        s.gf[q].sj = ex
        s.gf[q].sk = ex * ey
    return s


class Plan:
    """asx_batch plan over `patches` (device Patch objects).  Allocates the
    two extra state buffers per patch; buffer 0 is the patches' own state."""

    def __init__(self, lib, patches, par, dt, bc="copy"):
        import torch
        self.L = lib
        self.patches = patches
        self.par = par
        npch = len(patches)
        self.bufs = []
        ps = (abi.asx_patch * npch)()
        bs = (abi.asx_state * (2 * npch))()
        for i, p in enumerate(patches):
            ps[i] = p.struct()
            for k in range(2):
                d = {name: torch.empty_like(p.arr[name]) for name in abi.STATE}
                self.bufs.append(d)
                bs[2 * i + k] = state_struct(p, d)
        self._ps, self._bs = ps, bs
        self.h = C.c_void_p()
        abi.check("asx_batch_create", lib.asx_batch_create(
            npch, ps, bs, C.byref(par), float(dt), abi.BC[bc], C.byref(self.h)))

    def buffer(self, i, k):
        """State arrays of buffer k (0, 1, 2) of patch i."""
        if k == 0:
            return {name: self.patches[i].arr[name] for name in abi.STATE}
        return self.bufs[2 * i + k - 1]

    def stage(self, s, stream=None):
        abi.check("asx_batch_stage", self.L.asx_batch_stage(self.h, s, stream))

    def step(self, nsteps=1, graph=True, stream=None, persistent=False):
        """nsteps SSPRK3 steps: the captured graph (graph=True), plain
        launches (graph=False), or one persistent launch per stage
        (persistent=True)."""
        mode = 2 if persistent else int(graph)
        abi.check("asx_batch_step", self.L.asx_batch_step(self.h, nsteps, mode, stream))

    def autotune(self, reps=3, stream=None):
        abi.check("asx_batch_autotune", self.L.asx_batch_autotune(self.h, reps, stream))

    def variants(self):
        """The plan's launch shapes: {face_x, face_y, face_z, update}."""
        out = {}
        for name in ("face_x", "face_y", "face_z", "update"):
            v = C.c_int32()
            abi.check("asx_batch_get_variant", self.L.asx_batch_get_variant(
                self.h, abi.KCLASS[name], C.byref(v)))
            out[name] = v.value
        return out

    def info(self):
        z, b = C.c_int64(), C.c_int64()
        abi.check("asx_batch_info", self.L.asx_batch_info(self.h, C.byref(z), C.byref(b)))
        return z.value, b.value

    def close(self):
        if self.h:
            abi.check("asx_batch_destroy", self.L.asx_batch_destroy(self.h))
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def single_step(lib, patch, bufs, par, dt, bc="copy", stream=None):
    """One SSPRK3 step of one patch through the single-patch entry points,
    in the plan's order: stage s runs asx_substep from buffer STAGE_IN[s]
    into STAGE_OUT[s] (u0 = buffer 0), fills the state's ghosts, runs
    asx_con2prim_group and fills the synced primitives' ghosts.  bufs[k] is
    the dict of state arrays of buffer k (bufs[0] = the patch's own)."""
    kind = abi.BC[bc]
    st = [abi.SLOT[n] for n in abi.STATE]
    pr = [abi.SLOT[n] for n in SYNC_PRIMS]
    st_c = (C.c_int32 * len(st))(*st)
    pr_c = (C.c_int32 * len(pr))(*pr)
    own = dict(bufs[0])
    for s in range(3):
        c0, c1, cf = SSPRK3[s]
        for name in abi.STATE:
            patch.arr[name] = bufs[STAGE_IN[s]][name]
        ps = patch.struct()
        u0 = state_struct(patch, bufs[0])
        out = state_struct(patch, bufs[STAGE_OUT[s]])
        abi.check("asx_substep", lib.asx_substep(C.byref(ps), C.byref(u0), C.byref(out),
                                                 c0, c1, cf * dt, C.byref(par), stream))
        for name in abi.STATE:
            patch.arr[name] = bufs[STAGE_OUT[s]][name]
        ps = patch.struct()
        abi.check("asx_bc_fill_many", lib.asx_bc_fill_many(C.byref(ps), len(st), st_c, kind,
                                                           stream))
        abi.check("asx_con2prim_group", lib.asx_con2prim_group(C.byref(ps), C.byref(par),
                                                               None, stream))
        abi.check("asx_bc_fill_many", lib.asx_bc_fill_many(C.byref(ps), len(pr), pr_c, kind,
                                                           stream))
    for name in abi.STATE:
        patch.arr[name] = own[name]
