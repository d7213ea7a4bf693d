"""A single patch's substeps captured in a HIP graph (stream capture of the
C-ABI calls: asx_substep, the state fill, asx_con2prim_group, the primitive
fill) replay bit-identically to the same substeps launched one by one.  The
graph holds six substeps (two SSPRK3 steps, after which the three state
buffers are back in their places); it is replayed three times, with eager
work and synchronisations between the replays, and every replay is checked
against the eager run."""
import ctypes as C

import numpy as np
import pytest

from asterx_amd import abi
from states import make_state

pytestmark = pytest.mark.gpu

SSPRK3 = [(0.0, 1.0, 1.0), (0.75, 0.25, 0.25), (1.0 / 3.0, 2.0 / 3.0, 2.0 / 3.0)]
SYNC_PRIMS = ["rho", "velx", "vely", "velz", "eps", "press", "temperature", "entropy",
              "Ye", "Bvecx", "Bvecy", "Bvecz"]


def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch


class Run:
    """Substeps on one device patch with three rotating state buffers."""

    def __init__(self, d, par, dt, stream):
        import torch
        from asterx_amd import lib
        self.d, self.par, self.dt = d, par, dt
        self.L = lib.lib()
        self.sh = C.c_void_p(stream.cuda_stream)
        self.bufs = [{n: d[n] for n in abi.STATE}]
        for _ in range(2):
            self.bufs.append({n: torch.empty_like(d[n]) for n in abi.STATE})
        self.cur = self.base = self.stage = 0
        self.ss = (C.c_int32 * 11)(*[abi.SLOT[n] for n in abi.STATE])
        self.sp = (C.c_int32 * 12)(*[abi.SLOT[n] for n in SYNC_PRIMS])

    def state(self, k):
        s = abi.asx_state()
        for q, n in enumerate(abi.STATE):
            ex, ey, _ = abi.extents(self.d.n, self.d.ng, abi.centering(n))
            s.gf[q].data = self.d._ptr(self.bufs[k][n])
            s.gf[q].sj, s.gf[q].sk = ex, ex * ey
        return s

    def substep(self):
        c0, c1, cf = SSPRK3[self.stage]
        if self.stage == 0:
            self.base = self.cur
        out = [i for i in range(3) if i not in (self.cur, self.base)][0]
        u0, o = self.state(self.base), self.state(out)
        abi.check("asx_substep", self.L.asx_substep(
            C.byref(self.d.struct()), C.byref(u0), C.byref(o), c0, c1, cf * self.dt,
            C.byref(self.par), self.sh))
        for n in abi.STATE:
            self.d.arr[n] = self.bufs[out][n]
        self.cur = out
        ps = self.d.struct()
        abi.check("asx_bc_fill_many", self.L.asx_bc_fill_many(
            C.byref(ps), 11, self.ss, abi.BC["copy"], self.sh))
        abi.check("asx_con2prim_group", self.L.asx_con2prim_group(
            C.byref(ps), C.byref(self.par), None, self.sh))
        abi.check("asx_bc_fill_many", self.L.asx_bc_fill_many(
            C.byref(ps), 12, self.sp, abi.BC["copy"], self.sh))
        self.stage = (self.stage + 1) % 3

    def snapshot(self):
        return {n: self.d.numpy(n).copy() for n in list(abi.STATE) + SYNC_PRIMS}


def test_graph_replay_matches_eager_substeps():
    torch = _torch()
    h = make_state(curved=True, uct=False)
    par = abi.params(reconstruction_method="wenoz", flux_type="HLLE", gl_gamma=5.0 / 3.0)
    dt = 0.2 * min(h.dx)
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        eager = Run(h.to("cuda"), par, dt, stream)
        graph = Run(h.to("cuda"), par, dt, stream)
        # warm-up (library scratch and queue allocated on first use), as
        # before any capture
        eager.substep()
        graph.substep()
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=stream):
        for _ in range(6):
            graph.substep()
    assert (graph.cur, graph.base, graph.stage) == (eager.cur, eager.base, eager.stage)
    for rep in range(3):
        with torch.cuda.stream(stream):
            for _ in range(6):
                eager.substep()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        a, b = eager.snapshot(), graph.snapshot()
        for n in a:
            assert np.array_equal(a[n].view(np.uint64), b[n].view(np.uint64)), (rep, n)
            assert np.isfinite(a[n][h.interior(n)]).all(), (rep, n)


def test_release_scratch_then_substeps_again():
    """asx_release_scratch frees the stream's face scratch and con2prim
    queue; the next calls allocate them again and give the same bits."""
    torch = _torch()
    from asterx_amd import lib
    h = make_state(curved=True, uct=False)
    par = abi.params(reconstruction_method="wenoz", flux_type="HLLE", gl_gamma=5.0 / 3.0)
    dt = 0.2 * min(h.dx)
    stream = torch.cuda.Stream()
    snaps = []
    for release in (False, True):
        with torch.cuda.stream(stream):
            r = Run(h.to("cuda"), par, dt, stream)
            for q in range(3):
                r.substep()
                if release:
                    abi.check("asx_release_scratch",
                              lib.lib().asx_release_scratch(r.sh))
        torch.cuda.synchronize()
        snaps.append(r.snapshot())
    for n in snaps[0]:
        assert np.array_equal(snaps[0][n].view(np.uint64), snaps[1][n].view(np.uint64)), n
    # a stream the library holds nothing for: nothing to free
    abi.check("asx_release_scratch", lib.lib().asx_release_scratch(None))
