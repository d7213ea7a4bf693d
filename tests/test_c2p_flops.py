"""Algorithmic FLOPs of con2prim, pinned by the flop-counting oracle build.

oracle/flops.cpp compiles the oracle's con2prim with every double replaced
by a counting type (+, -, *, /, sqrt = 1 each; comparisons, fabs, min/max
= 0).  On the learned-guess workload (ideal gas Gamma 5/3, tolerance 1e-10,
Noble with Palenzuela backup) the counting build must reproduce the double
build's flags and iteration counts exactly, every converged Noble iteration
costs exactly C2P_FLOPS_PER_ITER, and the rest of a zone (setup, seeds,
floors, primitive recovery) averages C2P_FLOPS_SETUP.  bench.py prices its
con2prim FP64 roofline with these two numbers.
"""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import oracle
from asterx_amd import abi, zones

HERE = os.path.dirname(os.path.abspath(__file__))
ORC = os.path.join(os.path.dirname(HERE), "oracle")
C2P_FLOPS_PER_ITER = 95
C2P_FLOPS_SETUP = 445


def _counting_lib():
    subprocess.check_call(["make", "-s", "-C", ORC, "_flops/libc2pflops.so"])
    L = C.CDLL(os.path.join(ORC, "_flops", "libc2pflops.so"))
    L.c2p_flops.argtypes = [C.c_int64] + [C.c_void_p] * 3 + [C.POINTER(abi.asx_params)] + \
        [C.c_void_p] * 4
    return L


@pytest.mark.parametrize("cold", [False, True])
def test_c2p_flops_pinned(cold):
    L = _counting_lib()
    nz = 4096
    par = abi.params(gl_gamma=5.0 / 3.0, eps_max=1e8, rho_max=1e8, rho_abs_min=1e-12,
                     vw_lim=1e3, B_lim=1e3, c2p_tol=1e-10)
    cons, g, seed, _ = zones.make_batch(nz, seed=3, cold=cold)
    cons, g, seed = (np.ascontiguousarray(a) for a in (cons, g, seed))
    fl, tr, fg = np.zeros(nz), np.zeros(nz), np.zeros(nz)
    it = np.zeros(nz, dtype=np.int32)
    abi.check("c2p_flops", L.c2p_flops(nz, cons.ctypes.data, g.ctypes.data, seed.ctypes.data,
                                       C.byref(par), fl.ctypes.data, tr.ctypes.data,
                                       it.ctypes.data, fg.ctypes.data))
    prim_h, flag_h = np.zeros((8, nz)), np.zeros(nz)
    it_h = np.zeros(nz, dtype=np.int32)
    abi.check("orc_c2p_batch", oracle.lib().orc_c2p_batch(
        nz, cons.ctypes.data, g.ctypes.data, seed.ctypes.data, prim_h.ctypes.data,
        flag_h.ctypes.data, it_h.ctypes.data, C.byref(par)))
    assert np.array_equal(fg, flag_h) and np.array_equal(it, it_h)
    m = fg == 1
    assert m.mean() > 0.99
    # one more Noble iteration costs exactly C2P_FLOPS_PER_ITER: the cheapest
    # zone at k + 1 iterations costs that much more than at k
    lo = {k: fl[m & (it == k)].min() for k in np.unique(it[m]) if (m & (it == k)).sum() > 20}
    ks = sorted(lo)
    for a, b in zip(ks, ks[1:]):
        if b == a + 1:
            assert lo[b] - lo[a] == C2P_FLOPS_PER_ITER, (a, lo[a], lo[b])
    setup = (fl[m] - C2P_FLOPS_PER_ITER * it[m]).mean()
    assert abs(setup - C2P_FLOPS_SETUP) < 5, setup
    # the ideal-gas path calls pow/log a fixed handful of times per zone
    assert tr[m].max() <= 8
