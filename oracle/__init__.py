"""CPU oracle of the GRMHD substep: test infrastructure only.

Loaded by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg,
never by the product path (asterx_amd).  See orc_physics.hpp for what each
function restates and how it is pinned.
"""
import ctypes as C
import os

from asterx_amd import abi

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(_HERE, "liboracle.so")

_P = C.POINTER(abi.asx_patch)
_PP = C.POINTER(abi.asx_params)
_S = C.POINTER(abi.asx_state)
_D = C.POINTER(C.c_double)

_SIGS = {
    "orc_fluxes": [_P, _PP, C.c_int],
    "orc_source_terms": [_P, _PP],
    "orc_rhs": [_P, _PP],
    "orc_tmunu": [_P, _PP],
    "orc_estimate_error": [_P, _PP, C.c_int, C.c_void_p],
    "orc_bstag_from_A": [_P],
    "orc_bstag_from_A_halo": [_P, C.c_int],
    "orc_dB_from_bstag": [_P],
    "orc_B_from_dB": [_P],
    "orc_prim2con": [_P],
    "orc_check_prims": [_P, _PP],
    "orc_con2prim": [_P, _PP],
    "orc_state_lincomb": [_P, _S, C.c_double, C.c_int, _D, C.POINTER(_S)],
    "orc_bc_fill": [_P, C.c_int, C.c_int],
    "orc_substep": [_P, _S, _S, C.c_double, C.c_double, C.c_double, _PP],
    "orc_substep_bc": [_P, _S, _S, C.c_double, C.c_double, C.c_double, _PP, C.c_int],
    "orc_c2p_batch": [C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p,
                      C.c_void_p, C.c_void_p, C.c_void_p, _PP],
    "orc_prim2con_batch": [C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, _PP],
    "orc_c2p_batch_ye": [C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p,
                         C.c_void_p, C.c_void_p, C.c_void_p, _PP],
    "orc_prim2con_batch_ye": [C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p,
                              C.c_void_p, _PP],
    "orc_eos_from_rho_temp": [C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p,
                              C.c_void_p, _PP],
    "orc_eos_from_rho_eps": [C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p,
                             C.c_void_p, _PP],
    "orc_tab_interp": [C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, _PP],
    "orc_reconstruct": [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, _PP,
                        C.c_void_p],
    "orc_util": [C.c_int, C.c_void_p, C.c_void_p],
    "orc_set_threads": [C.c_int],
    "asx_gf_centering": [C.c_int],
}

_lib = None


def build():
    """Compile liboracle.so from the sources in this folder."""
    import subprocess
    subprocess.check_call(["make", "-s", "-C", _HERE, "-B", "liboracle.so"])


def build_planted(k, outdir):
    """Compile the oracle with planted mistake number k (ORC_PLANT=k, see
    orc_grid.cpp) into outdir/liboracle_plant<k>.so and return that path.
    Tests use it to show that a pin fails when the restatement is wrong."""
    import subprocess
    out = os.path.join(outdir, "liboracle_plant%d.so" % k)
    subprocess.check_call(["g++", "-O1", "-std=c++17", "-fPIC", "-ffp-contract=off",
                           "-fno-fast-math", "-fopenmp", "-DORC_PLANT=%d" % k,
                           "-I" + os.path.join(_HERE, "..", "include"), "-shared", "-o", out,
                           os.path.join(_HERE, "orc_grid.cpp")])
    return out


def load(path):
    """A ctypes handle of an oracle build at `path` with the signatures set."""
    L = C.CDLL(path)
    for name, args in _SIGS.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = C.c_int
    return L


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            build()
        _lib = load(LIBPATH)
    return _lib


def backend():
    from asterx_amd.driver import Backend
    return Backend(lib(), "orc_")
