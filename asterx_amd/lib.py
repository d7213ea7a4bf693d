"""Loader of the HIP library (libasterx_hip.so) with ctypes signatures.

The product path runs only through this library: if it is missing the
import fails loudly (there is no CPU fallback).
"""
import ctypes as C
import os
import subprocess

from . import abi

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(_HERE, "libasterx_hip.so")

_P = C.POINTER(abi.asx_patch)
_PP = C.POINTER(abi.asx_params)
_S = C.POINTER(abi.asx_state)
_CNT = C.POINTER(abi.asx_counters)
_V = C.c_void_p

SIGNATURES = {
    "asx_fluxes": ([_P, _PP, C.c_int32, _CNT, _V], C.c_int),
    "asx_source_terms": ([_P, _PP, _V], C.c_int),
    "asx_rhs": ([_P, _PP, _V], C.c_int),
    "asx_tmunu": ([_P, _PP, _V], C.c_int),
    "asx_estimate_error": ([_P, _PP, C.c_int32, C.c_void_p, _V], C.c_int),
    "asx_bstag_from_A": ([_P, _V], C.c_int),
    "asx_dB_from_bstag": ([_P, _V], C.c_int),
    "asx_B_from_dB": ([_P, _V], C.c_int),
    "asx_prim2con": ([_P, _V], C.c_int),
    "asx_con2prim": ([_P, _PP, _CNT, _V], C.c_int),
    "asx_con2prim_group": ([_P, _PP, _CNT, _V], C.c_int),
    "asx_con2prim_group_noalloc": ([_P, _PP, _CNT, _V], C.c_int),
    "asx_state_lincomb": ([_P, _S, C.c_double, C.c_int32, C.POINTER(C.c_double),
                           C.POINTER(_S), _V], C.c_int),
    "asx_bc_fill": ([_P, C.c_int32, C.c_int32, _V], C.c_int),
    "asx_bc_fill_many": ([_P, C.c_int32, C.POINTER(C.c_int32), C.c_int32, _V], C.c_int),
    "asx_release_scratch": ([_V], C.c_int),
    "asx_set_poison": ([C.c_int32], C.c_int),
    "asx_poison_report": ([C.POINTER(C.c_int32), C.POINTER(C.c_int64)], C.c_int),
    "asx_poison_check": ([_P, C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                          C.POINTER(C.c_int64), _V], C.c_int),
    "asx_prolongate": ([_P, _P, C.c_int32, C.POINTER(C.c_int32), C.c_int32, C.c_int32, _V],
                       C.c_int),
    "asx_restrict": ([_P, _P, C.c_int32, C.POINTER(C.c_int32), _V], C.c_int),
    "asx_regrid_boxes": ([_P, C.c_double, C.c_int32, C.POINTER(C.c_int32),
                          C.POINTER(C.c_int32), C.c_double, C.c_int32,
                          C.POINTER(abi.asx_box), C.POINTER(C.c_int32), _V], C.c_int),
    "asx_timing_enable": ([C.c_int32], C.c_int),
    "asx_timing_read": ([C.c_int32, C.POINTER(C.c_int64), C.POINTER(C.c_double)],
                        C.c_int),
    "asx_variant_count": ([C.c_int32], C.c_int32),
    "asx_set_variant": ([C.c_int32, C.c_int32], C.c_int),
    "asx_get_variant": ([C.c_int32, C.POINTER(C.c_int32)], C.c_int),
    "asx_autotune_substep": ([_P, _S, _S, C.c_double, C.c_double, C.c_double, _PP,
                              C.c_int32, _V], C.c_int),
    "asx_substep": ([_P, _S, _S, C.c_double, C.c_double, C.c_double, _PP, _V],
                    C.c_int),
    "asx_substep_bc": ([_P, _S, _S, C.c_double, C.c_double, C.c_double, _PP, C.c_int32,
                        _V], C.c_int),
    "asx_c2p_batch": ([C.c_int64, _V, _V, _V, _V, _V, _V, _PP, _V], C.c_int),
    "asx_zone_sample": ([C.c_int64, C.c_uint64, C.c_double, C.c_double, _V, _V, _V, _V,
                         _V], C.c_int),
    "asx_prim2con_batch": ([C.c_int64, _V, _V, _V, _V, _PP, _V], C.c_int),
    "asx_mlp_features": ([C.c_int64, _V, _V, _V, _V], C.c_int),
    "asx_mlp_forward": ([C.c_int64, _V, _V, _V, _V, _V], C.c_int),
    "asx_mlp_train_step": ([C.c_int64, _V, _V, _V, _V, _V, _V, C.c_float, C.c_int64, _V,
                            _V, _V, _V], C.c_int),
    "asx_c2p_batch_mlp": ([C.c_int64, _V, _V, _V, _V, C.c_int32, _V, _V, _V, _V, _PP, _V],
                          C.c_int),
    "asx_eos_from_rho_temp": ([C.c_int64, _V, _V, _V, _V, _PP, _V], C.c_int),
    "asx_eos_from_rho_eps": ([C.c_int64, _V, _V, _V, _V, _PP, _V], C.c_int),
    "asx_prim2con_batch_ye": ([C.c_int64, _V, _V, _V, _V, _PP, _V], C.c_int),
    "asx_c2p_batch_ye": ([C.c_int64, _V, _V, _V, _V, _V, _V, _PP, _V], C.c_int),
    "asx_zone_sample_ye": ([C.c_int64, C.c_uint64, _V, C.c_double, _V, _V, _V, _V, _PP, _V],
                           C.c_int),
    "asx_mlp_ye_features": ([C.c_int64, _V, _V, _V, _V], C.c_int),
    "asx_mlp_ye_forward": ([C.c_int64, _V, _V, _V, _V, _V], C.c_int),
    "asx_mlp_ye_train_step": ([C.c_int64, _V, _V, _V, _V, _V, _V, C.c_float, C.c_int64, _V,
                               _V, _V, _V], C.c_int),
    "asx_c2p_batch_ye_mlp": ([C.c_int64, _V, _V, _V, _V, C.c_double, _V, _V, _V, _PP, _V],
                             C.c_int),
    "asx_batch_create": ([C.c_int32, _P, _S, _PP, C.c_double, C.c_int32,
                          C.POINTER(C.c_void_p)], C.c_int),
    "asx_batch_stage": ([_V, C.c_int32, _V], C.c_int),
    "asx_batch_step": ([_V, C.c_int32, C.c_int32, _V], C.c_int),
    "asx_batch_autotune": ([_V, C.c_int32, _V], C.c_int),
    "asx_batch_get_variant": ([_V, C.c_int32, C.POINTER(C.c_int32)], C.c_int),
    "asx_batch_info": ([_V, C.POINTER(C.c_int64), C.POINTER(C.c_int64)], C.c_int),
    "asx_batch_destroy": ([_V], C.c_int),
    "asx_params_default": ([_PP], None),
    "asx_gf_centering": ([C.c_int], C.c_int),
    "asx_gf_strides": ([C.POINTER(C.c_int32), C.c_int32, C.c_int,
                        C.POINTER(C.c_int64), C.POINTER(C.c_int64)], None),
    "asx_version": ([], C.c_char_p),
}

_lib = None


def build():
    """Compile every HIP source for gfx950 into libasterx_hip.so (in-tree)."""
    subprocess.check_call(["make", "-s", "-B", "-j8", "-C",
                           os.path.join(_HERE, "csrc")])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise RuntimeError(
                "libasterx_hip.so is not built (%s); run __graft_entry__.build()"
                % LIBPATH)
        # torch ships its own HIP runtime under the same soname and loads it
        # by path; loaded first, it is the one this library binds to, so that
        # device pointers from torch tensors and from this library share one
        # runtime.  Loaded the other way round a process holds two runtimes
        # and the library's launches fail on torch's allocations.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(LIBPATH)
        for name, (args, res) in SIGNATURES.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


EOSIO_PATH = os.path.join(_HERE, "libasterx_eosio.so")
_eosio = None


class asx_eos_host_table(C.Structure):
    _fields_ = [("n", C.c_int32 * 3), ("data", C.POINTER(C.c_double)),
                ("axis", C.POINTER(C.c_double) * 3), ("energy_shift", C.c_double),
                ("eps_min", C.c_double), ("eps_max", C.c_double)]


def eosio():
    """The host table reader (libasterx_eosio.so, include/asterx_eosio.h)."""
    global _eosio
    if _eosio is None:
        if not os.path.exists(EOSIO_PATH):
            raise RuntimeError("libasterx_eosio.so is not built (%s); run "
                               "__graft_entry__.build()" % EOSIO_PATH)
        L = C.CDLL(EOSIO_PATH)
        L.asx_eos_read_stellarcollapse.argtypes = [C.c_char_p, C.POINTER(asx_eos_host_table)]
        L.asx_eos_read_stellarcollapse.restype = C.c_int32
        L.asx_eos_host_table_free.argtypes = [C.POINTER(asx_eos_host_table)]
        L.asx_eos_host_table_free.restype = None
        _eosio = L
    return _eosio


def backend(stream=None):
    from .driver import Backend
    return Backend(lib(), "asx_", stream)
