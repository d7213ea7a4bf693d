"""C ABI: header, ctypes mirror and the built library agree (no GPU calls)."""
import ctypes as C
import os
import re

import pytest

from asterx_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "asterx.h")


def header_text():
    return open(HEADER).read()


def test_slot_enum_matches_header():
    txt = header_text()
    body = txt[txt.index("ASX_ALP = 0"):txt.index("ASX_NUM_GF")]
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    names = [t.strip().split("=")[0].strip() for t in body.split(",") if t.strip()]
    assert len(names) == abi.NUM_GF
    for hname, pyname in zip(names, abi.SLOTS):
        assert hname == "ASX_" + pyname.upper().replace("ASTER_", ""), (hname, pyname)


def declared_functions():
    txt = re.sub(r"/\*.*?\*/", "", header_text(), flags=re.S)
    return sorted(set(re.findall(r"\b(asx_[a-z0-9_A-Z]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    from asterx_amd import lib
    L = lib.lib()
    missing = [f for f in declared_functions() if not hasattr(L, f)]
    assert not missing, missing
    assert set(declared_functions()) == set(lib.SIGNATURES), \
        set(declared_functions()) ^ set(lib.SIGNATURES)


def test_params_default_mirror():
    from asterx_amd import lib
    p = abi.asx_params()
    lib.lib().asx_params_default(C.byref(p))
    q = abi.params()
    for name, _ in abi.asx_params._fields_:
        assert getattr(p, name) == getattr(q, name), name


def test_centering_and_strides():
    from asterx_amd import lib
    L = lib.lib()
    n = (C.c_int32 * 3)(5, 6, 7)
    for i, name in enumerate(abi.SLOTS):
        assert L.asx_gf_centering(i) == abi.centering(name)
        sj, sk = C.c_int64(), C.c_int64()
        L.asx_gf_strides(n, 3, i, C.byref(sj), C.byref(sk))
        ex, ey, _ = abi.extents((5, 6, 7), 3, abi.centering(name))
        assert (sj.value, sk.value) == (ex, ex * ey)


def test_oracle_slot_table_agrees():
    import oracle
    for i, name in enumerate(abi.SLOTS):
        assert oracle.lib().asx_gf_centering(i) == abi.centering(name)


def test_errors_are_status_codes():
    """Invalid arguments return a status, never abort (host-side checks)."""
    from asterx_amd import lib
    L = lib.lib()
    par = abi.params()
    assert L.asx_fluxes(None, C.byref(par), 0, None, None) == 1  # ASX_ERR_NULL
    from asterx_amd.patch import Patch
    p = Patch((4, 4, 4), 3, (0, 0, 0), (1, 1, 1), slots=["rho"])
    ps = p.struct()
    bad = abi.params(reconstruction_method="eppm")
    assert L.asx_fluxes(C.byref(ps), C.byref(bad), 0, None, None) == 4
    bad = abi.params(evolution_eos="Hybrid")
    assert L.asx_con2prim(C.byref(ps), C.byref(bad), None, None) == 4
    ok = abi.params()
    assert L.asx_fluxes(C.byref(ps), C.byref(ok), 0, None, None) == 1  # slots missing
    ps.gf[abi.SLOT["rho"]].sj += 1
    assert L.asx_bc_fill(C.byref(ps), abi.SLOT["rho"], 0, None) == 2  # layout


def test_one_hip_runtime_when_library_loads_first():
    """Loading the library before torch must still leave one HIP runtime in
    the process (torch's), so torch's device pointers are valid for it."""
    import subprocess
    import sys
    code = ("import asterx_amd.lib as l; l.lib(); import torch; "
            "m = open('/proc/self/maps').read(); "
            "print(len({x.split()[-1] for x in m.splitlines() if 'libamdhip64' in x}))")
    out = subprocess.run([sys.executable, "-c", code], cwd=ROOT, capture_output=True,
                         text=True, timeout=600)
    assert out.returncode == 0, out.stderr[-2000:]
    assert out.stdout.strip().splitlines()[-1] == "1", out.stdout


def test_eosio_library_exports_its_header():
    """libasterx_eosio.so exports every function include/asterx_eosio.h
    declares."""
    from asterx_amd import lib
    txt = re.sub(r"/\*.*?\*/", "",
                 open(os.path.join(ROOT, "include", "asterx_eosio.h")).read(), flags=re.S)
    names = sorted(set(re.findall(r"\b(asx_[a-z0-9_A-Z]+)\s*\(", txt)))
    assert names == ["asx_eos_host_table_free", "asx_eos_read_stellarcollapse"]
    L = lib.eosio()
    assert all(hasattr(L, f) for f in names)
