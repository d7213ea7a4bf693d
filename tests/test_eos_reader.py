"""The StellarCollapse table reader (asx_eos_read_stellarcollapse, EOSX's
eos_3p_tabulated3d.hxx:53-193) on a file in that layout written here with
the HDF5 C library: the synthetic table's CGS arrays go in, and what comes
out must equal the same arrays converted by tabeos.code_units (EOSX's unit
conversion and variable-fastest order), node for node; missing datasets,
short axes and non-uniform axes are refused."""
import ctypes as C
import math
import os

import numpy as np
import pytest

from asterx_amd import lib, tabeos

H5 = "/opt/conda/lib/libhdf5.so"
hid_t = C.c_int64


def _h5():
    if not os.path.exists(H5):
        pytest.skip("no HDF5 library")
    h = C.CDLL(H5)
    h.H5open()
    for f, res in (("H5Fcreate", hid_t), ("H5Screate_simple", hid_t), ("H5Dcreate2", hid_t)):
        getattr(h, f).restype = res
    h.H5Fcreate.argtypes = [C.c_char_p, C.c_uint, hid_t, hid_t]
    h.H5Screate_simple.argtypes = [C.c_int, C.POINTER(C.c_ulonglong), C.c_void_p]
    h.H5Dcreate2.argtypes = [hid_t, C.c_char_p, hid_t, hid_t, hid_t, hid_t, hid_t]
    h.H5Dwrite.argtypes = [hid_t, hid_t, hid_t, hid_t, hid_t, C.c_void_p]
    for f in ("H5Dclose", "H5Sclose", "H5Fclose"):
        getattr(h, f).argtypes = [hid_t]
    return h


def write_stellarcollapse(path, datasets):
    """datasets: name -> numpy array (float64 or int32), written with its shape."""
    h = _h5()
    dbl = hid_t.in_dll(h, "H5T_NATIVE_DOUBLE_g").value
    nint = hid_t.in_dll(h, "H5T_NATIVE_INT_g").value
    f = h.H5Fcreate(os.fsencode(path), 2, 0, 0)  # H5F_ACC_TRUNC, default lists
    assert f >= 0
    for name, a in datasets.items():
        a = np.ascontiguousarray(a)
        dims = (C.c_ulonglong * a.ndim)(*a.shape)
        sp = h.H5Screate_simple(a.ndim, dims, None)
        t = nint if a.dtype == np.int32 else dbl
        d = h.H5Dcreate2(f, name.encode(), t, sp, 0, 0, 0)
        assert d >= 0
        assert h.H5Dwrite(d, t, 0, 0, 0, a.ctypes.data) >= 0
        h.H5Dclose(d)
        h.H5Sclose(sp)
    h.H5Fclose(f)


def _datasets(nrho=23, ntemp=17, nye=9):
    lr, lt, ye, lp, le, cs2, ent = tabeos.make_cgs(nrho, ntemp, nye)
    return dict(pointsrho=np.array([nrho], np.int32), pointstemp=np.array([ntemp], np.int32),
                pointsye=np.array([nye], np.int32), logrho=lr, logtemp=lt, ye=ye,
                energy_shift=np.array([tabeos.SHIFT_CGS]), logpress=lp, logenergy=le,
                cs2=cs2, entropy=ent,
                munu=np.zeros_like(lp)), (lr, lt, ye, lp, le, cs2, ent)


def test_reader_matches_conversion(tmp_path):
    ds, cgs = _datasets()
    path = str(tmp_path / "table.h5")
    write_stellarcollapse(path, ds)
    got = tabeos.read_stellarcollapse(path)
    want = tabeos.code_units(*cgs)
    assert np.array_equal(got["data"], want["data"])
    for k in range(3):
        assert np.array_equal(got["axes"][k], want["axes"][k])
    assert got["shift"] == want["shift"]
    assert math.isclose(got["eps_min"], want["eps_min"], rel_tol=1e-15)
    assert math.isclose(got["eps_max"], want["eps_max"], rel_tol=1e-15)
    # spot check of the conversion itself at one node: P in code units
    i, j, k = 5, 7, 3
    p_cgs = 10.0 ** cgs[3][k, j, i]
    assert math.isclose(math.exp(got["data"][k, j, i, 0]), p_cgs * tabeos.PRESSGF,
                        rel_tol=1e-12)
    # a table built from the read arrays carries them into asx_eos_table
    t = tabeos.Table(got, device=None)
    assert t.desc.n[0] == 23 and t.desc.nvars == 4


def test_reader_refuses_bad_files(tmp_path):
    L = lib.eosio()
    t = lib.asx_eos_host_table()
    assert L.asx_eos_read_stellarcollapse(str(tmp_path / "none.h5").encode(), C.byref(t)) == 2
    ds, _ = _datasets()
    del ds["cs2"]
    path = str(tmp_path / "missing.h5")
    write_stellarcollapse(path, ds)
    assert L.asx_eos_read_stellarcollapse(path.encode(), C.byref(t)) == 2
    ds, _ = _datasets()
    ds["logrho"] = ds["logrho"].copy()
    ds["logrho"][4] += 1e-3
    path = str(tmp_path / "nonuniform.h5")
    write_stellarcollapse(path, ds)
    assert L.asx_eos_read_stellarcollapse(path.encode(), C.byref(t)) == 3
    assert L.asx_eos_read_stellarcollapse(None, C.byref(t)) == 1
