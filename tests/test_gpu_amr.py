"""asx_prolongate / asx_restrict / asx_regrid_boxes (HIP) against the
oracle's restatement (oracle/amr_ref.py) on seeded data.  The kernels sum
in the oracle's order without contraction, so the bar is bit-identical
(array_equal) for every centering, order and region; the box table must be
the same list.  Also the C-ABI refusals."""
import ctypes as C

import numpy as np
import pytest

from oracle import amr_ref
from asterx_amd import abi
from asterx_amd.patch import Patch

pytestmark = pytest.mark.gpu

ALL = ["rho", "alp", "dBx_stag", "dBy_stag", "dBz_stag", "Avec_x", "Avec_y", "Avec_z"]


def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch


def _slots(names):
    return (C.c_int32 * max(1, len(names)))(*[abi.SLOT[n] for n in names])


def levels(cn, cng, off, fn, fng, x0=(-0.3, 0.1, 0.2), dx=(0.1, 0.12, 0.08)):
    coarse = Patch(cn, cng, x0, dx)
    fdx = tuple(d / 2 for d in dx)
    fine = Patch(fn, fng, tuple(x0[d] + off[d] * fdx[d] for d in range(3)), fdx)
    return coarse, fine


GEOMS = [  # coarse n, ng; fine offset, n, ng
    ((12, 12, 12), 3, (8, 6, 10), (8, 10, 6), 3),
    ((10, 9, 8), 4, (0, -2, 2), (20, 18, 12), 2),   # fine reaches into coarse ghosts
    ((16, 6, 6), 3, (14, 2, 0), (6, 4, 12), 1),
]


@pytest.mark.parametrize("order", [1, 3, 5])
@pytest.mark.parametrize("region", [0, 1])
@pytest.mark.parametrize("geom", range(len(GEOMS)))
def test_prolongate_matches_oracle(order, region, geom):
    torch = _torch()
    from asterx_amd import lib
    cn, cng, off, fn, fng = GEOMS[geom]
    ch, fh = levels(cn, cng, off, fn, fng)
    rng = np.random.default_rng(100 * geom + 10 * order + region)
    for name in ALL:
        ch[name] = rng.standard_normal(ch.shape(name))
        fh.alloc(name, 5.0)
    cd, fd = ch.to("cuda"), fh.to("cuda")
    st = lib.lib().asx_prolongate(C.byref(cd.struct()), C.byref(fd.struct()), len(ALL),
                                  _slots(ALL), order, region, None)
    abi.check("asx_prolongate", st)
    torch.cuda.synchronize()
    amr_ref.prolongate(ch, fh, ALL, order, region)
    for name in ALL:
        assert np.array_equal(fd.numpy(name), fh[name]), name


@pytest.mark.parametrize("geom", range(len(GEOMS)))
def test_restrict_matches_oracle(geom):
    torch = _torch()
    from asterx_amd import lib
    cn, cng, off, fn, fng = GEOMS[geom]
    ch, fh = levels(cn, cng, off, fn, fng)
    rng = np.random.default_rng(7 + geom)
    for name in ALL:
        fh[name] = rng.standard_normal(fh.shape(name))
        ch.alloc(name, np.nan)
    cd, fd = ch.to("cuda"), fh.to("cuda")
    abi.check("asx_restrict", lib.lib().asx_restrict(
        C.byref(fd.struct()), C.byref(cd.struct()), len(ALL), _slots(ALL), None))
    torch.cuda.synchronize()
    amr_ref.restrict(fh, ch, ALL)
    for name in ALL:
        a, b = cd.numpy(name), ch[name]
        assert np.array_equal(np.isnan(a), np.isnan(b)), name
        assert np.array_equal(a[~np.isnan(a)], b[~np.isnan(b)]), name
        assert (~np.isnan(b)).sum() > 0


def test_amr_refusals():
    _torch()
    from asterx_amd import lib
    L = lib.lib()
    ch, fh = levels((12, 12, 12), 3, (8, 6, 10), (8, 10, 6), 3)
    ch["rho"] = np.zeros(ch.shape("rho"))
    fh["rho"] = np.zeros(fh.shape("rho"))
    cd, fd = ch.to("cuda"), fh.to("cuda")
    c, f = cd.struct(), fd.struct()
    s = _slots(["rho"])
    assert L.asx_prolongate(C.byref(c), C.byref(f), 1, s, 2, 0, None) == 4  # UNSUPPORTED
    assert L.asx_prolongate(C.byref(c), C.byref(f), 1, s, 3, 2, None) == 3  # PARAM
    assert L.asx_prolongate(C.byref(c), C.byref(f), 65, s, 3, 0, None) == 3
    assert L.asx_prolongate(C.byref(c), C.byref(f), 1, None, 3, 0, None) == 1
    # odd offset, wrong ratio
    _, fo = levels((12, 12, 12), 3, (7, 6, 10), (8, 10, 6), 3)
    fo["rho"] = np.zeros(fo.shape("rho"))
    assert L.asx_prolongate(C.byref(c), C.byref(fo.to("cuda").struct()), 1, s, 1, 0,
                            None) == 2  # LAYOUT
    fw = Patch((8, 10, 6), 3, fh.x0, (0.05, 0.06, 0.05), slots=["rho"])
    assert L.asx_restrict(C.byref(fw.to("cuda").struct()), C.byref(c), 1, s, None) == 2
    # stencil past the coarse patch's points
    _, ff = levels((12, 12, 12), 3, (-6, 6, 10), (8, 10, 6), 3)
    ff["rho"] = np.zeros(ff.shape("rho"))
    assert L.asx_prolongate(C.byref(c), C.byref(ff.to("cuda").struct()), 1, s, 5, 0,
                            None) == 5  # GHOSTS


def _boxes(d, thr, buf, bf, mgs, eff, cap=4096):
    from asterx_amd import lib
    out = (abi.asx_box * max(1, cap))()
    nb = C.c_int32()
    st = lib.lib().asx_regrid_boxes(C.byref(d.struct()), thr, buf, (C.c_int32 * 3)(*bf),
                                    (C.c_int32 * 3)(*mgs), eff, cap, out, C.byref(nb), None)
    return st, [((tuple(out[q].lo)), tuple(out[q].hi)) for q in range(min(nb.value, cap))], \
        nb.value


def blob_error(n, seed, nblob=4):
    rng = np.random.default_rng(seed)
    z, y, x = np.meshgrid(np.arange(n[2]), np.arange(n[1]), np.arange(n[0]), indexing="ij")
    e = np.zeros((n[2], n[1], n[0]))
    for _ in range(nblob):
        c = [rng.uniform(0, n[d]) for d in range(3)]
        r = rng.uniform(2, 6)
        e += np.exp(-(((x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2) / r ** 2))
    return e


@pytest.mark.parametrize("seed", range(3))
@pytest.mark.parametrize("bf,mgs,buf,eff", [((1, 1, 1), (64, 64, 64), 0, 0.7),
                                            ((4, 4, 4), (16, 16, 16), 1, 0.7),
                                            ((8, 4, 2), (16, 8, 32), 2, 0.9)])
def test_regrid_boxes_match_oracle(seed, bf, mgs, buf, eff):
    torch = _torch()
    n = (32, 24, 16)
    h = Patch(n, 3, (0, 0, 0), (1, 1, 1))
    h.alloc("regrid_error", np.nan)  # ghosts are not read
    err = blob_error(n, seed)
    h["regrid_error"][3:-3, 3:-3, 3:-3] = err
    d = h.to("cuda")
    st, got, nb = _boxes(d, 0.3, buf, bf, mgs, eff)
    abi.check("asx_regrid_boxes", st)
    torch.cuda.synchronize()
    want = amr_ref.regrid_boxes(err, 0.3, buf, bf, mgs, eff)
    assert nb == len(want) and got == want
    # too small a table: the count, ASX_ERR_SIZE
    if len(want) > 1:
        st, got, nb = _boxes(d, 0.3, buf, bf, mgs, eff, cap=1)
        assert st == 7 and nb == len(want) and got == want[:1]


def test_regrid_boxes_refusals():
    _torch()
    h = Patch((12, 12, 12), 2, (0, 0, 0), (1, 1, 1), slots=["regrid_error"])
    d = h.to("cuda")
    assert _boxes(d, 0.1, 0, (3, 1, 1), (12, 12, 12), 0.7)[0] == 3   # not a power of 2
    assert _boxes(d, 0.1, 0, (8, 1, 1), (12, 12, 12), 0.7)[0] == 3   # mgs % bf
    assert _boxes(d, 0.1, 0, (16, 1, 1), (16, 12, 12), 0.7)[0] == 3  # n % (bf/2)
    assert _boxes(d, 0.1, -1, (1, 1, 1), (12, 12, 12), 0.7)[0] == 3
    assert _boxes(d, 0.1, 0, (1, 1, 1), (12, 12, 12), 0.0)[0] == 3
    st, got, nb = _boxes(d, 0.1, 0, (1, 1, 1), (12, 12, 12), 0.7)
    assert st == 0 and nb == 0 and got == []  # nothing above threshold
