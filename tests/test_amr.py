"""Pins of the oracle's refinement operators and regrid box table
(oracle/amr_ref.py) by mathematics, on the CPU.

* Prolongation of order p reproduces tensor-product polynomials exactly:
  degree p along vertex axes (point values), degree p-1 along cell axes
  (cell averages in, fine cell averages out); one degree more does not
  reproduce, so the order is exactly the claimed one.
* Conservation: the two children of a cell average to the parent.
* Restriction after prolongation is the identity on the covered points.
* Weights: the closed forms (1/2, 1/2), (-1, 9, 9, -1)/16, (1/8, 1, -1/8)
  and sum rules.
* Planted mistakes (a weight table of the wrong order, swapped children, a
  dropped sign) fail these checks.
* Box table: every tag covered, boxes disjoint, inside the domain, aligned
  to the blocking factor, at most max_grid_size, each cluster at grid_eff or
  one block; known answers on simple tag patterns.
"""
import numpy as np
import pytest

from oracle import amr_ref
from asterx_amd import abi
from asterx_amd.patch import Patch

NAMES = {0: "rho", 7: "alp", 1: "dBx_stag", 2: "dBy_stag", 4: "dBz_stag",
         6: "Avec_x", 5: "Avec_y", 3: "Avec_z"}


def levels(cn=(12, 12, 12), cng=3, off=(8, 6, 10), fn=(8, 10, 6), fng=3, x0=(-0.3, 0.1, 0.2),
           dx=(0.1, 0.12, 0.08)):
    coarse = Patch(cn, cng, x0, dx)
    fdx = tuple(d / 2 for d in dx)
    fx0 = tuple(x0[d] + off[d] * fdx[d] for d in range(3))
    fine = Patch(fn, fng, fx0, fdx)
    return coarse, fine


def coords(p, name):
    """Per-axis coordinates of a centering's points and whether the axis is
    a vertex axis."""
    c = abi.centering(name)
    out = []
    for d in range(3):
        vc = (c >> d) & 1
        E = p.n[d] + vc + 2 * p.ng
        x = p.x0[d] + (np.arange(E) - p.ng + (0.0 if vc else 0.5)) * p.dx[d]
        out.append((x, vc, p.dx[d]))
    return out


def mono_values(p, name, deg):
    """Tensor-product monomial prod_d (x_d - s_d)^deg[d]: point values along
    vertex axes, exact cell averages along cell axes."""
    shift = (0.07, -0.11, 0.05)
    vals = []
    for d, (x, vc, h) in enumerate(coords(p, name)):
        a = deg[d]
        y = x - shift[d]
        if vc:
            v = y ** a
        else:
            v = ((y + h / 2) ** (a + 1) - (y - h / 2) ** (a + 1)) / ((a + 1) * h)
        vals.append(v)
    return vals[2][:, None, None] * vals[1][None, :, None] * vals[0][None, None, :]


def exactness_error(order, name, deg, weights=None):
    coarse, fine = levels()
    coarse[name] = mono_values(coarse, name, deg)
    fine.alloc(name, np.nan)
    amr_ref.prolongate(coarse, fine, [name], order, 0, weights)
    want = mono_values(fine, name, deg)
    return np.abs(fine[name] - want).max() / np.abs(want).max()


@pytest.mark.parametrize("order", [1, 3, 5])
@pytest.mark.parametrize("cent", sorted(NAMES))
def test_prolongation_polynomial_exactness(order, cent):
    name = NAMES[cent]
    deg = [order if (cent >> d) & 1 else order - 1 for d in range(3)]
    assert exactness_error(order, name, deg) <= 1e-12
    for d in range(3):
        more = list(deg)
        more[d] += 1
        assert exactness_error(order, name, more) > 1e-8, (d, more)


def test_weights_closed_forms_and_sums():
    assert amr_ref.poly_mid_weights(1) == [0.5, 0.5]
    assert amr_ref.poly_mid_weights(3) == [-1 / 16, 9 / 16, 9 / 16, -1 / 16]
    assert amr_ref.cons_child_weights(0, 0) == [1.0]
    assert amr_ref.cons_child_weights(2, 0) == [1 / 8, 1.0, -1 / 8]
    assert amr_ref.cons_child_weights(2, 1) == [-1 / 8, 1.0, 1 / 8]
    for p in (1, 3, 5):
        assert sum(amr_ref.poly_mid_weights(p)) == 1.0
    for q in (0, 2, 4):
        lo, up = amr_ref.cons_child_weights(q, 0), amr_ref.cons_child_weights(q, 1)
        assert lo == up[::-1]
        assert [a + b for a, b in zip(lo, up)] == [2.0 if m == q // 2 else 0.0
                                                  for m in range(q + 1)]


@pytest.mark.parametrize("order", [1, 3, 5])
def test_conservation_and_restriction_identity(order):
    rng = np.random.default_rng(order)
    coarse, fine = levels()
    for name in ("rho", "alp", "Avec_x", "dBy_stag"):
        coarse[name] = rng.standard_normal(coarse.shape(name))
        fine.alloc(name, np.nan)
    names = ["rho", "alp", "Avec_x", "dBy_stag"]
    amr_ref.prolongate(coarse, fine, names, order, 0)
    back = coarse.to(None)
    for name in names:
        back[name] = np.full(coarse.shape(name), np.nan)
    amr_ref.restrict(fine, back, names)
    off = amr_ref._geometry(coarse, fine)
    for name in names:
        c = abi.centering(name)
        lo, cnt = amr_ref.restrict_box(c, off, fine.n, coarse.n, coarse.ng)
        sl = tuple(slice(lo[d], lo[d] + cnt[d]) for d in (2, 1, 0))
        a, b = back[name][sl], coarse[name][sl]
        assert a.size > 0 and not np.isnan(a).any()
        if c == 7:
            assert np.array_equal(a, b)  # injection of coinciding points
        else:
            assert np.abs(a - b).max() <= 1e-14 * np.abs(b).max()
        outside = np.ones(back[name].shape, bool)
        outside[sl] = False
        assert np.isnan(back[name][outside]).all()


def test_ghost_region_leaves_own_points():
    coarse, fine = levels()
    coarse["Avec_y"] = np.random.default_rng(3).standard_normal(coarse.shape("Avec_y"))
    fine.alloc("Avec_y", 7.0)
    amr_ref.prolongate(coarse, fine, ["Avec_y"], 3, 1)
    g = amr_ref.ghost_mask(fine.shape("Avec_y"), fine.ng)
    assert (fine["Avec_y"][~g] == 7.0).all()
    assert not (fine["Avec_y"][g] == 7.0).any()


def planted(order, kind):
    w = amr_ref.default_weights(order)
    if kind == "wrong_order":
        w["poly"] = [0.0] * ((order + 1) // 2 - 1) + [0.5, 0.5] + [0.0] * ((order + 1) // 2 - 1)
    elif kind == "swapped_children":
        w["cons"] = [w["cons"][1], w["cons"][0]]
    elif kind == "dropped_sign":
        w["poly"] = [abs(x) / sum(abs(y) for y in w["poly"]) for x in w["poly"]]
    return w


@pytest.mark.parametrize("kind,name", [("wrong_order", "alp"), ("swapped_children", "rho"),
                                       ("dropped_sign", "Avec_z")])
def test_planted_weight_mistakes_fail(kind, name):
    order = 3
    cent = abi.centering(name)
    deg = [order if (cent >> d) & 1 else order - 1 for d in range(3)]
    assert exactness_error(order, name, deg) <= 1e-12
    assert exactness_error(order, name, deg, planted(order, kind)) > 1e-6


# ---------------------------------------------------------------------------
# Box table
# ---------------------------------------------------------------------------
def check_boxes(err, thr, buf, bf, mgs, eff, boxes):
    nz, ny, nx = err.shape
    n = (nx, ny, nz)
    blk = [max(1, b // 2) for b in bf]
    unit = [2 * b for b in blk]
    tags = amr_ref.block_tags(err, thr, buf, blk)
    cover = np.zeros((2 * nz, 2 * ny, 2 * nx), dtype=np.int32)
    for lo, hi in boxes:
        for d in range(3):
            assert 0 <= lo[d] <= hi[d] < 2 * n[d]
            assert lo[d] % unit[d] == 0 and (hi[d] + 1) % unit[d] == 0
            assert hi[d] - lo[d] + 1 <= mgs[d]
        cover[lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1] += 1
    assert cover.max(initial=0) <= 1  # disjoint
    # every tagged block's fine cells are covered
    fine_tags = tags.repeat(unit[2], 0).repeat(unit[1], 1).repeat(unit[0], 2)
    assert (cover[fine_tags] == 1).all()
    # each cluster meets grid_eff or is one block
    for lo, hi in amr_ref.cluster(tags, eff):
        sub = tags[lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1]
        assert sub.size == 1 or sub.sum() >= eff * sub.size
        for d, ax in ((0, (0, 1)), (1, (0, 2)), (2, (1, 2))):
            s = sub.sum(axis=ax)
            assert s[0] > 0 and s[-1] > 0  # shrunk to the tags


def blob_error(n=(32, 24, 16), seed=0, nblob=3):
    rng = np.random.default_rng(seed)
    z, y, x = np.meshgrid(np.arange(n[2]), np.arange(n[1]), np.arange(n[0]), indexing="ij")
    e = np.zeros((n[2], n[1], n[0]))
    for _ in range(nblob):
        c = [rng.uniform(0, n[d]) for d in range(3)]
        r = rng.uniform(2, 5)
        e += np.exp(-(((x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2) / r ** 2))
    return e


@pytest.mark.parametrize("seed", range(4))
@pytest.mark.parametrize("bf,mgs,buf,eff", [((1, 1, 1), (64, 64, 64), 0, 0.7),
                                            ((4, 4, 4), (16, 16, 16), 1, 0.7),
                                            ((8, 4, 2), (16, 8, 32), 2, 0.9),
                                            ((2, 2, 2), (8, 8, 8), 1, 0.5)])
def test_box_invariants(seed, bf, mgs, buf, eff):
    err = blob_error(seed=seed)
    boxes = amr_ref.regrid_boxes(err, 0.3, buf, bf, mgs, eff)
    assert boxes
    check_boxes(err, 0.3, buf, bf, mgs, eff, boxes)


def test_box_known_answers():
    e = np.zeros((8, 8, 8))
    assert amr_ref.regrid_boxes(e, 0.5, 1, (2, 2, 2), (16, 16, 16), 0.7) == []
    e[3, 4, 5] = 1.0  # one cell, no buffer, blocking factor 1: its two fine cells per axis
    assert amr_ref.regrid_boxes(e, 0.5, 0, (1, 1, 1), (16, 16, 16), 0.7) == \
        [((10, 8, 6), (11, 9, 7))]
    # two separated slabs: one box each (cut at the empty plane between)
    e = np.zeros((8, 8, 8))
    e[:, :, 0:2] = 1.0
    e[:, :, 6:8] = 1.0
    assert amr_ref.regrid_boxes(e, 0.5, 0, (1, 1, 1), (16, 16, 16), 0.7) == \
        [((0, 0, 0), (3, 15, 15)), ((12, 0, 0), (15, 15, 15))]
    # everything tagged, max_grid_size 8: cut into 2 x 2 x 2 pieces
    e = np.ones((8, 8, 8))
    b = amr_ref.regrid_boxes(e, 0.5, 0, (2, 2, 2), (8, 8, 8), 0.7)
    assert len(b) == 8 and b[0] == ((0, 0, 0), (7, 7, 7)) and b[-1] == ((8, 8, 8), (15, 15, 15))
    # an L shape below grid_eff is cut at the signature's inflection
    e = np.zeros((4, 8, 8))
    e[:, 0:2, :] = 1.0
    e[:, :, 0:2] = 1.0
    b = amr_ref.regrid_boxes(e, 0.5, 0, (1, 1, 1), (32, 32, 32), 0.9)
    check_boxes(e, 0.5, 0, (1, 1, 1), (32, 32, 32), 0.9, b)
    assert len(b) == 2


def check_clusters(tags, eff, clusters):
    cover = np.zeros(tags.shape, dtype=np.int32)
    for lo, hi in clusters:
        sub = tags[lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1]
        assert sub.size == 1 or sub.sum() >= eff * sub.size
        cover[lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1] += 1
    assert cover.max(initial=0) <= 1
    assert (cover[tags] == 1).all()


@pytest.mark.parametrize("plant", ["no_split", "dropped_box", "duplicated_box"])
def test_planted_clustering_mistakes_fail(plant):
    """Clusterings with a planted mistake (bounding box kept below grid_eff,
    a cluster dropped, a cluster listed twice) fail the invariants the
    oracle's clustering meets."""
    tags = amr_ref.block_tags(blob_error(seed=1), 0.3, 0, (1, 1, 1))
    good = amr_ref.cluster(tags, 0.7)
    assert len(good) > 1
    check_clusters(tags, 0.7, good)
    bad = {"no_split": amr_ref.cluster(tags, 1e-9),
           "dropped_box": good[:-1],
           "duplicated_box": good + good[:1]}[plant]
    with pytest.raises(AssertionError):
        check_clusters(tags, 0.7, bad)
