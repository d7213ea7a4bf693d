"""CPU restatement of the refinement operators and the regrid box table:
test infrastructure only (tests/ and bench.py's CPU legs), never the
product path.

What it restates.  CarpetX, which owns these operators, is not part of the
reference tree; the reference selects them by parameter
(CarpetX::prolongation_type = "ddf", AsterX/test/Balsara1_shocktube_xdir_
ppm.par:59; CarpetX::regrid_error_threshold / blocking_factor_* /
max_grid_size_*, :52-57 and AsterX/par/Scaling_magTOV_Z4c_unigrid.par:61-63).
So the restatement follows the operators' definitions:

* prolongation "ddf" (discrete differential forms) of order p, refinement 2:
  along a vertex axis, fine points on coarse points copy them and fine points
  between take the centred Lagrange interpolant of degree p; along a cell
  axis, the two children of a coarse cell take the averages over their
  halves of the degree p-1 polynomial whose cell averages match the p coarse
  cells centred on the parent.  3-D by tensor product.
* restriction: vertex axis injection, cell axis mean of the two children.
* box table: tags (error > threshold), grown by n_error_buf cells, coarsened
  to blocks of blocking_factor/2 cells, clustered by Berger & Rigoutsos,
  "An algorithm for point clustering and grid generation", IEEE Trans.
  Systems, Man and Cybernetics 21 (1991) 1278: accept a box at grid_eff,
  else cut at a hole, else at the strongest inflection of the signature,
  else in half; boxes cut to max_grid_size.

How it is pinned (tests/test_amr.py).  The weights are derived here from
the definitions with exact rational arithmetic, not typed in; the
operators are checked for polynomial exactness of the claimed degree and
failure one degree higher, for conservation (children average to the
parent) and for restriction after prolongation being the identity; the
clustering for its invariants (every tag covered, boxes disjoint and inside
the domain, aligned, at most max_grid_size, each accepted cluster at
grid_eff or a single block) and known answers.  Planted mistakes in the
weights and the clustering fail those checks.  Parity against CarpetX's own
code is unpinned (not in the reference tree).

Summation order equals the HIP kernels' (x innermost, then y, then z, each
partial sum from 0 adding weight * value in stencil order), so results are
bit-identical on the same inputs.
"""
from fractions import Fraction

import numpy as np


# ---------------------------------------------------------------------------
# Weights from the definitions
# ---------------------------------------------------------------------------
def _solve(A, b):
    n = len(A)
    M = [list(r) + [v] for r, v in zip(A, b)]
    for c in range(n):
        p = next(r for r in range(c, n) if M[r][c] != 0)
        M[c], M[p] = M[p], M[c]
        for r in range(n):
            if r != c and M[r][c] != 0:
                f = M[r][c] / M[c][c]
                M[r] = [x - f * y for x, y in zip(M[r], M[c])]
    return [M[i][n] / M[i][i] for i in range(n)]


def poly_mid_weights(order):
    """Weights at the order+1 coarse points -order/2 .. order/2 (unit
    spacing) of the Lagrange interpolant's value at the midpoint 0."""
    pts = [Fraction(2 * m - order, 2) for m in range(order + 1)]
    A = [[x ** k for x in pts] for k in range(order + 1)]
    b = [Fraction(1) if k == 0 else Fraction(0) for k in range(order + 1)]
    return [float(w) for w in _solve(A, b)]


def cons_child_weights(q, child):
    """Weights at the q+1 coarse cells -q/2 .. q/2 of the average over the
    lower (child 0: [-1/2, 0]) or upper (1: [0, 1/2]) half of the parent of
    the degree-q polynomial with those cell averages."""
    def integ(k, a, b):
        return (b ** (k + 1) - a ** (k + 1)) / (k + 1)
    cells = range(-(q // 2), q // 2 + 1)
    half = Fraction(1, 2)
    A = [[integ(k, Fraction(j) - half, Fraction(j) + half) for j in cells] for k in range(q + 1)]
    lo, hi = (-half, Fraction(0)) if child == 0 else (Fraction(0), half)
    b = [2 * integ(k, lo, hi) for k in range(q + 1)]
    return [float(w) for w in _solve(A, b)]


# ---------------------------------------------------------------------------
# Prolongation and restriction
# ---------------------------------------------------------------------------
def _extent(n, ng, vc):
    return n + vc + 2 * ng


def _geometry(coarse, fine):
    """Fine index of the fine patch's first interior point counted from
    twice the coarse one's, per axis (refinement 2, aligned patches)."""
    off = []
    for d in range(3):
        assert abs(2 * fine.dx[d] - coarse.dx[d]) <= 1e-12 * coarse.dx[d]
        r = (fine.x0[d] - coarse.x0[d]) / fine.dx[d]
        ri = int(round(r))
        assert abs(r - ri) <= 1e-6 and ri % 2 == 0 and fine.n[d] % 2 == 0
        off.append(ri)
    return off


def _axis_classes(I, vertex, order, weights):
    """Split fine indices I (array, counted from twice the coarse origin)
    into classes of equal stencil shape: (positions in I, first coarse
    index counted from the coarse origin, weights)."""
    out = []
    if vertex:
        ev = np.nonzero(I % 2 == 0)[0]
        od = np.nonzero(I % 2 == 1)[0]
        if ev.size:
            out.append((ev, I[ev] // 2, [1.0]))
        if od.size:
            out.append((od, (I[od] - 1) // 2 - (order - 1) // 2, weights["poly"]))
    else:
        for child in (0, 1):
            sel = np.nonzero(I % 2 == child)[0]
            if sel.size:
                w = [1.0] if order == 1 else weights["cons"][child]
                out.append((sel, I[sel] // 2 - (order - 1) // 2, w))
    return out


def default_weights(order):
    return {"poly": poly_mid_weights(order),
            "cons": [cons_child_weights(order - 1, 0), cons_child_weights(order - 1, 1)]}


def prolongate_array(U, cent, cng, off, fn, fng, order, weights=None):
    """Every point of one fine grid function (numpy (z, y, x) array) from
    the coarse array U of the same centering code `cent`."""
    w = weights or default_weights(order)
    E = [_extent(fn[d], fng, (cent >> d) & 1) for d in range(3)]
    F = np.empty((E[2], E[1], E[0]))
    cls = [_axis_classes(off[d] + np.arange(E[d]) - fng, (cent >> d) & 1, order, w)
           for d in range(3)]
    for px, sx, wx in cls[0]:
        for py, sy, wy in cls[1]:
            for pz, sz, wz in cls[2]:
                az = np.zeros((pz.size, py.size, px.size))
                for c in range(len(wz)):
                    ay = np.zeros_like(az)
                    for b in range(len(wy)):
                        ax = np.zeros_like(az)
                        for a in range(len(wx)):
                            blk = U[np.ix_(sz + c + cng, sy + b + cng, sx + a + cng)]
                            ax = ax + wx[a] * blk
                        ay = ay + wy[b] * ax
                    az = az + wz[c] * ay
                F[np.ix_(pz, py, px)] = az
    return F


def ghost_mask(shape, ng):
    m = np.ones(shape, dtype=bool)
    m[ng:shape[0] - ng, ng:shape[1] - ng, ng:shape[2] - ng] = False
    return m


def prolongate(coarse, fine, names, order, region, weights=None):
    """asx_prolongate on host Patch objects (in place on `fine`)."""
    from asterx_amd import abi
    off = _geometry(coarse, fine)
    for name in names:
        c = abi.centering(name)
        F = prolongate_array(coarse[name], c, coarse.ng, off, fine.n, fine.ng, order, weights)
        if region == 0:
            fine[name][...] = F
        else:
            m = ghost_mask(F.shape, fine.ng)
            fine[name][m] = F[m]


def restrict_box(c, off, fn, cn, cng):
    """Coarse points (local indices, per axis lo and count) that asx_restrict
    writes for centering code c."""
    lo, cnt = [], []
    for d in range(3):
        vc = (c >> d) & 1
        a = max(off[d] // 2, 0)
        b = min((off[d] + fn[d]) // 2 if vc else off[d] // 2 + fn[d] // 2 - 1, cn[d] + vc - 1)
        lo.append(a + cng)
        cnt.append(max(0, b - a + 1))
    return lo, cnt


def restrict(fine, coarse, names):
    """asx_restrict on host Patch objects (in place on `coarse`)."""
    from asterx_amd import abi
    off = _geometry(coarse, fine)
    for name in names:
        c = abi.centering(name)
        lo, cnt = restrict_box(c, off, fine.n, coarse.n, coarse.ng)
        if min(cnt) == 0:
            continue
        U = fine[name]
        idx, wts = [], []
        for d in range(3):
            ic = lo[d] + np.arange(cnt[d])
            st = 2 * (ic - coarse.ng) - off[d] + fine.ng
            if (c >> d) & 1:
                idx.append([st])
                wts.append([1.0])
            else:
                idx.append([st, st + 1])
                wts.append([0.5, 0.5])
        az = np.zeros((cnt[2], cnt[1], cnt[0]))
        for c2 in range(len(wts[2])):
            ay = np.zeros_like(az)
            for b in range(len(wts[1])):
                ax = np.zeros_like(az)
                for a in range(len(wts[0])):
                    ax = ax + wts[0][a] * U[np.ix_(idx[2][c2], idx[1][b], idx[0][a])]
                ay = ay + wts[1][b] * ax
            az = az + wts[2][c2] * ay
        coarse[name][lo[2]:lo[2] + cnt[2], lo[1]:lo[1] + cnt[1], lo[0]:lo[0] + cnt[0]] = az


# ---------------------------------------------------------------------------
# Regrid box table
# ---------------------------------------------------------------------------
def block_tags(err, threshold, n_error_buf, blk):
    """Tags of blocks of blk[d] cells from the interior regrid_error (numpy
    (z, y, x)): a block is tagged when a cell within n_error_buf cells of it
    has error > threshold."""
    t = err > threshold
    b = n_error_buf
    if b > 0:
        g = np.zeros(t.shape, dtype=bool)
        nz, ny, nx = t.shape
        p = np.pad(t, b)
        for dz in range(2 * b + 1):
            for dy in range(2 * b + 1):
                for dx in range(2 * b + 1):
                    g |= p[dz:dz + nz, dy:dy + ny, dx:dx + nx]
        t = g
    nz, ny, nx = t.shape
    bz, by, bx = blk[2], blk[1], blk[0]
    return t.reshape(nz // bz, bz, ny // by, by, nx // bx, bx).any(axis=(1, 3, 5))


def cluster(tags, grid_eff):
    """Berger-Rigoutsos clusters of a (z, y, x) bool block array, as
    inclusive (lo[x,y,z], hi[x,y,z]) boxes in depth-first order (lower part
    before upper part)."""
    T = np.asarray(tags, dtype=bool)
    out = []

    def run(lo, hi):
        sub = T[lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1]
        total = int(sub.sum())
        if total == 0:
            return
        sig = [sub.sum(axis=(0, 1)), sub.sum(axis=(0, 2)), sub.sum(axis=(1, 2))]
        lo, hi = list(lo), list(hi)
        for d in range(3):
            nz = np.nonzero(sig[d])[0]
            a, e = int(nz[0]), int(nz[-1])
            sig[d] = sig[d][a:e + 1].astype(np.int64)
            hi[d] = lo[d] + e
            lo[d] += a
        vol = sig[0].size * sig[1].size * sig[2].size
        if vol == 1 or total >= grid_eff * vol:
            out.append((tuple(lo), tuple(hi)))
            return
        cd, cut = -1, 0
        best = 2.0
        for d in range(3):
            L = sig[d].size
            for x in range(1, L - 1):
                if sig[d][x] == 0:
                    dist = abs(2.0 * x - (L - 1)) / L
                    if dist < best:
                        best, cd, cut = dist, d, x
        if cd < 0:
            strength, bdist = 0, 2.0
            for d in range(3):
                s = sig[d]
                L = s.size
                for x in range(1, L - 2):
                    l0 = s[x + 1] - 2 * s[x] + s[x - 1]
                    l1 = s[x + 2] - 2 * s[x + 1] + s[x]
                    if (l0 < 0 < l1) or (l0 > 0 > l1):
                        st = abs(int(l1) - int(l0))
                        dist = abs(2.0 * (x + 1) - L) / L
                        if st > strength or (st == strength and dist < bdist):
                            strength, bdist, cd, cut = st, dist, d, x + 1
        if cd < 0:
            cd = int(np.argmax([sig[0].size, sig[1].size, sig[2].size]))
            cut = sig[cd].size // 2
        lo1, hi1 = list(lo), list(hi)
        hi1[cd] = lo[cd] + cut - 1
        lo2, hi2 = list(lo), list(hi)
        lo2[cd] = lo[cd] + cut
        run(lo1, hi1)
        run(lo2, hi2)

    nz, ny, nx = T.shape
    run([0, 0, 0], [nx - 1, ny - 1, nz - 1])
    return out


def regrid_boxes(err, threshold, n_error_buf, blocking_factor, max_grid_size, grid_eff):
    """asx_regrid_boxes on the interior regrid_error (numpy (z, y, x)):
    list of (lo, hi) in fine cells, inclusive, x/y/z order."""
    blk = [max(1, b // 2) for b in blocking_factor]
    unit = [2 * b for b in blk]
    per = [m // u for m, u in zip(max_grid_size, unit)]
    tags = block_tags(err, threshold, n_error_buf, blk)
    boxes = []
    for lo, hi in cluster(tags, grid_eff):
        ln = [hi[d] - lo[d] + 1 for d in range(3)]
        npc = [(ln[d] + per[d] - 1) // per[d] for d in range(3)]
        for pz in range(npc[2]):
            for py in range(npc[1]):
                for px in range(npc[0]):
                    p = (px, py, pz)
                    blo, bhi = [], []
                    for d in range(3):
                        base, extra = divmod(ln[d], npc[d])
                        a = lo[d] + p[d] * base + min(p[d], extra)
                        w = base + (1 if p[d] < extra else 0)
                        blo.append(a * unit[d])
                        bhi.append((a + w) * unit[d] - 1)
                    boxes.append((tuple(blo), tuple(bhi)))
    return boxes
