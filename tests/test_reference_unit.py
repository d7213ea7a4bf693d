"""Replays of the reference's own unit tests on the oracle.

AsterX/src/unit_tests/*.cxx (run by AsterX/src/test.cxx:14-28 when
unit_test = yes) check the AsterUtils algebra, minmod and MP5 with draws
from std::mt19937_64 seeded 100 (test.hxx:13) through libstdc++'s
uniform_real_distribution, comparing with isapprox (relative tolerance
sqrt(machine epsilon), test.hxx:28-37).  The reference only prints PASSED /
FAILED; here every check is an assertion.  The engine is restated below and
checked against the C++ standard's known value, so the replays draw the
reference's numbers: the reference's configuration (unit_test_repetitions
= 1, AsterX/param.ccl:65-68) and 200 repetitions.

Con2PrimFactory/src/test.cxx:21-247 is the con2prim known-answer case:
rho 0.125, eps 0.8, B = (0.5, -0.5, 0) at rest in flat space, through
Noble, Palenzuela and Entropy at tolerance 1e-8; each must return the
state it was built from.
"""
import ctypes as C

import numpy as np
import pytest

import oracle
from asterx_amd import abi
from states import kat_params, kat_state

M64 = (1 << 64) - 1


class MT19937_64:
    """std::mt19937_64 (C++ [rand.predef]: w 64, n 312, m 156, r 31)."""

    def __init__(self, seed):
        self.mt = [seed & M64]
        for i in range(1, 312):
            x = self.mt[-1]
            self.mt.append((6364136223846793005 * (x ^ (x >> 62)) + i) & M64)
        self.i = 312

    def __call__(self):
        mt = self.mt
        if self.i >= 312:
            for k in range(312):
                x = (mt[k] & 0xFFFFFFFF80000000) | (mt[(k + 1) % 312] & 0x7FFFFFFF)
                xa = x >> 1
                if x & 1:
                    xa ^= 0xB5026F5AA96619E9
                mt[k] = mt[(k + 156) % 312] ^ xa
            self.i = 0
        x = mt[self.i]
        self.i += 1
        x ^= (x >> 29) & 0x5555555555555555
        x ^= (x << 17) & 0x71D67FFFEDA60000
        x ^= (x << 37) & 0xFFF7EEE000000000
        x ^= x >> 43
        return x & M64

    def uniform(self, a, b):
        """libstdc++ uniform_real_distribution<double>: generate_canonical
        with one 64-bit draw, then a + u (b - a)."""
        u = float(self()) / 18446744073709551616.0
        if u >= 1.0:
            u = np.nextafter(1.0, 0.0)
        return u * (b - a) + a


def isapprox(x, y):
    return abs(x - y) <= np.sqrt(np.finfo(float).eps) * max(abs(x), abs(y))


def test_engine_known_value():
    """[rand.predef]: the 10000th draw of a default-constructed
    mt19937_64 (seed 5489) is 9981545732273789042."""
    g = MT19937_64(5489)
    for _ in range(9999):
        g()
    assert g() == 9981545732273789042


def _util(op, vals, nout):
    a = np.ascontiguousarray(vals, dtype=np.float64)
    out = np.zeros(nout)
    abi.check("orc_util", oracle.lib().orc_util(op, a.ctypes.data, out.ctypes.data))
    return out


def _mp5(s):
    a = np.ascontiguousarray(s, dtype=np.float64)
    out = np.zeros(2)
    par = abi.params()
    abi.check("orc_reconstruct", oracle.lib().orc_reconstruct(
        abi.RECON["mp5"], a.ctypes.data, None, None, 0, C.byref(par), out.ctypes.data))
    return out


def replay_all(g, reps):
    """AsterX/src/test.cxx:19-28 order, one engine.  Returns the number of
    MP5 cubic cases whose reconstruction leaves the range of the face
    values (the other checks assert)."""
    # contraction_smat_upvec.cxx
    for _ in range(reps):
        m = [g.uniform(-1, 1) for _ in range(6)]
        v = [g.uniform(-1, 1) for _ in range(3)]
        full = np.array([[m[0], m[1], m[2]], [m[1], m[3], m[4]], [m[2], m[4], m[5]]])
        got = _util(3, m + v, 3)
        for i in range(3):
            assert isapprox(got[i], full[i, 0] * v[0] + full[i, 1] * v[1] + full[i, 2] * v[2])
    # contraction_upvec_downvec.cxx
    for _ in range(reps):
        a = [g.uniform(-1, 1) for _ in range(6)]
        assert isapprox(_util(4, a, 1)[0], a[0] * a[3] + a[1] * a[4] + a[2] * a[5])
    # cross_product.cxx
    for _ in range(reps):
        a = [g.uniform(-1, 1) for _ in range(6)]
        got = _util(2, a, 3)
        want = [a[1] * a[5] - a[2] * a[4], a[2] * a[3] - a[0] * a[5], a[0] * a[4] - a[1] * a[3]]
        for i in range(3):
            assert isapprox(got[i], want[i])
    # wlorentz.cxx: both sides are NaN when the draws give v.v > 1
    for _ in range(reps):
        a = [g.uniform(0.0, 0.9) for _ in range(6)]
        c = a[0] * a[3] + a[1] * a[4] + a[2] * a[5]
        got = _util(1, a, 1)[0]
        if c >= 1.0:
            assert np.isnan(got) or np.isinf(got)
        else:
            assert isapprox(got, 1.0 / np.sqrt(1.0 - c))
    # minmod.cxx
    mm = lambda x, y: _util(0, [x, y], 1)[0]
    for _ in range(reps):
        p = [g.uniform(0.0, 1.0) for _ in range(4)]
        n = [g.uniform(-1.0, 0.0) for _ in range(4)]
        r = [g.uniform(-1.0, 1.0) for _ in range(2)]
        assert isapprox(mm(p[0], p[1]), min(p[0], p[1]))
        assert isapprox(mm(n[0], n[1]), max(n[0], n[1]))
        assert isapprox(mm(n[0], p[0]), 0.0) and isapprox(mm(n[1], p[1]), 0.0)
        mx = max(abs(r[0]), abs(r[1]))
        assert mm(r[0], r[1]) < mx or isapprox(mm(r[0], r[1]), mx)
        assert isapprox(mm(mm(p[0], p[1]), mm(p[2], p[3])), min(p))
        assert isapprox(mm(mm(n[0], n[1]), mm(n[2], n[3])), max(n))
    # mp5.cxx: 7 cells on [-1, 1], the stencil of cells 0, 1, 2, 4, 5, 6
    dx = 2.0 / 7
    xc = np.array([-1.0 + dx * (i + 0.5) for i in range(7)])
    xf = np.array([-1.0 + i * dx for i in range(8)])
    st = lambda f: [f[0], f[1], f[2], f[4], f[5], f[6]]
    cubic_out = 0
    for _ in range(reps):
        y = g.uniform(-1, 1)
        lo, hi = _mp5(st(np.full(7, y)))
        assert isapprox(lo, y) and isapprox(hi, y)
        for ncoef in (2, 3, 4):
            cf = [g.uniform(-1, 1) for _ in range(ncoef)]
            poly = lambda x: sum(c * x ** (ncoef - 1 - k) for k, c in enumerate(cf))
            faces = poly(xf)
            lo, hi = _mp5(st(poly(xc)))
            inside = all(faces.min() <= v <= faces.max() for v in (lo, hi))
            if ncoef < 4:
                assert inside, (ncoef, cf, lo, hi)
            elif not inside:
                cubic_out += 1
    return cubic_out


def test_unit_tests_replayed():
    """The reference's configuration (one repetition): every check passes."""
    assert replay_all(MT19937_64(100), 1) == 0


def test_unit_tests_replayed_200():
    """200 repetitions.  MP5 on a cubic sampled at cells 0, 1, 2, 4, 5, 6
    (the stencil skips cell 3, mp5.cxx:193-195) is not bounded by the face
    values: 4 of these 200 draws leave that range, and the reference's own
    test prints FAILED for them (it never aborts).  Lines, parabolas and
    constants always pass."""
    assert replay_all(MT19937_64(100), 200) == 4


@pytest.mark.parametrize("prime", ["Noble", "Palenzuela", "Entropy"])
def test_c2p_known_answer(prime):
    p = kat_state()
    par = kat_params(prime)
    abi.check("orc_con2prim", oracle.lib().orc_con2prim(C.byref(p.struct()), C.byref(par)))
    sl = p.interior("rho")
    assert np.all(p.numpy("con2prim_flag")[sl] == 1.0)
    rho, eps = 0.125, 0.8
    for name, want in (("rho", rho), ("eps", eps), ("press", rho * eps), ("Ye", 0.5),
                       ("Bvecx", 0.5), ("Bvecy", -0.5)):
        got = p.numpy(name)[sl]
        assert np.abs(got - want).max() <= 1e-7 * abs(want), (prime, name, got.ravel()[0])
    for name in ("velx", "vely", "velz", "Bvecz"):
        assert np.abs(p.numpy(name)[sl]).max() <= 1e-12, (prime, name)
