"""The oracle's tabulated EOS and con2prim on it, pinned by what the
mathematics fixes (no reference table or build is available):

* the synthetic table: its sound speed against finite differences of
  P(rho, eps), monotonicity in T, and the unit conversion of EOSX's reader;
* interpolation: table values at the nodes, exactness for multilinear
  functions, agreement with scipy's RegularGridInterpolator (multilinear
  interpolation on a grid is unique);
* the temperature inversion (EOSX zero_brent): T(rho, eps(rho, T, Ye)) = T,
  the clamping at the table edges, and the root against scipy's brentq;
* Palenzuela con2prim on the table: prim -> cons -> prim round trip.
"""
import ctypes as C
import math

import numpy as np
import pytest
from scipy.interpolate import RegularGridInterpolator
from scipy.optimize import brentq

import oracle
from asterx_amd import abi, tabeos, zones


@pytest.fixture(scope="module")
def table():
    tab = tabeos.build()
    return tabeos.Table(tab, device=None)


def _par(table, **kw):
    base = dict(c2p_prime="Palenzuela", c2p_second="None", use_temperature=1,
                use_press_atmo=0, thermal_eos_atmo=1, rho_abs_min=1e-20, t_atmo=0.01,
                Ye_atmo=0.3, c2p_tol=1e-10, max_iter=100, vw_lim=1e3, B_lim=1e3,
                ye_min=0.035, ye_max=0.55)
    base.update(kw)
    return table.attach(abi.params(**base))


def _d(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def test_table_sound_speed_and_monotonicity():
    rho = np.array([1e5, 1e11, 3e14, 8e14, 2e15])
    temp = np.array([0.05, 1.0, 10.0, 30.0, 100.0])
    ye = np.array([0.1, 0.25, 0.4, 0.5, 0.3])
    P, eps, cs2 = tabeos.state_cgs(rho, temp, ye)
    # dP/drho at fixed eps and dP/deps at fixed rho by central differences
    # of P(rho, eps): P = S P_c(rho) + (2/3) rho (eps - S eps_c(rho))
    S = tabeos.ye_factor(ye)

    def P_re(r, e):
        Pc, ec, _ = tabeos.cold(r)
        return S * Pc + (2.0 / 3.0) * r * (e - S * ec)
    hr, he = 1e-6 * rho, 1e-6 * eps
    dPdr = (P_re(rho + hr, eps) - P_re(rho - hr, eps)) / (2 * hr)
    dPde = (P_re(rho, eps + he) - P_re(rho, eps - he)) / (2 * he)
    h = 1 + (eps + P / rho) / tabeos.C2_CGS
    assert np.allclose(cs2, (dPdr + P / rho ** 2 * dPde) / h, rtol=1e-6)
    assert np.allclose(P_re(rho, eps), P, rtol=1e-12)
    assert (cs2 > 0).all() and (cs2 < tabeos.C2_CGS).all()
    lr, lt, ye_ax, lp, le, c2, ent = tabeos.make_cgs(20, 30, 5)
    assert (ent > 0).all()
    assert (np.diff(le, axis=1) > 0).all()  # eps increases with T: invertible


def test_code_units_follow_the_reader():
    lr, lt, ye, lp, le, c2, ent = tabeos.make_cgs(6, 5, 4)
    t = tabeos.code_units(lr, lt, ye, lp, le, c2, ent)
    assert np.allclose(np.exp(t["axes"][0]), 10.0 ** lr * tabeos.RHOGF, rtol=1e-13)
    assert np.allclose(np.exp(t["axes"][1]), 10.0 ** lt, rtol=1e-13)
    k, j, i = 2, 3, 4
    row = t["data"][k, j, i]
    assert math.isclose(math.exp(row[0]), 10.0 ** lp[k, j, i] * tabeos.PRESSGF, rel_tol=1e-12)
    assert math.isclose(math.exp(row[1]), (10.0 ** le[k, j, i]) * tabeos.EPSGF, rel_tol=1e-12)
    # c_s^2 in units of c^2
    assert math.isclose(row[2], c2[k, j, i] / tabeos.C2_CGS, rel_tol=1e-9)
    assert row[3] == ent[k, j, i]  # entropy per baryon, no units
    assert t["eps_min"] >= 1e-15 and t["eps_max"] > t["eps_min"]


def _interp(table, lr, lt, ye):
    lr, lt, ye = _d(lr), _d(lt), _d(ye)  # keep the arrays alive over the call
    n = len(lr)
    out = np.zeros((4, n))
    abi.check("orc_tab_interp", oracle.lib().orc_tab_interp(
        n, lr.ctypes.data, lt.ctypes.data, ye.ctypes.data, out.ctypes.data,
        C.byref(_par(table))))
    return out


def test_interpolation_nodes_and_scipy(table):
    ax = table.host["axes"]
    data = table.host["data"]
    rng = np.random.default_rng(0)
    # interior nodes: exactly the table values
    i = rng.integers(0, len(ax[0]) - 1, 50)
    j = rng.integers(0, len(ax[1]) - 1, 50)
    k = rng.integers(0, len(ax[2]) - 1, 50)
    out = _interp(table, ax[0][i], ax[1][j], ax[2][k])
    assert np.array_equal(out.T, data[k, j, i])
    # random interior points against scipy's multilinear interpolator
    n = 2000
    pts = [rng.uniform(a[0], a[-1], n) for a in ax]
    out = _interp(table, *pts)
    for v in range(4):
        f = RegularGridInterpolator(ax, np.transpose(data[..., v], (2, 1, 0)))
        ref = f(np.stack(pts, axis=1))
        assert np.allclose(out[v], ref, rtol=1e-12, atol=1e-14 * np.abs(ref).max()), v


def test_interpolation_exact_for_multilinear(table):
    ax = table.host["axes"]
    Y, T, R = np.meshgrid(ax[2], ax[1], ax[0], indexing="ij")
    f = lambda r, t, y: 0.3 + 1.7 * r - 0.4 * t + 2.5 * y + 0.01 * r * t * y - 0.2 * t * y
    data = np.stack([f(R, T, Y), 2 * f(R, T, Y), -f(R, T, Y), 3 * f(R, T, Y)], axis=-1)
    t2 = dict(table.host, data=np.ascontiguousarray(data))
    tab2 = tabeos.Table(t2, device=None)
    rng = np.random.default_rng(1)
    pts = [rng.uniform(a[0], a[-1], 500) for a in ax]
    out = _interp(tab2, *pts)
    ref = f(*pts)
    assert np.allclose(out[0], ref, rtol=1e-12, atol=1e-12)
    assert np.allclose(out[1], 2 * ref, rtol=1e-12, atol=1e-12)
    assert np.allclose(out[3], 3 * ref, rtol=1e-12, atol=1e-12)


def _eos_rt(table, rho, temp, ye):
    rho, temp, ye = _d(rho), _d(temp), _d(ye)
    n = len(rho)
    out = np.zeros((4, n))
    abi.check("orc_eos_from_rho_temp", oracle.lib().orc_eos_from_rho_temp(
        n, rho.ctypes.data, temp.ctypes.data, ye.ctypes.data, out.ctypes.data,
        C.byref(_par(table))))
    return out


def _eos_re(table, rho, eps, ye):
    rho, eps, ye = _d(rho), _d(eps), _d(ye)
    n = len(rho)
    out = np.zeros((3, n))
    abi.check("orc_eos_from_rho_eps", oracle.lib().orc_eos_from_rho_eps(
        n, rho.ctypes.data, eps.ctypes.data, ye.ctypes.data, out.ctypes.data,
        C.byref(_par(table))))
    return out


def test_temperature_inversion_round_trip(table):
    (r0, r1), (t0, t1), (y0, y1) = table.ranges()
    rng = np.random.default_rng(2)
    n = 3000
    rho = np.exp(rng.uniform(np.log(r0) + 0.1, np.log(r1) - 0.1, n))
    temp = np.exp(rng.uniform(np.log(t0) + 0.05, np.log(t1) - 0.05, n))
    ye = rng.uniform(y0, y1, n)
    p, eps, cs, ent = _eos_rt(table, rho, temp, ye)
    assert np.allclose(ent, tabeos.entropy(rho / tabeos.RHOGF, temp), rtol=1e-6)
    T2, p2, eps2 = _eos_re(table, rho, eps, ye)
    assert np.max(np.abs(np.log(T2 / temp))) < 1e-12
    assert np.allclose(p2, p, rtol=1e-11)
    assert np.allclose(eps2, eps, rtol=1e-13)
    # below / above the table at fixed rho: T clamps to the axis end and eps
    # to the table's value there
    lo = _eos_rt(table, rho[:20], np.full(20, t0), ye[:20])[1]
    hi = _eos_rt(table, rho[:20], np.full(20, t1), ye[:20])[1]
    Tl, _, el = _eos_re(table, rho[:20], 0.5 * lo, ye[:20])
    Th, _, eh = _eos_re(table, rho[:20], 2.0 * hi, ye[:20])
    assert np.allclose(Tl, t0, rtol=1e-14) and np.allclose(el, lo, rtol=1e-12)
    assert np.allclose(Th, t1, rtol=1e-14) and np.allclose(eh, hi, rtol=1e-12)


def test_temperature_root_matches_brentq(table):
    """The inversion's root against scipy's brentq on the same residual
    (log eps(rho, T, Ye) of the interpolated table)."""
    ax = table.host["axes"]
    rng = np.random.default_rng(3)
    for _ in range(20):
        lrho = rng.uniform(ax[0][1], ax[0][-2])
        ye = rng.uniform(ax[2][0], ax[2][-1])
        lt_true = rng.uniform(ax[1][2], ax[1][-3])
        le = _interp(table, [lrho], [lt_true], [ye])[1][0]
        eps = math.exp(le) - table.host["shift"]
        T = _eos_re(table, np.array([math.exp(lrho)]), np.array([eps]), np.array([ye]))[0][0]

        def f(lt):
            return le - _interp(table, [lrho], [lt], [ye])[1][0]
        root = brentq(f, ax[1][0], ax[1][-1], xtol=1e-14)
        assert abs(math.log(T) - root) < 1e-12


def test_palenzuela_round_trip_on_table(table):
    (r0, r1), (t0, t1), (y0, y1) = table.ranges()
    par = _par(table)
    nz = 1500
    prim, g = zones.make_batch_ye(nz, 5, (1e-12, 1e-4), (0.05, 80.0), (0.06, 0.5))
    cons = np.zeros((9, nz))
    thermo = np.zeros((2, nz))
    L = oracle.lib()
    abi.check("orc_prim2con_batch_ye", L.orc_prim2con_batch_ye(
        nz, prim.ctypes.data, g.ctypes.data, cons.ctypes.data, thermo.ctypes.data,
        C.byref(par)))
    rng = np.random.default_rng(6)
    u = lambda: 1 + 1e-2 * rng.uniform(-1, 1, nz)
    seed = _d(np.stack([prim[0] * u(), thermo[0] * u(), prim[3] * u(), prim[4] * u(),
                        prim[5] * u(), prim[2], prim[1]]))
    out = np.zeros((9, nz))
    flag = np.zeros(nz)
    iters = np.zeros(nz, dtype=np.int32)
    abi.check("orc_c2p_batch_ye", L.orc_c2p_batch_ye(
        nz, cons.ctypes.data, g.ctypes.data, seed.ctypes.data, out.ctypes.data,
        flag.ctypes.data, iters.ctypes.data, C.byref(par)))
    ok = flag == 1
    assert ok.mean() > 0.99, np.unique(flag, return_counts=True)
    assert np.max(np.abs(out[0, ok] / prim[0, ok] - 1)) < 1e-8
    assert np.max(np.abs(out[7, ok] / prim[1, ok] - 1)) < 1e-6  # temperature
    assert np.max(np.abs(out[8, ok] - prim[2, ok])) < 1e-12  # Ye = DYe / D
    assert np.max(np.abs(out[1, ok] / thermo[0, ok] - 1)) < 1e-7
    for i in range(3):
        assert np.max(np.abs(out[3 + i, ok] - prim[3 + i, ok])) < 1e-7
    assert (iters[ok] > 0).all()
