"""Seeded synthetic patch states for parity tests.

Smooth fields plus seeded noise on every point (ghosts included), with a
curved, non-diagonal metric, non-zero shift, lapse and extrinsic curvature,
so every term of the stencil and of con2prim is exercised.  Conserved
variables come from the Valencia prim -> cons map written here in numpy.
"""
import numpy as np

from asterx_amd import abi
from asterx_amd.driver import all_slots
from asterx_amd.patch import Patch


def _v2c(a):
    s = 0.0
    for dk in (0, 1):
        for dj in (0, 1):
            for di in (0, 1):
                s = s + a[dk:a.shape[0] - 1 + dk, dj:a.shape[1] - 1 + dj,
                          di:a.shape[2] - 1 + di]
    return s / 8.0


def prim2con_np(g, rho, eps, press, vel, B, Ye, ent):
    """Valencia prim -> cons (densitized), as Con2PrimFactory cons.hxx:49-91."""
    gxx, gxy, gxz, gyy, gyz, gzz = g
    det = 2 * gxy * gxz * gyz + gxx * gyy * gzz - gyy * gxz ** 2 - \
        gxx * gyz ** 2 - gzz * gxy ** 2
    sq = np.sqrt(det)

    def low(v):
        return [gxx * v[0] + gxy * v[1] + gxz * v[2],
                gxy * v[0] + gyy * v[1] + gyz * v[2],
                gxz * v[0] + gyz * v[1] + gzz * v[2]]
    vl = low(vel)
    Bl = low(B)
    w = 1.0 / np.sqrt(1.0 - sum(vl[i] * vel[i] for i in range(3)))
    bst = w * sum(B[i] * vl[i] for i in range(3))
    bl = [Bl[i] / w + bst * vl[i] for i in range(3)]
    bs2 = (sum(B[i] * Bl[i] for i in range(3)) + bst * bst) / (w * w)
    dens = sq * rho * w
    hW2 = w * w * (rho * (1.0 + eps) + press + bs2)
    mom = [sq * (hW2 * vl[i] - bst * bl[i]) for i in range(3)]
    tau = sq * (hW2 - (press + 0.5 * bs2) - bst * bst) - dens
    return dens, mom, tau, [sq * B[i] for i in range(3)], dens * Ye, dens * ent


def _table_thermo(table, par, rho, temp, ye):
    """press, eps, entropy from the oracle's EOS at (rho, T, Ye)."""
    import ctypes as C
    import oracle
    shp = rho.shape
    r, t, y = (np.ascontiguousarray(a, dtype=np.float64).ravel() for a in (rho, temp, ye))
    out = np.zeros((4, r.size))
    abi.check("orc_eos_from_rho_temp", oracle.lib().orc_eos_from_rho_temp(
        r.size, r.ctypes.data, t.ctypes.data, y.ctypes.data, out.ctypes.data, C.byref(par)))
    return out[0].reshape(shp), out[1].reshape(shp), out[3].reshape(shp)


def make_state(n=(18, 14, 12), ng=3, curved=True, seed=7, gamma=5.0 / 3.0,
               uct=True, zero_entropy_block=False, jump=False, table_par=None):
    """With table_par (asx_params of a host table): rho ~ 1e-6 (code),
    T ~ 5 MeV, Ye ~ 0.3, and eps, press, entropy from the table; magnetic
    field and vector potential scaled by 1e-3."""
    rng = np.random.default_rng(seed)
    dx = [0.1, 0.12, 0.09]
    p = Patch(n, ng, (-0.9, -0.8, -0.6), dx, slots=all_slots(use_uct=uct))

    def noise(name, amp):
        return amp * rng.uniform(-1.0, 1.0, size=p.shape(name))

    # ---- metric on vertices -------------------------------------------
    z, y, x = p.coords(7)
    r2 = x * x + y * y + z * z
    shp = p.shape("alp")
    if curved:
        p["alp"] = np.broadcast_to(1.0 - 0.2 * np.exp(-r2), shp)
        p["betax"] = np.broadcast_to(0.05 * np.sin(x + 2 * y), shp)
        p["betay"] = np.broadcast_to(-0.04 * np.cos(y - z), shp)
        p["betaz"] = np.broadcast_to(0.03 * x * z, shp)
        p["gxx"] = np.broadcast_to(1.0 + 0.1 * np.exp(-r2) + 0.02 * x * x, shp)
        p["gxy"] = np.broadcast_to(0.03 * np.sin(x * y), shp)
        p["gxz"] = np.broadcast_to(-0.02 * np.cos(z) * x, shp)
        p["gyy"] = np.broadcast_to(1.0 + 0.08 * np.exp(-r2) + 0.01 * y * y, shp)
        p["gyz"] = np.broadcast_to(0.015 * y * z, shp)
        p["gzz"] = np.broadcast_to(1.0 + 0.12 * np.exp(-r2), shp)
        for c, a in zip(["kxx", "kxy", "kxz", "kyy", "kyz", "kzz"],
                        [0.01, 0.004, -0.003, 0.02, 0.002, -0.01]):
            p[c] = np.broadcast_to(a * np.cos(x + y - z), shp)
    else:
        from asterx_amd import seeds
        seeds.minkowski(p)

    # ---- primitives on cells ------------------------------------------
    z, y, x = p.coords(0)
    cs = p.shape("rho")
    rho = 1.0 + 0.3 * np.sin(2.0 * x) * np.cos(1.5 * y) + 0.1 * z
    press = 0.6 + 0.2 * np.cos(x - y) * np.sin(z + 0.3)
    if jump:
        rho = np.where(x > 0.05, 1e-3 * rho, rho)
        press = np.where(x > 0.05, 1e-4 * press, press)
    rho = np.broadcast_to(rho, cs) * (1.0 + noise("rho", 1e-3))
    press = np.broadcast_to(press, cs) * (1.0 + noise("press", 1e-3))
    vel = [np.broadcast_to(0.2 * np.sin(y + z), cs) + noise("velx", 0.01),
           np.broadcast_to(-0.15 * np.cos(x), cs) + noise("vely", 0.01),
           np.broadcast_to(0.1 * np.sin(x * y), cs) + noise("velz", 0.01)]
    B = [np.broadcast_to(0.4 + 0.1 * np.cos(y), cs) + noise("Bvecx", 0.01),
         np.broadcast_to(0.3 * np.sin(x + z), cs) + noise("Bvecy", 0.01),
         np.broadcast_to(-0.2 + 0.05 * x, cs) + noise("Bvecz", 0.01)]
    eps = press / ((gamma - 1.0) * rho)
    ent = (gamma - 1.0) * eps * np.power(rho, 1.0 - gamma)
    Ye = 0.5 + noise("Ye", 0.01)
    temp = (gamma - 1.0) * eps
    bscale = 1.0
    if table_par is not None:
        rho = 1e-6 * rho
        temp = 5.0 * press / 0.6
        Ye = 0.3 + noise("Ye", 0.02)
        press, eps, ent = _table_thermo(None, table_par, rho, temp, Ye)
        bscale = 1e-3
        B = [bscale * b for b in B]
    if zero_entropy_block:
        ent = ent.copy()
        ent[ng + 2:ng + 4, ng + 3:ng + 5, ng + 4:ng + 6] = 0.0
    p["rho"], p["press"], p["eps"], p["entropy"], p["Ye"] = rho, press, eps, ent, Ye
    p["temperature"] = temp
    for i, c in enumerate("xyz"):
        p["vel" + c] = vel[i]
        p["Bvec" + c] = B[i]

    # metric at cell centres for zvec/svec and cons
    gc = [_v2c(p.numpy(c)) for c in ("gxx", "gxy", "gxz", "gyy", "gyz", "gzz")]
    vl = [gc[0] * vel[0] + gc[1] * vel[1] + gc[2] * vel[2],
          gc[1] * vel[0] + gc[3] * vel[1] + gc[4] * vel[2],
          gc[2] * vel[0] + gc[4] * vel[1] + gc[5] * vel[2]]
    w = 1.0 / np.sqrt(1.0 - (vl[0] * vel[0] + vl[1] * vel[1] + vl[2] * vel[2]))
    for i, c in enumerate("xyz"):
        p["zvec_" + c] = w * vel[i]
        p["svec_" + c] = (rho + rho * eps + press) * w * w * vel[i]
    dens, mom, tau, dB, DYe, DEnt = prim2con_np(gc, rho, eps, press, vel, B, Ye, ent)
    p["dens"], p["tau"], p["DYe"], p["DEnt"] = dens, tau, DYe, DEnt
    for i, c in enumerate("xyz"):
        p["mom" + c] = mom[i]
        p["dB" + c] = dB[i]
    # seeds: true prims, perturbed
    p["saved_rho"] = rho * (1.0 + noise("saved_rho", 1e-3))
    p["saved_eps"] = eps * (1.0 + noise("saved_eps", 1e-3))
    p["saved_Ye"] = Ye
    for c in "xyz":
        p["saved_vel" + c] = p.numpy("vel" + c) * (1.0 + noise("saved_vel" + c, 1e-3))

    # ---- vector potential, Psi, staggered field -----------------------
    for name, f in (("Avec_x", lambda z, y, x: 0.3 * z - 0.1 * np.sin(y)),
                    ("Avec_y", lambda z, y, x: 0.2 * x + 0.05 * np.cos(z)),
                    ("Avec_z", lambda z, y, x: 0.4 * y + 0.03 * x * x)):
        zz, yy, xx = p.coords(abi.centering(name))
        p[name] = bscale * (np.broadcast_to(f(zz, yy, xx), p.shape(name)) + noise(name, 1e-3))
    zz, yy, xx = p.coords(7)
    p["Psi"] = np.broadcast_to(0.01 * np.sin(xx + yy + zz), p.shape("Psi")) + \
        noise("Psi", 1e-3)
    for d, name in enumerate(abi.DBSTAG):
        zz, yy, xx = p.coords(1 << d)
        p[name] = np.broadcast_to(B[d].mean() + bscale * 0.01 * np.sin(xx - zz),
                                  p.shape(name)) + bscale * noise(name, 1e-3)
    return p


def interior_vals(p, name, halo=0):
    return p.numpy(name)[p.interior(name, halo)]


def rel_err(a, b):
    """max|a-b| / max(|b|, tiny) over finite entries; NaN positions must match."""
    a = np.asarray(a)
    b = np.asarray(b)
    na, nb = np.isnan(a), np.isnan(b)
    if not np.array_equal(na, nb):
        return np.inf
    m = ~nb
    if not m.any():
        return 0.0
    d = np.abs(a[m] - b[m]).max()
    s = np.abs(b[m]).max()
    return d / s if s > 0 else d


def kat_state(gamma=2.0):
    """The round-trip state of Con2PrimFactory/src/test.cxx:21-247 on a
    2^3-cell flat patch: rho 0.125, eps 0.8, Ye 0.5, v = 0, B = (0.5, -0.5,
    0), entropy = kappa = (Gamma - 1) eps rho^(1 - Gamma); cons from the
    Valencia map (densitized), saved primitives = the state."""
    p = Patch((2, 2, 2), 1, (0.0, 0.0, 0.0), (0.1, 0.1, 0.1), slots=all_slots())
    p["alp"] = 1.0
    for c in ("gxx", "gyy", "gzz"):
        p[c] = 1.0
    rho, eps, Ye = 0.125, 0.8, 0.5
    press = (gamma - 1.0) * rho * eps
    ent = (gamma - 1.0) * eps * rho ** (1.0 - gamma)
    p["rho"], p["eps"], p["press"], p["Ye"], p["entropy"] = rho, eps, press, Ye, ent
    p["temperature"] = (gamma - 1.0) * eps
    p["Bvecx"], p["Bvecy"] = 0.5, -0.5
    g = [np.ones(p.shape("rho")) if c in (0, 3, 5) else np.zeros(p.shape("rho"))
         for c in range(6)]
    vel = [np.zeros(p.shape("rho"))] * 3
    B = [p.numpy("Bvec" + c) for c in "xyz"]
    dens, mom, tau, dB, DYe, DEnt = prim2con_np(g, p.numpy("rho"), p.numpy("eps"),
                                                p.numpy("press"), vel, B, Ye, ent)
    p["dens"], p["tau"], p["DYe"], p["DEnt"] = dens, tau, DYe, DEnt
    for i, c in enumerate("xyz"):
        p["mom" + c] = mom[i]
        p["dB" + c] = dB[i]
    p["saved_rho"], p["saved_eps"], p["saved_Ye"] = rho, eps, Ye
    return p


def kat_params(prime):
    """Con2PrimFactory/src/test.cxx:37-65: atmosphere rho 1e-10, eps 1e-8;
    max_iter 100, tol 1e-8, alp_thresh -1, cons_error_limit -1, vw_lim 1,
    B_lim 1, BH limits 1e20, Ye_lenient; one solver."""
    return abi.params(c2p_prime=prime, c2p_second="None", gl_gamma=2.0, max_iter=100,
                      c2p_tol=1e-8, alp_thresh=-1.0, cons_error_limit=-1.0, vw_lim=1.0,
                      B_lim=1.0, rho_BH=1e20, eps_BH=1e20, vwlim_BH=1e20, Ye_lenient=1,
                      rho_abs_min=1e-10, eps_atmo=1e-8)
