"""Fishbone-Moncrief torus on a Kerr-Schild black hole: the curved-spacetime
workload of the substep.

Reference behaviour followed (FishboneMoncriefIDX):
  * metric: "spherical" Kerr-Schild, i.e. the Kerr metric in ingoing
    Kerr-Schild coordinates (r, theta, phi) with r the Euclidean radius
    sqrt(x^2+y^2+z^2), mapped to Cartesian components and stored on vertices
    (FM_disk_ID.cxx:86-131, FM_disk_implementation.hxx:57-230);
  * disk: Fishbone & Moncrief (1976) constant-l = u_phi u^t torus with l the
    Keplerian value at r_at_max_density, rho from the enthalpy of a polytrope,
    Valencia velocity from u^mu (FM_disk_ID.cxx:134-220,
    FM_disk_implementation.hxx:9-55, 232-296);
  * pressure perturbation p <- p (1 + U[lo, hi]) inside the disk, eps from
    the ideal gas (FM_disk_ID.cxx:224-260); seeded here, where the reference
    uses rand();
  * vector potential A_phi = A_b max(rho/rho_max - rho_cut, 0) on the edges,
    rho averaged from the four cells around each edge (FM_disk_ID.cxx:262-340,
    FM_disk_implementation.hxx:311-340; the reference averages rho along x
    for all three edge directions, here each edge averages across its own
    transverse axes), scaled so that min(2p/b^2) = beta_min.

The formulas are written from the physics, not from the reference's
generated code: gamma_ij = J^a_i J^b_j g_ab with J = d(r,theta,phi)/d(x,y,z),
and K_ij = (D_i beta_j + D_j beta_i) / (2 alpha) (the slices are stationary),
whose derivatives are taken by complex-step differentiation of the closed
form, exact to rounding.  Computation runs in torch on the patch's device.
"""
import math

import numpy as np

from . import abi

PAIRS = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]
GNAMES = ["gxx", "gxy", "gxz", "gyy", "gyz", "gzz"]
KNAMES = ["kxx", "kxy", "kxz", "kyy", "kyz", "kzz"]


def _torch():
    import torch
    return torch


# ---------------------------------------------------------------------------
# Kerr-Schild metric
# ---------------------------------------------------------------------------
def ks_metric(x, y, z, M, a):
    """Lapse, shift (up), 3-metric (6 comps, xx xy xz yy yz zz) and shift
    (down) of spherical Kerr-Schild at Cartesian points.  Works on real or
    complex tensors (every operation is analytic off the z axis)."""
    t = _torch()
    a2 = a * a
    rc2 = x * x + y * y
    R2 = rc2 + z * z
    r = t.sqrt(R2)
    rc = t.sqrt(rc2)
    cos2 = z * z / R2
    sin2 = rc2 / R2
    Sig = R2 + a2 * cos2
    Del = R2 - 2.0 * M * r + a2
    H = 2.0 * M * r / Sig
    A = 1.0 + H
    g_rr = A
    g_rp = -a * sin2 * A
    g_hh = Sig
    g_pp = ((R2 + a2) ** 2 - a2 * Del * sin2) * sin2 / Sig
    zero = t.zeros_like(x)
    Jr = (x / r, y / r, z / r)
    Jh = (x * z / (R2 * rc), y * z / (R2 * rc), -rc / R2)
    Jp = (-y / rc2, x / rc2, zero)
    gam = []
    for i, j in PAIRS:
        gam.append(g_rr * Jr[i] * Jr[j] + g_hh * Jh[i] * Jh[j] +
                   g_pp * Jp[i] * Jp[j] + g_rp * (Jr[i] * Jp[j] + Jp[i] * Jr[j]))
    alp = 1.0 / t.sqrt(A)
    br = H / A                       # beta^r; beta^theta = beta^phi = 0
    beta_up = [(x / r) * br, (y / r) * br, (z / r) * br]
    b_r = H                          # beta_r = g_tr
    b_p = -a * sin2 * H              # beta_phi = g_tphi
    beta_dn = [b_r * Jr[i] + b_p * Jp[i] for i in range(3)]
    return alp, beta_up, gam, beta_dn


def ks_adm(x, y, z, M, a, h=1e-30):
    """All 16 ADM quantities (alp, beta^i, gamma_ij, K_ij) at real points.

    K_ij = [d_i beta_j + d_j beta_i - beta^l (d_i g_jl + d_j g_il - d_l g_ij)]
           / (2 alpha),
    the Killing-slice form of K_ij = -(d_t g_ij - L_beta g_ij) / (2 alpha),
    with every derivative d_l f = Im f(x + i h e_l) / h."""
    t = _torch()
    alp, bu, gam, bd = ks_metric(x, y, z, M, a)
    dg, db = [], []
    xs = [x.to(t.complex128), y.to(t.complex128), z.to(t.complex128)]
    for l in range(3):
        xc = list(xs)
        xc[l] = xc[l] + 1j * h
        _, _, g_c, b_c = ks_metric(xc[0], xc[1], xc[2], M, a)
        dg.append([g.imag / h for g in g_c])
        db.append([b.imag / h for b in b_c])

    def gidx(i, j):
        return PAIRS.index((min(i, j), max(i, j)))

    K = []
    for i, j in PAIRS:
        s = db[i][j] + db[j][i]
        for l in range(3):
            s = s - bu[l] * (dg[i][gidx(j, l)] + dg[j][gidx(i, l)] -
                             dg[l][gidx(i, j)])
        K.append(s / (2.0 * alp))
    out = {"alp": alp, "betax": bu[0], "betay": bu[1], "betaz": bu[2]}
    for n, g in zip(GNAMES, gam):
        out[n] = g
    for n, k in zip(KNAMES, K):
        out[n] = k
    return out


# ---------------------------------------------------------------------------
# Fishbone-Moncrief disk
# ---------------------------------------------------------------------------
def kepler_ell(M, a, r):
    """l = u_phi u^t of the prograde circular equatorial geodesic at r
    (Bardeen, Press & Teukolsky 1972, eqs. 2.12-2.13)."""
    sM = math.sqrt(M)
    den = r ** 0.75 * math.sqrt(r ** 1.5 - 3.0 * M * math.sqrt(r) + 2.0 * a * sM)
    ut = (r ** 1.5 + a * sM) / den
    uphi_dn = sM * (r * r - 2.0 * a * math.sqrt(M * r) + a * a) / den
    return ut * uphi_dn


def _bl(M, a, r, sin2):
    """Sigma, Delta, A = (r^2+a^2)^2 - a^2 Delta sin^2 of Boyer-Lindquist."""
    cos2 = 1.0 - sin2
    Sig = r * r + a * a * cos2
    Del = r * r - 2.0 * M * r + a * a
    A = (r * r + a * a) ** 2 - a * a * Del * sin2
    return Sig, Del, A


def fm_lnh_raw(M, a, ell, r, sin2, log=math.log, sqrt=math.sqrt):
    """ln h up to its additive constant, Fishbone & Moncrief (1976) eq. 3.6
    with l = u_phi u^t: 1/2 ln[(1+X) A/(Sigma Delta)] - X/2 - 2 a M r l / A,
    X = sqrt(1 + 4 l^2 Sigma^2 Delta / (A^2 sin^2))."""
    Sig, Del, A = _bl(M, a, r, sin2)
    X = sqrt(1.0 + 4.0 * ell * ell * Sig * Sig * Del / (A * A * sin2))
    return 0.5 * log((1.0 + X) * A / (Sig * Del)) - 0.5 * X - \
        2.0 * a * M * r * ell / A


def fm_four_velocity(M, a, ell, r, sin2, sqrt=math.sqrt):
    """(u^t, u^phi) of the torus (u^r = u^theta = 0), built in the frame of
    the zero-angular-momentum observer: u^(phi)^2 = (X - 1)/2, u^t =
    sqrt(1 + u^(phi)^2) sqrt(A/(Sigma Delta)), u^phi = omega u^t +
    u^(phi) sqrt(Sigma/A)/sin(theta), omega = 2 a M r / A."""
    Sig, Del, A = _bl(M, a, r, sin2)
    X = sqrt(1.0 + 4.0 * ell * ell * Sig * Sig * Del / (A * A * sin2))
    up = sqrt(0.5 * (X - 1.0))
    gz = sqrt(1.0 + up * up)
    ut = gz * sqrt(A / (Sig * Del))
    uphi = 2.0 * a * M * r * gz / sqrt(A * Sig * Del) + sqrt(Sig / A) * up / \
        sqrt(sin2)
    return ut, uphi


class TorusSpec:
    """Parameters of the torus workload (defaults: a = 0.9375, r_in = 6,
    r_max = 12, Gamma = 4/3 polytrope K = 1e-3, 2% perturbations,
    min beta = 100; free-graded atmosphere as FMdisk_unigrid.par)."""

    def __init__(self, M=1.0, a=0.9375, r_in=6.0, r_max=12.0, gamma=4.0 / 3.0,
                 kappa=1e-3, pert=0.02, rho_cut=0.2, beta_min=100.0,
                 atmo_rho=1e-5, atmo_nrho=1.5, atmo_press=1e-7 / 3.0,
                 atmo_npress=2.5, seed=42):
        self.M, self.a, self.r_in, self.r_max = M, a, r_in, r_max
        self.gamma, self.kappa, self.pert = gamma, kappa, pert
        self.rho_cut, self.beta_min = rho_cut, beta_min
        self.atmo_rho, self.atmo_nrho = atmo_rho, atmo_nrho
        self.atmo_press, self.atmo_npress = atmo_press, atmo_npress
        self.seed = seed
        self.ell = kepler_ell(M, a, r_max)
        self.lnh_in = fm_lnh_raw(M, a, self.ell, r_in, 1.0)
        hm1 = math.exp(fm_lnh_raw(M, a, self.ell, r_max, 1.0) - self.lnh_in) - 1
        self.rho_max = self.rho_of_hm1(hm1)

    def rho_of_hm1(self, hm1):
        g = self.gamma
        return (hm1 * (g - 1.0) / (self.kappa * g)) ** (1.0 / (g - 1.0))


def torus_patch_geometry(n=192, L=40.0, ng=3, offset=0.25):
    """n^3 cells over [-L, L]^3 shifted by offset*dx on every axis, so that no
    vertex lies on the z axis or at r = 0 (as FMdisk_unigrid.par)."""
    dx = 2.0 * L / n
    x0 = -L + offset * dx
    return (n, n, n), ng, (x0, x0, x0), (dx, dx, dx)


def _coords(patch, cent, dev):
    t = _torch()
    z, y, x = patch.coords(cent)
    f = lambda v: t.from_numpy(np.ascontiguousarray(v)).to(dev)
    return f(z), f(y), f(x)


def _set(patch, name, v):
    t = _torch()
    v = v.expand(patch.shape(name)).contiguous() if v.dim() == 3 else v
    if patch.device is None:
        patch[name] = v.cpu().numpy()
    else:
        patch.arr[name] = v.to(t.float64).contiguous()


def _dev(patch):
    return "cpu" if patch.device is None else patch.device


def set_metric(patch, spec, zchunk=16):
    """Kerr-Schild ADM variables on the patch's vertices."""
    t = _torch()
    dev = _dev(patch)
    z, y, x = _coords(patch, 7, dev)
    ez, ey, ex = patch.shape("alp")
    out = {n: t.empty((ez, ey, ex), dtype=t.float64, device=dev)
           for n in abi.ADM}
    for k0 in range(0, ez, zchunk):
        k1 = min(ez, k0 + zchunk)
        zz = z[k0:k1].expand(k1 - k0, ey, ex)
        yy = y.expand(k1 - k0, ey, ex)
        xx = x.expand(k1 - k0, ey, ex)
        res = ks_adm(xx, yy, zz, spec.M, spec.a)
        for n in abi.ADM:
            out[n][k0:k1] = res[n]
    for n in abi.ADM:
        _set(patch, n, out[n])


def _disk_fields(spec, x, y, z):
    """rho, press, eps, v^i and the in-disk mask at cell points."""
    t = _torch()
    M, a, ell = spec.M, spec.a, spec.ell
    R2 = x * x + y * y + z * z
    r = t.sqrt(R2)
    sin2 = (x * x + y * y) / R2
    outside = r <= spec.r_in
    rs = t.where(outside, t.full_like(r, spec.r_max), r)
    lnh = fm_lnh_raw(M, a, ell, rs, sin2, log=t.log, sqrt=t.sqrt) - spec.lnh_in
    hm1 = t.expm1(lnh)
    disk = (~outside) & (hm1 > 0)
    hm1p = t.where(disk, hm1, t.ones_like(hm1))
    g = spec.gamma
    rho_d = (hm1p * (g - 1.0) / (spec.kappa * g)) ** (1.0 / (g - 1.0))
    p_d = spec.kappa * rho_d ** g
    # Valencia velocity v^i = (u^i/u^t + beta^i)/alpha in KS; u^r = u^th = 0
    ut, uphi = fm_four_velocity(M, a, ell, rs, sin2, sqrt=t.sqrt)
    Sig = rs * rs + a * a * (1.0 - sin2)
    H = 2.0 * M * rs / Sig
    alp = 1.0 / t.sqrt(1.0 + H)
    vr = (H / (1.0 + H)) / alp
    vphi = uphi / (alp * ut)
    vx = vr * x / rs - vphi * y
    vy = vr * y / rs + vphi * x
    vz = vr * z / rs
    rho_a = spec.atmo_rho * (r + 1e-100) ** (-spec.atmo_nrho)
    p_a = spec.atmo_press * (r + 1e-100) ** (-spec.atmo_npress)
    zero = t.zeros_like(r)
    rho = t.where(disk, rho_d, rho_a)
    press = t.where(disk, p_d, p_a)
    v = [t.where(disk, c, zero) for c in (vx, vy, vz)]
    return rho, press, v, disk


def set_disk(patch, spec):
    """Hydro primitives on cells, with the seeded pressure perturbation."""
    t = _torch()
    dev = _dev(patch)
    z, y, x = _coords(patch, 0, dev)
    shp = patch.shape("rho")
    z, y, x = z.expand(shp), y.expand(shp), x.expand(shp)
    rho, press, v, disk = _disk_fields(spec, x, y, z)
    rng = np.random.default_rng(spec.seed)
    u = t.from_numpy(rng.uniform(-spec.pert, spec.pert, size=shp)).to(dev)
    press = t.where(disk, press * (1.0 + u), press)
    eps = press / ((rho + 1e-300) * (spec.gamma - 1.0))
    for n, val in (("rho", rho), ("press", press), ("eps", eps),
                   ("velx", v[0]), ("vely", v[1]), ("velz", v[2])):
        _set(patch, n, val)
    for n in ("Bvecx", "Bvecy", "Bvecz"):
        _set(patch, n, t.zeros(shp, dtype=t.float64, device=dev))
    return disk


def _edge_avg(rho, axis):
    """Average of the four cells around each edge along `axis` (0=x, 1=y,
    2=z; array axes are (z, y, x)); cells past the patch edge repeat the
    outermost one."""
    t = _torch()
    tr = [d for d in (0, 1, 2) if d != axis]     # transverse axes (x=0 ...)
    P = rho
    for d in tr:
        ad = 2 - d
        first = P.narrow(ad, 0, 1)
        last = P.narrow(ad, P.shape[ad] - 1, 1)
        P = t.cat([first, P, last], dim=ad)
    a1, a2 = 2 - tr[0], 2 - tr[1]
    n1, n2 = P.shape[a1] - 1, P.shape[a2] - 1
    s = (P.narrow(a1, 0, n1).narrow(a2, 0, n2) +
         P.narrow(a1, 1, n1).narrow(a2, 0, n2) +
         P.narrow(a1, 0, n1).narrow(a2, 1, n2) +
         P.narrow(a1, 1, n1).narrow(a2, 1, n2))
    return 0.25 * s


def set_avec(patch, spec, A_b=1.0):
    """A_x = -A_phi y / rc^2, A_y = A_phi x / rc^2, A_z = 0 with
    A_phi = A_b max(rho/rho_max - rho_cut, 0) on the edges."""
    t = _torch()
    dev = _dev(patch)
    rho = patch["rho"]
    rho = rho if patch.device is not None else t.from_numpy(rho)
    rho = rho.to(dev)
    cents = {"Avec_x": 6, "Avec_y": 5, "Avec_z": 3}
    for axis, name in enumerate(("Avec_x", "Avec_y", "Avec_z")):
        z, y, x = _coords(patch, cents[name], dev)
        shp = patch.shape(name)
        if axis == 2:
            _set(patch, name, t.zeros(shp, dtype=t.float64, device=dev))
            continue
        re = _edge_avg(rho, axis)
        assert tuple(re.shape) == tuple(shp), (re.shape, shp)
        x, y = x.expand(shp), y.expand(shp)
        rc2 = t.clamp(x * x + y * y, min=1e-30)
        Aphi = A_b * t.clamp(re / spec.rho_max - spec.rho_cut, min=0.0)
        comp = -Aphi * y / rc2 if axis == 0 else Aphi * x / rc2
        _set(patch, name, comp)


def _v2c(a):
    return 0.125 * (a[:-1, :-1, :-1] + a[1:, :-1, :-1] + a[:-1, 1:, :-1] +
                    a[1:, 1:, :-1] + a[:-1, :-1, 1:] + a[1:, :-1, 1:] +
                    a[:-1, 1:, 1:] + a[1:, 1:, 1:])


def plasma_beta_min(patch, disk):
    """min over interior disk cells of 2 p / b^2, with
    b^2 = B^2 / W^2 + (B^i v_i)^2 (Valencia v, cell-averaged metric)."""
    t = _torch()
    dev = _dev(patch)

    def T(n):
        v = patch[n]
        return (v if patch.device is not None else t.from_numpy(v)).to(dev)

    g = [_v2c(T(n)) for n in GNAMES]
    gm = [[g[0], g[1], g[2]], [g[1], g[3], g[4]], [g[2], g[4], g[5]]]
    B = [T("Bvecx"), T("Bvecy"), T("Bvecz")]
    v = [T("velx"), T("vely"), T("velz")]
    vl = [sum(gm[i][j] * v[j] for j in range(3)) for i in range(3)]
    Bl = [sum(gm[i][j] * B[j] for j in range(3)) for i in range(3)]
    v2 = sum(vl[i] * v[i] for i in range(3))
    B2 = sum(Bl[i] * B[i] for i in range(3))
    Bv = sum(B[i] * vl[i] for i in range(3))
    b2 = B2 * (1.0 - v2) + Bv * Bv
    sl = patch.interior("rho")
    p = T("press")[sl]
    b2 = b2[sl]
    m = disk[sl] & (b2 > 0)
    if not bool(m.any()):
        raise RuntimeError("torus: no magnetised cell on the patch")
    return float((2.0 * p[m] / b2[m]).min())


def eos_params(spec, **kw):
    """asx_params of the torus workload: ideal gas Gamma, floors rho 1e-8 and
    p 1e-10 (use_press_atmo), EOS/c2p caps above the disk's values."""
    base = dict(gl_gamma=spec.gamma, poly_gamma=spec.gamma,
                poly_k=spec.kappa, rho_abs_min=1e-8, p_atmo=1e-10,
                use_press_atmo=1, rho_max=1e8, eps_max=1e8,
                c2p_prime="Noble", c2p_second="Palenzuela")
    base.update(kw)
    return abi.params(**base)


def build_torus(patch, backend, par, spec=None, bc="copy"):
    """Fill `patch` (host or device) with the torus workload and run the
    initial group: metric, disk, A scaled to min beta, B from A,
    check_prims, prim2con.  Returns (spec, Evolution, disk mask)."""
    from .driver import Evolution, check_prims_ideal
    from .seeds import post_initial
    spec = spec or TorusSpec()
    set_metric(patch, spec)
    disk = set_disk(patch, spec)
    set_avec(patch, spec, 1.0)
    ev = Evolution(backend, patch, par, bc=bc)
    b = backend
    b.bstag_from_A(patch.struct())
    ev.sync(abi.DBSTAG)
    b.dB_from_bstag(patch.struct())
    ev.sync(abi.DB)
    b.B_from_dB(patch.struct())
    beta0 = plasma_beta_min(patch, disk)
    scale = math.sqrt(beta0 / spec.beta_min)
    for n in ("Avec_x", "Avec_y"):
        if patch.device is None:
            patch[n] = patch[n] * scale
        else:
            patch.arr[n].mul_(scale)
    post_initial(patch, spec.gamma)
    ev.initial(check_prims_ideal)
    return spec, ev, disk


def eos_from_rho_temp(backend, par, rho, temp, ye):
    """press, eps, entropy from the EOS at (rho, T, Ye), arrays of one
    shape on the backend's side (torch on the device, numpy on the host)."""
    import ctypes as C
    t = _torch()
    if backend.prefix == "asx_":
        r, T, y = (a.contiguous() for a in (rho, temp, ye))
        out = t.empty((4,) + tuple(r.shape), dtype=t.float64, device=r.device)
        backend._call("eos_from_rho_temp", r.numel(), C.c_void_p(r.data_ptr()),
                      C.c_void_p(T.data_ptr()), C.c_void_p(y.data_ptr()),
                      C.c_void_p(out.data_ptr()), C.byref(par), backend.stream)
        return out[0], out[1], out[3]
    r, T, y = (np.ascontiguousarray(np.asarray(a), dtype=np.float64) for a in (rho, temp, ye))
    out = np.zeros((4,) + r.shape)
    backend._call("eos_from_rho_temp", r.size, r.ctypes.data, T.ctypes.data, y.ctypes.data,
                  out.ctypes.data, C.byref(par))
    out = t.from_numpy(out)
    return out[0], out[1], out[3]



def _saved_prims_table(patch, par):
    """AsterX_CheckPrims (AsterX/src/check_prims.cxx:24-228) for initial
    data that already lies inside the floors and ceilings (rho above the
    atmosphere cut and below the table, T at or above the atmosphere's,
    Ye inside the table, |v| below the limit; build_torus_tab makes it so
    and checks it): the saved primitives and the z and s vectors.  numpy on
    interior cells, cell-averaged metric."""
    sl = patch.interior("rho")
    g = {c: _v2c(_arr(patch, c)).numpy() if patch.device is None else
         _v2c(_arr(patch, c)).cpu().numpy() for c in GNAMES}
    rho, eps, press, ye = (patch.numpy(n) for n in ("rho", "eps", "press", "Ye"))
    v = [patch.numpy(n) for n in ("velx", "vely", "velz")]
    gm = [[g["gxx"], g["gxy"], g["gxz"]], [g["gxy"], g["gyy"], g["gyz"]],
          [g["gxz"], g["gyz"], g["gzz"]]]
    vl = [sum(gm[i][j] * v[j] for j in range(3)) for i in range(3)]
    v2 = sum(vl[i] * v[i] for i in range(3))
    w_lim = math.sqrt(1.0 + par.vw_lim ** 2)
    if float(np.sqrt(v2[sl]).max()) > par.vw_lim / w_lim:
        raise RuntimeError("torus on a table: velocity above the limit")
    w = 1.0 / np.sqrt(1.0 - v2)
    vals = {"saved_rho": rho, "saved_eps": eps, "saved_Ye": ye,
            "saved_velx": v[0], "saved_vely": v[1], "saved_velz": v[2]}
    for i, c in enumerate("xyz"):
        vals["zvec_" + c] = w * v[i]
        vals["svec_" + c] = (rho + rho * eps + press) * w * w * v[i]
    for name, val in vals.items():
        if name not in patch:
            patch.alloc(name)
        full = patch.numpy(name).copy()
        full[sl] = val[sl]
        patch[name] = full


def _arr(patch, name):
    t = _torch()
    a = patch[name]
    return a if patch.device is not None else t.from_numpy(np.asarray(a))


def build_torus_tab(patch, backend, par, table_ranges, spec=None, bc="copy", rho_disk=1e-4,
                    t_disk=10.0, t_atmo=0.02):
    """The torus on a table (par carries the table): the ideal-gas disk of
    build_torus (velocities, and A_phi from its density) with the density
    scaled to rho_disk at the disk maximum (code units),
    T = t_atmo + t_disk (rho/rho_max)^(2/3) MeV,
    Ye = 0.5 - 0.4 (rho/rho_max)^(1/2), and eps, p and the entropy from the
    table at (rho, T, Ye); A scaled to the minimum beta; B from A; the
    initial-data check; prim2con.  table_ranges = ((rho_lo, rho_hi),
    (T_lo, T_hi), (Ye_lo, Ye_hi)) of the table, in code units.  Returns
    (spec, Evolution, disk mask)."""
    from .driver import Evolution
    t = _torch()
    spec = spec or TorusSpec()
    set_metric(patch, spec)
    disk = set_disk(patch, spec)
    set_avec(patch, spec, 1.0)               # from the ideal-gas density
    rho = _arr(patch, "rho")
    rmax = float(rho[disk].max())
    rel = rho / rmax
    temp = t_atmo + t_disk * rel ** (2.0 / 3.0)
    ye = t.clamp(0.5 - 0.4 * t.sqrt(rel), 0.05, 0.5)
    (r0, r1), (T0, T1), (y0, y1) = table_ranges
    # the atmosphere of check_prims.cxx:47-53 (ungraded: r_atmo beyond the
    # patch) where the scaled density falls below its cut
    rho_atm = max(par.rho_abs_min, r0)
    rho_t = rho * (rho_disk / rmax)
    rho_t = t.where(rho_t < rho_atm * (1.0 + par.atmo_tol), t.full_like(rho_t, rho_atm), rho_t)
    if not (par.r_atmo > 1e3 and par.rho_min <= rho_atm and float(rho_t.max()) < r1 and
            t_atmo >= max(T0, par.t_atmo) and float(temp.max()) < T1 and 0.05 >= y0 and
            0.5 <= y1):
        raise RuntimeError("torus on a table: initial data outside the table or the floors")
    press, eps, ent = eos_from_rho_temp(backend, par, rho_t, temp, ye)
    for n, val in (("rho", rho_t), ("temperature", temp), ("Ye", ye), ("press", press),
                   ("eps", eps), ("entropy", ent)):
        _set(patch, n, val)
    ev = Evolution(backend, patch, par, bc=bc)
    b = backend
    b.bstag_from_A(patch.struct())
    ev.sync(abi.DBSTAG)
    b.dB_from_bstag(patch.struct())
    ev.sync(abi.DB)
    b.B_from_dB(patch.struct())
    beta0 = plasma_beta_min(patch, disk)
    scale = math.sqrt(beta0 / spec.beta_min)
    for n in ("Avec_x", "Avec_y"):
        if patch.device is None:
            patch[n] = patch[n] * scale
        else:
            patch.arr[n].mul_(scale)
    ev.initial(_saved_prims_table)
    return spec, ev, disk
