"""Initial-data recipes (numpy, host side) for the patch workloads.

These restate the reference's initial-data thorns as input generators:
  balsara()        AsterSeeds/src/1D_tests.cxx:165-481 (Balsara 1-5 shock
                   tubes, vector potential of :384-478)
  post_initial()   AsterSeeds/src/SetTempEntropyYe.cxx (Ye, T, entropy)
  minkowski()      ADMBaseX "Cartesian Minkowski"
  kerr_schild(), fm_torus()  in torus.py
They produce inputs only; nothing here is on the measured path.
"""
import math

import numpy as np

BALSARA = {
    # rho, vx, vy, vz, p, Bx, By, Bz  (left, right)
    "Balsara1": ((1.0, 0.0, 0.0, 0.0, 1.0, 0.5, 1.0, 0.0),
                 (0.125, 0.0, 0.0, 0.0, 0.1, 0.5, -1.0, 0.0)),
    "Balsara2": ((1.0, 0.0, 0.0, 0.0, 30.0, 5.0, 6.0, 6.0),
                 (1.0, 0.0, 0.0, 0.0, 1.0, 5.0, 0.7, 0.7)),
    "Balsara3": ((1.0, 0.0, 0.0, 0.0, 1000.0, 10.0, 7.0, 7.0),
                 (1.0, 0.0, 0.0, 0.0, 0.1, 10.0, 0.7, 0.7)),
    "Balsara4": ((1.0, 0.999, 0.0, 0.0, 0.1, 10.0, 7.0, 7.0),
                 (1.0, -0.999, 0.0, 0.0, 0.1, 10.0, -7.0, -7.0)),
    "Balsara5": ((1.08, 0.4, 0.3, 0.2, 0.95, 2.0, 0.3, 0.3),
                 (1.0, -0.45, -0.2, 0.2, 1.0, 2.0, -0.7, 0.5)),
}


def _rot(th_x, th_y, th_z, seq):
    """Rotation of 1D_tests.cxx:311-331: R = Rs[seq2] Rs[seq1] Rs[seq0]."""
    cx, sx = math.cos(th_x), math.sin(th_x)
    cy, sy = math.cos(th_y), math.sin(th_y)
    cz, sz = math.cos(th_z), math.sin(th_z)
    Rx = np.array([[1, 0, 0], [0, cx, sx], [0, -sx, cx]])
    Ry = np.array([[cy, 0, -sy], [0, 1, 0], [sy, 0, cy]])
    Rz = np.array([[cz, sz, 0], [-sz, cz, 0], [0, 0, 1]])
    Rs = [Rx, Ry, Rz]
    return Rs[seq[2]] @ Rs[seq[1]] @ Rs[seq[0]]


def minkowski(patch):
    for name in ("alp", "gxx", "gyy", "gzz"):
        patch[name] = 1.0
    for name in ("betax", "betay", "betaz", "gxy", "gxz", "gyz",
                 "kxx", "kxy", "kxz", "kyy", "kyz", "kzz"):
        patch[name] = 0.0


def balsara(patch, case="Balsara1", shock_dir="x", gamma=2.0, perturb=None,
            seed=100):
    """Balsara shock tube on every point of the patch (ghosts included).

    perturb=a multiplies every cell's rho and p by (1 + a*u), u ~ U[-1, 1]
    from numpy's PCG64 seeded with `seed`, before eps is computed.
    """
    L, R = BALSARA[case]
    th = {"x": (0.0, 0.0, 0.0), "y": (0.5 * math.pi, 0.0, 0.5 * math.pi),
          "z": (0.0, -0.5 * math.pi, -0.5 * math.pi)}[shock_dir]
    Rm = _rot(th[0], th[1], th[2], [2, 1, 0])
    Rinv = _rot(-th[0], -th[1], -th[2], [0, 1, 2])
    Bl, Br = Rinv @ np.array(L[5:8]), Rinv @ np.array(R[5:8])
    vl, vr = Rinv @ np.array(L[1:4]), Rinv @ np.array(R[1:4])

    def newx(z, y, x):
        return Rm[0, 0] * x + Rm[0, 1] * y + Rm[0, 2] * z

    z, y, x = patch.coords(0)
    left = np.broadcast_to(newx(z, y, x) <= 0.0, patch.shape("rho"))
    rho = np.where(left, L[0], R[0])
    press = np.where(left, L[4], R[4])
    if perturb:
        rng = np.random.default_rng(seed)
        u1 = rng.uniform(-1.0, 1.0, size=rho.shape)
        u2 = rng.uniform(-1.0, 1.0, size=rho.shape)
        rho = rho * (1.0 + perturb * u1)
        press = press * (1.0 + perturb * u2)
    patch["rho"] = rho
    patch["press"] = press
    for i, name in enumerate(("velx", "vely", "velz")):
        patch[name] = np.where(left, vl[i], vr[i])
    patch["eps"] = press / (rho * (gamma - 1.0))
    for i, B in enumerate(("Bvecx", "Bvecy", "Bvecz")):
        patch[B] = np.where(left, Bl[i], Br[i])

    # 1D_tests.cxx:392-426 (rotation angles zero): A_x = B_y z, A_y = B_z x,
    # A_z = B_x y, the side chosen by the rotated x of the A point.
    def avec(name, comp, coord_axis):
        z, y, x = patch.coords({"Avec_x": 6, "Avec_y": 5, "Avec_z": 3}[name])
        lft = newx(z, y, x) <= 0.0
        c = {0: x, 1: y, 2: z}[coord_axis]
        patch[name] = np.broadcast_to(np.where(lft, Bl[comp], Br[comp]) * c,
                                      patch.shape(name))

    avec("Avec_x", 1, 2)
    avec("Avec_y", 2, 0)
    avec("Avec_z", 0, 1)


def post_initial(patch, gamma, particle_mass=1.0, Ye_atmo=0.5):
    """SetYe / SetTemp / SetEntropy (AsterSeeds/src/SetTempEntropyYe.cxx),
    ideal gas: T = (gamma-1) m eps, entropy = kappa = (gamma-1) eps rho^(1-gamma)."""
    patch["Ye"] = Ye_atmo
    rho = patch.numpy("rho")
    eps = patch.numpy("eps")
    patch["temperature"] = (gamma - 1) * particle_mass * eps
    patch["entropy"] = (gamma - 1.0) * eps * np.power(rho, 1.0 - gamma)


def _set_hydro(patch, rho, vel, press, gamma):
    shp = patch.shape("rho")
    patch["rho"] = np.broadcast_to(rho, shp)
    for name, v in zip(("velx", "vely", "velz"), vel):
        patch[name] = np.broadcast_to(v, shp)
    patch["press"] = np.broadcast_to(press, shp)
    patch["eps"] = np.broadcast_to(press / (rho * (gamma - 1.0)), shp)
    for name in ("Bvecx", "Bvecy", "Bvecz"):
        patch[name] = 0.0


def _set_avec(patch, fx, fy, fz):
    for name, cent, f in (("Avec_x", 6, fx), ("Avec_y", 5, fy), ("Avec_z", 3, fz)):
        z, y, x = patch.coords(cent)
        patch[name] = np.broadcast_to(f(x, y, z), patch.shape(name)).astype(np.float64)


def sound_wave(patch, gamma, amplitude=1e-3):
    """AsterSeeds/src/1D_tests.cxx:66-92 ("sound wave"; amplitude default
    AsterSeeds/param.ccl:78-81): rho = p = 1, v^x = amplitude sin(pi x), A = 0."""
    z, y, x = patch.coords(0)
    _set_hydro(patch, 1.0, (amplitude * np.sin(math.pi * x), 0.0, 0.0), 1.0, gamma)
    zero = lambda x, y, z: 0.0 * (x + y + z)
    _set_avec(patch, zero, zero, zero)


def alfven_wave(patch, gamma):
    """AsterSeeds/src/1D_tests.cxx:94-127 ("Alfven wave"): rho = 1, p = 0.5,
    v = -va A0 (0, cos kx, sin kx), A = (z cos kx - y sin kx, -z/2, y/2),
    A0 = 1, va = 0.5, k = 2 pi."""
    A0, va, k = 1.0, 0.5, 2.0 * math.pi
    z, y, x = patch.coords(0)
    _set_hydro(patch, 1.0, (0.0, -va * A0 * np.cos(k * x), -va * A0 * np.sin(k * x)),
               0.5, gamma)
    _set_avec(patch, lambda x, y, z: z * np.cos(k * x) - y * np.sin(k * x),
              lambda x, y, z: -z / 2.0 + 0.0 * (x + y),
              lambda x, y, z: y / 2.0 + 0.0 * (x + z))


def magnetic_rotor(patch, gamma):
    """AsterSeeds/src/2D_tests.cxx:33-73 ("magnetic rotor"): inside
    r_cyl < 0.1 rho = 10 rotating with omega = 9.95, outside rho = 1 at rest,
    p = 1; A = (0, 0, y)."""
    z, y, x = patch.coords(0)
    shp = patch.shape("rho")
    x, y = np.broadcast_to(x, shp), np.broadcast_to(y, shp)
    inside = np.sqrt(x * x + y * y) < 0.1
    omega = 9.95
    rho = np.where(inside, 10.0, 1.0)
    vx = np.where(inside, -omega * y, 0.0)
    vy = np.where(inside, omega * x, 0.0)
    _set_hydro(patch, rho, (vx, vy, 0.0), 1.0, gamma)
    zero = lambda x, y, z: 0.0 * (x + y + z)
    _set_avec(patch, zero, zero, lambda x, y, z: y + 0.0 * (x + z))


def cylindrical_blast(patch, gamma):
    """AsterSeeds/src/2D_tests.cxx:135-190 ("cylindrical blast"): rho 1e-2,
    p 1 inside r_cyl < 0.8, rho 1e-4, p 3e-5 outside r_cyl > 1, both
    log-linear in r_cyl between; at rest; A = (0, 0, 0.1 y)."""
    z, y, x = patch.coords(0)
    shp = patch.shape("rho")
    x, y = np.broadcast_to(x, shp), np.broadcast_to(y, shp)
    r = np.sqrt(x * x + y * y)
    r_in, r_out = 0.8, 1.0
    rho_in, rho_out, p_in, p_out = 1e-2, 1e-4, 1.0, 3.0e-5

    def prof(a_in, a_out):
        mid = np.exp(((r_out - r) * math.log(a_in) + (r - r_in) * math.log(a_out)) /
                     (r_out - r_in))
        return np.where(r < r_in, a_in, np.where(r > r_out, a_out, mid))
    _set_hydro(patch, prof(rho_in, rho_out), (0.0, 0.0, 0.0), prof(p_in, p_out), gamma)
    zero = lambda x, y, z: 0.0 * (x + y + z)
    _set_avec(patch, zero, zero, lambda x, y, z: 0.1 * y + 0.0 * (x + z))
