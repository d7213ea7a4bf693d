"""Independent-zone con2prim workloads (no grid).

make_batch() draws primitive states from the learned-guess distribution
(rho log-uniform in [1e-10, 1e-2], p/rho log-uniform in [1e-6, 1e2], Lorentz
factor uniform in [1, 10] along a random direction, magnetisation b^2/rho
log-uniform in [1e-4, 1e2] along a random direction, SPD 3-metric within 10%
of the identity), maps them to densitized conserved variables with the
Valencia prim -> cons map of Con2PrimFactory/src/cons.hxx:49-91, and builds
the plain con2prim seed (saved prims) as the true primitives times
(1 + s*u), u ~ U[-1, 1] per component (s = seed_rel, default 1e-2), or a
cold start (v = 0, eps = tau/D) with cold=True.

Layouts match asx_c2p_batch: cons[8][nz] = dens, momx, momy, momz, tau,
dBx, dBy, dBz; g[6][nz]; seed[5][nz] = rho, eps, velx, vely, velz.
"""
import numpy as np


def _lower(g, v):
    gxx, gxy, gxz, gyy, gyz, gzz = g
    return np.stack([gxx * v[0] + gxy * v[1] + gxz * v[2],
                     gxy * v[0] + gyy * v[1] + gyz * v[2],
                     gxz * v[0] + gyz * v[1] + gzz * v[2]])


def make_batch(nz, seed=0, gamma=5.0 / 3.0, seed_rel=1e-2, cold=False):
    rng = np.random.default_rng(seed)
    lu = lambda a, b: np.exp(rng.uniform(np.log(a), np.log(b), nz))
    rho = lu(1e-10, 1e-2)
    press = rho * lu(1e-6, 1e2)
    eps = press / ((gamma - 1.0) * rho)
    W = rng.uniform(1.0, 10.0, nz)
    M = rng.uniform(-1.0, 1.0, (6, nz)) * (0.1 / 3.0)
    g = np.stack([1 + M[0], M[1], M[2], 1 + M[3], M[4], 1 + M[5]])

    def unit(vec):
        n = np.sqrt(np.einsum("in,in->n", vec, _lower(g, vec)))
        return vec / n
    nv = unit(rng.normal(size=(3, nz)))
    vabs = np.sqrt(1.0 - 1.0 / (W * W))
    vel = nv * vabs
    nb = unit(rng.normal(size=(3, nz)))
    b2 = rho * lu(1e-4, 1e2)
    vl = _lower(g, vel)
    ndotv = np.einsum("in,in->n", nb, vl)
    beta = np.sqrt(b2 / (1.0 / (W * W) + ndotv * ndotv))
    B = nb * beta
    det = 2 * g[1] * g[2] * g[4] + g[0] * g[3] * g[5] - g[3] * g[2] ** 2 - \
        g[0] * g[4] ** 2 - g[5] * g[1] ** 2
    sq = np.sqrt(det)
    Bl = _lower(g, B)
    bst = W * np.einsum("in,in->n", B, vl)
    bl = Bl / W + bst * vl
    bs2 = (np.einsum("in,in->n", B, Bl) + bst * bst) / (W * W)
    dens = sq * rho * W
    hW2 = W * W * (rho * (1.0 + eps) + press + bs2)
    mom = sq * (hW2 * vl - bst * bl)
    tau = sq * (hW2 - (press + 0.5 * bs2) - bst * bst) - dens
    cons = np.ascontiguousarray(np.concatenate(
        [dens[None], mom, tau[None], sq * B]))
    if cold:
        sd = np.stack([dens / sq, tau / dens, 0 * rho, 0 * rho, 0 * rho])
    else:
        u = lambda: 1.0 + seed_rel * rng.uniform(-1.0, 1.0, nz)
        sd = np.stack([rho * u(), eps * u(), vel[0] * u(), vel[1] * u(),
                       vel[2] * u()])
    true = dict(rho=rho, eps=eps, press=press, vel=vel, W=W, B=B, b2=b2)
    return cons, np.ascontiguousarray(g), np.ascontiguousarray(sd), true


def make_batch_ye(nz, seed, rho_range, temp_range, ye_range, W_max=10.0):
    """Zones with composition for a 3-parameter EOS: rho log-uniform in
    rho_range, T log-uniform in temp_range, Ye uniform in ye_range, W
    uniform in [1, W_max] along a random direction, b^2/rho log-uniform in
    [1e-4, 1e2] along a random direction, SPD metric within 10% of the
    identity.  Returns prim[9][nz] = rho, T, Ye, v^i, B^i (the layout of
    asx_prim2con_batch_ye) and g[6][nz]."""
    rng = np.random.default_rng(seed)
    lu = lambda a, b: np.exp(rng.uniform(np.log(a), np.log(b), nz))
    rho = lu(*rho_range)
    temp = lu(*temp_range)
    ye = rng.uniform(ye_range[0], ye_range[1], nz)
    W = rng.uniform(1.0, W_max, nz)
    M = rng.uniform(-1.0, 1.0, (6, nz)) * (0.1 / 3.0)
    g = np.stack([1 + M[0], M[1], M[2], 1 + M[3], M[4], 1 + M[5]])

    def unit(vec):
        n = np.sqrt(np.einsum("in,in->n", vec, _lower(g, vec)))
        return vec / n
    vel = unit(rng.normal(size=(3, nz))) * np.sqrt(1.0 - 1.0 / (W * W))
    nb = unit(rng.normal(size=(3, nz)))
    b2 = rho * lu(1e-4, 1e2)
    ndotv = np.einsum("in,in->n", nb, _lower(g, vel))
    B = nb * np.sqrt(b2 / (1.0 / (W * W) + ndotv * ndotv))
    prim = np.ascontiguousarray(np.concatenate([rho[None], temp[None], ye[None], vel, B]))
    return prim, np.ascontiguousarray(g)
