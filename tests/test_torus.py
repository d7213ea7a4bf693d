"""Torus workload: the Kerr-Schild metric and the Fishbone-Moncrief disk are
checked against what the mathematics fixes, not against a retyped formula:

  * a = 0 reduces to Schwarzschild in Kerr-Schild form, whose closed forms
    gamma_ij = delta_ij + 2H l_i l_j, alpha^-2 = 1 + 2H, beta^i = 2H l^i/(1+2H)
    and K_ij = 2 H alpha / r [delta_ij - (2 + H) l_i l_j] (H = M/r) are
    textbook;
  * at a = 0.9375 the data are vacuum: the Hamiltonian and momentum
    constraints vanish (derivatives by finite differences of independently
    evaluated points);
  * the disk has u_phi u^t = l everywhere, u.u = -1, h = 1 at the inner edge
    and a pressure maximum (d_r h = 0) at r_max on the equator.
"""
import math

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from asterx_amd import torus  # noqa: E402

M, A = 1.0, 0.9375


def _pts(n=40, seed=0, rmin=2.0, rmax=30.0):
    rng = np.random.default_rng(seed)
    r = rng.uniform(rmin, rmax, n)
    cth = rng.uniform(-0.95, 0.95, n)
    ph = rng.uniform(0, 2 * np.pi, n)
    sth = np.sqrt(1 - cth ** 2)
    x, y, z = r * sth * np.cos(ph), r * sth * np.sin(ph), r * cth
    return [torch.from_numpy(v) for v in (x, y, z)]


def _sym(q, pre):
    m = torch.empty(q[pre + "xx"].shape + (3, 3), dtype=torch.float64)
    for i, a in enumerate("xyz"):
        for j, b in enumerate("xyz"):
            key = pre + (a + b if i <= j else b + a)
            m[..., i, j] = q[key]
    return m


def test_schwarzschild_closed_form():
    x, y, z = _pts()
    q = torus.ks_adm(x, y, z, M, 0.0)
    r = torch.sqrt(x * x + y * y + z * z)
    l = torch.stack([x / r, y / r, z / r], -1)
    H = M / r
    ll = l[..., :, None] * l[..., None, :]
    eye = torch.eye(3, dtype=torch.float64)
    g = eye + 2 * H[..., None, None] * ll
    alp = (1 + 2 * H) ** -0.5
    K = (2 * H * alp / r)[..., None, None] * (eye - (2 + H)[..., None, None] * ll)
    assert torch.allclose(_sym(q, "g"), g, rtol=1e-13, atol=1e-14)
    assert torch.allclose(q["alp"], alp, rtol=1e-14)
    beta = torch.stack([q["betax"], q["betay"], q["betaz"]], -1)
    assert torch.allclose(beta, (2 * H / (1 + 2 * H))[..., None] * l, atol=1e-14)
    assert torch.allclose(_sym(q, "k"), K, rtol=1e-10, atol=1e-13)


def test_shift_lowered_by_metric():
    x, y, z = _pts(seed=1)
    alp, bu, gam, bd = torus.ks_metric(x, y, z, M, A)
    g = torch.stack([torch.stack([gam[torus.PAIRS.index((min(i, j), max(i, j)))]
                                  for j in range(3)], -1) for i in range(3)], -2)
    b = torch.stack(bu, -1)
    low = (g @ b[..., None])[..., 0]
    assert torch.allclose(low, torch.stack(bd, -1), rtol=1e-13, atol=1e-14)
    # lapse from the 4-metric: alpha^2 = beta_i beta^i - g_tt
    r = torch.sqrt(x * x + y * y + z * z)
    Sig = r * r + A * A * (z / r) ** 2
    g_tt = -(1 - 2 * M * r / Sig)
    assert torch.allclose(alp ** 2, (low * b).sum(-1) - g_tt, rtol=1e-13)


def _fields(p):
    q = torus.ks_adm(p[0], p[1], p[2], M, A)
    g = _sym(q, "g")
    K = _sym(q, "k")
    return g, K


def _christoffel(p):
    """Gamma^k_ij at points p, by complex-step derivatives of gamma."""
    h = 1e-30
    dg = []
    for l in range(3):
        pc = [c.to(torch.complex128) for c in p]
        pc[l] = pc[l] + 1j * h
        _, _, gam, _ = torus.ks_metric(pc[0], pc[1], pc[2], M, A)
        dg.append(torch.stack([torch.stack([gam[torus.PAIRS.index(
            (min(i, j), max(i, j)))].imag / h for j in range(3)], -1)
            for i in range(3)], -2))
    dg = torch.stack(dg, -3)                     # [..., l, i, j] = d_l g_ij
    g, _ = _fields(p)
    gi = torch.linalg.inv(g)
    # Gamma_kij = 1/2 (d_i g_jk + d_j g_ik - d_k g_ij)
    low = 0.5 * (dg.permute(0, 3, 1, 2) + dg.permute(0, 3, 2, 1) - dg)
    # low[..., k, i, j]: permute so index order is (k, i, j)
    return torch.einsum("...kl,...lij->...kij", gi, low)


def _d(fn, p, m, h=2e-3):
    """4th-order central difference of fn along axis m."""
    def at(s):
        q = list(p)
        q[m] = q[m] + s * h
        return fn(q)
    return (-at(2) + 8 * at(1) - 8 * at(-1) + at(-2)) / (12 * h)


def test_vacuum_constraints():
    p = _pts(n=12, seed=2, rmin=3.0, rmax=20.0)
    g, K = _fields(p)
    gi = torch.linalg.inv(g)
    G = _christoffel(p)
    dG = torch.stack([_d(_christoffel, p, m) for m in range(3)], -4)
    # R_ij = d_k G^k_ij - d_j G^k_ik + G^k_kl G^l_ij - G^k_jl G^l_ik
    R = (torch.einsum("...kkij->...ij", dG)
         - torch.einsum("...jkik->...ij", dG)
         + torch.einsum("...kkl,...lij->...ij", G, G)
         - torch.einsum("...kjl,...lik->...ij", G, G))
    Rs = torch.einsum("...ij,...ij->...", gi, R)
    Ku = gi @ K @ gi
    trK = torch.einsum("...ij,...ij->...", gi, K)
    KK = torch.einsum("...ij,...ij->...", K, Ku)
    ham = Rs + trK ** 2 - KK
    scale = Rs.abs() + trK ** 2 + KK.abs()
    assert (ham.abs() / scale).max() < 1e-8
    # momentum: D_j K^j_i - d_i K = 0, K^j_i = g^jk K_ki
    def mixed(q):
        gg, KK_ = _fields(q)
        return torch.linalg.inv(gg) @ KK_

    def trace(q):
        gg, KK_ = _fields(q)
        return torch.einsum("...ij,...ij->...", torch.linalg.inv(gg), KK_)
    Km = gi @ K                                   # [.., j, i] = K^j_i
    dKm = torch.stack([_d(mixed, p, m) for m in range(3)], -3)  # [.., m, j, i]
    div = torch.einsum("...jji->...i", dKm) \
        + torch.einsum("...jjl,...li->...i", G, Km) \
        - torch.einsum("...lji,...jl->...i", G, Km)
    dtr = torch.stack([_d(trace, p, m) for m in range(3)], -1)
    mom = div - dtr
    mscale = dKm.abs().amax((-1, -2, -3)) + 1e-30
    assert (mom.abs().amax(-1) / mscale).max() < 1e-7


def test_fm_disk_invariants():
    spec = torus.TorusSpec()
    ell = spec.ell
    rng = np.random.default_rng(3)
    for _ in range(50):
        r = rng.uniform(7.0, 20.0)
        s2 = rng.uniform(0.6, 1.0)
        ut, uphi = torus.fm_four_velocity(M, A, ell, r, s2)
        Sig, Del, Abl = torus._bl(M, A, r, s2)
        g_tt = -(1 - 2 * M * r / Sig)
        g_tp = -2 * A * M * r * s2 / Sig
        g_pp = Abl * s2 / Sig
        u_phi = g_tp * ut + g_pp * uphi
        assert abs(u_phi * ut - ell) < 1e-12 * ell
        norm = g_tt * ut * ut + 2 * g_tp * ut * uphi + g_pp * uphi * uphi
        assert abs(norm + 1) < 1e-12
    # h = 1 on the inner edge, pressure maximum at r_max on the equator
    lnh = lambda r: torus.fm_lnh_raw(M, A, ell, r, 1.0) - spec.lnh_in
    assert abs(lnh(spec.r_in)) < 1e-15
    h = 1e-4
    d = (lnh(spec.r_max + h) - lnh(spec.r_max - h)) / (2 * h)
    assert abs(d) < 1e-9
    assert lnh(spec.r_max) > lnh(spec.r_max - 1) and \
        lnh(spec.r_max) > lnh(spec.r_max + 1)
    assert spec.rho_max > 1.0


def test_torus_patch_on_oracle():
    """The full initial group on a coarse patch through the CPU oracle:
    finite state, min beta = 100 in the disk, the disk present."""
    import oracle
    from asterx_amd.patch import Patch
    from asterx_amd import abi
    from asterx_amd.driver import all_slots
    n, ng, x0, dx = torus.torus_patch_geometry(n=24)
    p = Patch(n, ng, x0, dx, all_slots())
    spec = torus.TorusSpec()
    par = torus.eos_params(spec, reconstruction_method="wenoz")
    spec, ev, disk = torus.build_torus(p, oracle.backend(), par, spec)
    for name in abi.CONS + abi.PRIMS + abi.ADM + abi.AVEC:
        assert np.isfinite(p.numpy(name)).all(), name
    assert int(disk.sum()) > 20
    beta = torus.plasma_beta_min(p, disk)
    assert abs(beta / spec.beta_min - 1) < 1e-6
    # one substep runs and stays finite
    ev.rhs()
    for name in abi.RHS:
        assert np.isfinite(p.numpy(name)[p.interior(name)]).all(), name
