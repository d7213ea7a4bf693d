"""Regression cases of the reference (AsterX/test/*.par) as runnable setups.

Each entry restates the parameters of one parfile; run_case() evolves it
with the single-patch driver on a backend (CPU oracle or HIP library) and
returns the primitive fields at every iteration the golden TSV files hold,
on every cut direction the golden holds.
"""
import math
import os

import numpy as np

from asterx_amd import abi, seeds
from asterx_amd.driver import Evolution, all_slots, check_prims_ideal
from asterx_amd.patch import Patch

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

_B1 = dict(kind="balsara", test="Balsara1", shock_dir="x", ncells=50, ng=3,
           bc="copy", recon="ppm", flux="HLLE", prime="Noble",
           second="Palenzuela", tol=1e-8, max_iter=100, rho_abs_min=1e-8,
           gamma=2.0, t_final=0.40, rk="rk4", dtfac=0.25, dirs=None)

CASES = {
    # AsterX/test/Balsara1_shocktube_xdir_ppm.par
    "Balsara1_shocktube_xdir_ppm": dict(_B1, bc="linear"),
    # AsterX/test/Balsara1_shocktube_xdir_wenoz.par
    "Balsara1_shocktube_xdir_wenoz": dict(_B1, recon="wenoz"),
    # AsterX/test/Balsara1_shocktube_xdir_mp5.par
    "Balsara1_shocktube_xdir_mp5": dict(_B1, recon="mp5"),
    # AsterX/test/Balsara1_shocktube_xdir_godunov.par
    "Balsara1_shocktube_xdir_godunov": dict(_B1, recon="Godunov", ng=2),
    # AsterX/test/Balsara1_shocktube_xdir_monocentral.par
    "Balsara1_shocktube_xdir_monocentral": dict(_B1, recon="monocentral", ng=2),
    # AsterX/test/Balsara1_shocktube_xdir_ppm_PalenzuelaC2P.par
    "Balsara1_shocktube_xdir_ppm_PalenzuelaC2P": dict(
        _B1, prime="Palenzuela", second="None", tol=1e-14),
    # AsterX/test/Balsara1_shocktube_ydir_ppm.par
    "Balsara1_shocktube_ydir_ppm": dict(_B1, shock_dir="y"),
    # AsterX/test/Balsara1_shocktube_zdir_ppm.par
    "Balsara1_shocktube_zdir_ppm": dict(_B1, shock_dir="z"),
    # AsterX/test/Balsara2_shocktube_xdir_ppm.par
    "Balsara2_shocktube_xdir_ppm": dict(_B1, test="Balsara2",
                                        gamma=1.66666666666667),
    # AsterX/test/Balsara3_shocktube_xdir_minmod_LxF.par
    "Balsara3_shocktube_xdir_minmod_LxF": dict(
        _B1, test="Balsara3", ncells=80, ng=2, recon="minmod", flux="LxF",
        rho_abs_min=1e-6, gamma=1.66666666666667),
    # AsterX/test/Balsara4_shocktube_xdir_ppm.par
    "Balsara4_shocktube_xdir_ppm": dict(_B1, test="Balsara4",
                                        gamma=1.66666666666667),
    # AsterX/test/Balsara5_shocktube_xdir_ppm.par
    "Balsara5_shocktube_xdir_ppm": dict(_B1, test="Balsara5",
                                        gamma=1.66666666666667, t_final=0.55),
    # AsterX/test/Alfven_wave.par: periodic, minmod, algebraic gauge, RK2
    "Alfven_wave": dict(
        _B1, kind="alfven", n=(17, 3, 3), lo=(-0.53125,) * 3, hi=(0.53125,) * 3,
        ng=2, bc="periodic", recon="minmod", gauge="algebraic", rk="rk2",
        dtfac=0.2, rho_abs_min=1e-5, atmo_tol=1e-1, gamma=1.6666666,
        eps_min=1e-8, dirs="xyz"),
    # AsterX/test/Sound_wave.par: 3-D PPM, linear extrapolation, RK4
    "Sound_wave": dict(
        _B1, kind="sound", n=(10, 10, 10), lo=(-0.5,) * 3, hi=(0.5,) * 3,
        ng=3, bc="linear", dirs="xyz"),
    # AsterX/test/Magnetic_rotor.par: 2-D PPM, neumann, RK4
    "Magnetic_rotor": dict(
        _B1, kind="rotor", n=(40, 40, 2), lo=(-0.5, -0.5, -0.004),
        hi=(0.5, 0.5, 0.004), ng=3, bc="copy", dtfac=0.35,
        gamma=1.66666666666667, eps_min=1e-6, dirs="xyz"),
    # AsterX/test/Cylindrical_blast.par: 2-D minmod, upwind CT, z_vec,
    # entropy fix, neumann, RK4
    "Cylindrical_blast": dict(
        _B1, kind="blast", n=(50, 50, 2), lo=(-6.0, -6.0, -0.06),
        hi=(6.0, 6.0, 0.06), ng=2, bc="copy", recon="minmod", dtfac=0.35,
        max_iter=1000, rho_abs_min=1e-6, gamma=1.3333333333, eps_min=1e-7,
        uct=1, recon_type="z_vec", entropy_fix=1, dirs="xyz"),
}


def case_params(c):
    kw = {}
    if "gauge" in c:
        kw["vector_potential_gauge"] = c["gauge"]
    if "atmo_tol" in c:
        kw["atmo_tol"] = c["atmo_tol"]
    if "eps_min" in c:
        kw["eps_min"] = c["eps_min"]
    for k, name in (("uct", "use_uct"), ("recon_type", "recon_type"),
                    ("entropy_fix", "use_entropy_fix")):
        if k in c:
            kw[name] = c[k]
    return abi.params(
        reconstruction_method=c["recon"], flux_type=c["flux"],
        c2p_prime=c["prime"], c2p_second=c["second"], c2p_tol=c["tol"],
        max_iter=c["max_iter"], rho_abs_min=c["rho_abs_min"], B_lim=1e8,
        vw_lim=1e8, Ye_lenient=1, gl_gamma=c["gamma"], poly_gamma=c["gamma"],
        rho_max=1e8, eps_max=1e8, **kw)


def cut_dirs(c):
    """Golden cut directions compared for a case."""
    return c["dirs"] or c["shock_dir"]


def make_patch(c, device=None):
    if c["kind"] == "balsara":
        d = "xyz".index(c["shock_dir"])
        n = [2, 2, 2]
        n[d] = c["ncells"]
        lo, hi = (-0.5,) * 3, (0.5,) * 3
    else:
        n, lo, hi = list(c["n"]), c["lo"], c["hi"]
    dx = [(hi[q] - lo[q]) / n[q] for q in range(3)]
    p = Patch(n, c["ng"], lo, dx, slots=all_slots(use_uct=bool(c.get("uct"))), device=None)
    seeds.minkowski(p)
    if c["kind"] == "balsara":
        seeds.balsara(p, c["test"], c["shock_dir"], c["gamma"])
    elif c["kind"] == "alfven":
        seeds.alfven_wave(p, c["gamma"])
    elif c["kind"] == "sound":
        seeds.sound_wave(p, c["gamma"])
    elif c["kind"] == "blast":
        seeds.cylindrical_blast(p, c["gamma"])
    else:
        seeds.magnetic_rotor(p, c["gamma"])
    seeds.post_initial(p, c["gamma"])
    if device is not None:
        p = p.to(device)
    return p


def cut(p, d, idx):
    """Values of every TSV variable of patch p along the golden cut d."""
    ng = p.ng
    ii = idx[:, 0] + ng
    jj = idx[:, 1] + ng
    kk = idx[:, 2] + ng
    out = {}
    for var, names in (("rho", ["rho"]), ("vel", ["velx", "vely", "velz"]),
                       ("eps", ["eps"]), ("press", ["press"]),
                       ("bvec", ["Bvecx", "Bvecy", "Bvecz"])):
        out[var] = np.stack([p.numpy(n)[kk, jj, ii] for n in names], axis=-1)
    return out


def run_case(name, backend, device=None, iterations=None, fused=False):
    """Evolve a case; returns {iteration: {var: values on the golden cut}}.
    fused: the RHS group through the fused substep with SYNCed faces."""
    c = CASES[name]
    gold = np.load(os.path.join(GOLDEN, "tsv_%s.npz" % name))
    dirs = cut_dirs(c)
    its = list(gold["it_" + dirs[0]])
    p = make_patch(c, device)
    par = case_params(c)
    ev = Evolution(backend, p, par, bc=c["bc"], fused=fused)
    ev.initial(_check_prims_for(device))
    dt = c["dtfac"] * min(p.dx)
    step = ev.rk4_step if c["rk"] == "rk4" else ev.rk2_step
    nmax = its[-1] if iterations is None else iterations

    def cuts():
        return {d: cut(p, d, gold["idx_" + d]) for d in dirs}
    res = {}
    if 0 in its:
        res[0] = cuts()
    for it in range(1, nmax + 1):
        step(dt)
        if it in its:
            res[it] = cuts()
    return res, gold


def _check_prims_for(device):
    if device is None:
        return check_prims_ideal

    def cp(p, par):
        names = ["gxx", "gxy", "gxz", "gyy", "gyz", "gzz", "rho", "eps",
                 "press", "Ye", "temperature", "velx", "vely", "velz"]
        h = p.to(None, names)
        check_prims_ideal(h, par)
        for k, v in h.arr.items():
            if k not in ("gxx", "gxy", "gxz", "gyy", "gyz", "gzz"):
                p[k] = v
    return cp


def compare(res, gold, atol=1e-8, rtol=1e-8):
    """Max violation of |a-b| <= atol + rtol*|b| over every iteration, cut
    direction and variable."""
    worst = 0.0
    where = None
    for it, by_dir in res.items():
        for d, vals in by_dir.items():
            t = list(gold["it_" + d]).index(it)
            for var, a in vals.items():
                key = "%s_%s" % (var, d)
                if key not in gold.files:  # not written by the reference test
                    continue
                b = gold[key][t]
                err = np.abs(a - b) - (atol + rtol * np.abs(b))
                m = float(err.max())
                if m > worst:
                    worst = m
                    where = (it, d, var,
                             int(np.unravel_index(err.argmax(), err.shape)[0]))
    return worst, where
