"""Single-patch time integration around the scheduled functions.

This is the part of the reference's run that CarpetX supplies (ODESolvers
RK update, SYNC + physical boundary conditions), arranged the way
AsterX/schedule.ccl wires the thorn's functions:

  initial   AsterX_InitialGroup   (schedule.ccl:39-100)
  rhs       AsterX_RHSGroup       (schedule.ccl:169-228): Fluxes, SourceTerms,
            RHS, each followed by the SYNC of what it wrote
  poststep  ODESolvers_PostStep   (schedule.ccl:110-165): Sync of the state,
            ComputedBstagFromA, ComputedBFromdBstag, Con2Prim

A SYNC on a single patch is the physical boundary fill of every listed grid
function.  The backend is any object with the asx_* entry points (the HIP
library, or the CPU oracle in tests).
"""
import numpy as np

from . import abi
from .patch import StateBuf, ptr_array

SYNC_FLUXES = abi.FLUX + abi.AUX
SYNC_UCT = abi.UCT
SYNC_RHS_HYDRO = ["densrhs", "momxrhs", "momyrhs", "momzrhs", "taurhs",
                  "DEntrhs", "DYe_rhs"]
SYNC_RHS_EM = ["Avec_x_rhs", "Avec_y_rhs", "Avec_z_rhs", "Psi_rhs"]
SYNC_PRIMS = ["rho", "velx", "vely", "velz", "eps", "press", "Bvecx", "Bvecy",
              "Bvecz", "temperature", "entropy", "Ye"]
SYNC_C2P = ["con2prim_flag"] + SYNC_PRIMS + abi.SAVED + abi.ZS + \
    ["dens", "tau", "DEnt", "momx", "momy", "momz", "DYe", "dBx", "dBy", "dBz"]


def all_slots(use_uct=False):
    s = abi.ADM + abi.PRIMS + abi.CONS + abi.DB + abi.SAVED + abi.ZS + \
        ["con2prim_flag", "aster_mask_cc"] + abi.DBSTAG + abi.AVEC + \
        abi.FLUX + abi.AUX + abi.RHS
    if use_uct:
        s = s + abi.UCT
    return s


class Evolution:
    def __init__(self, backend, patch, par, bc="linear", tmunu=False, fused=False):
        """tmunu: also set the stress-energy tensor after every con2prim
        (TmunuBaseX's SetTmunuVars, which runs after AsterX_Con2PrimGroup in
        ODESolvers_PostStep, AsterX/schedule.ccl:119), as a spacetime
        coupled to the fluid needs.  fused: the RHS group is one call of
        substep_bc (the fused substep with the face data SYNCed by the
        boundary fill) instead of Fluxes, SourceTerms and RHS."""
        self.b = backend
        self.p = patch
        self.par = par
        self.bc = abi.BC[bc] if isinstance(bc, str) else bc
        self.uct = bool(par.use_uct)
        self.with_tmunu = tmunu
        self.fused = fused

    def sync(self, names):
        ps = self.p.struct()
        self.b.bc_fill_many(ps, [abi.SLOT[n] for n in names if n in self.p],
                            self.bc)

    # AsterX_InitialGroup (check_prims is given: it is initial-data only)
    def initial(self, check_prims):
        p, b = self.p, self.b
        b.bstag_from_A(p.struct())
        self.sync(abi.DBSTAG)
        b.dB_from_bstag(p.struct())
        self.sync(abi.DB)
        b.B_from_dB(p.struct())
        self.sync(["Bvecx", "Bvecy", "Bvecz"])
        check_prims(p, self.par)
        self.sync(SYNC_PRIMS + abi.SAVED + abi.ZS)
        b.prim2con(p.struct())
        self.sync(["dens", "tau", "DEnt", "momx", "momy", "momz", "DYe"] + abi.DB)
        p["Psi"] = 0.0

    def rhs(self):
        p, b, par = self.p, self.b, self.par
        if self.fused:
            # out = 0 u0 + 0 u + 1 rhs(u) into the RHS grid functions
            b.substep_bc(p.struct(), None, p.state(abi.STATE_RHS), 0.0, 0.0, 1.0, par,
                         self.bc)
            return
        b.fluxes(p.struct(), par, 0)
        self.sync(SYNC_FLUXES + (SYNC_UCT if self.uct else []))
        b.source_terms(p.struct(), par)
        self.sync(SYNC_RHS_HYDRO)
        b.rhs(p.struct(), par)
        self.sync(SYNC_RHS_HYDRO + SYNC_RHS_EM)

    def poststep(self):
        p, b = self.p, self.b
        self.sync(abi.STATE)
        b.bstag_from_A(p.struct())
        self.sync(abi.DBSTAG)
        b.dB_from_bstag(p.struct())
        self.sync(abi.DB)
        b.con2prim(p.struct(), self.par)
        self.sync(SYNC_C2P)
        if self.with_tmunu:
            self.tmunu()

    # TmunuBaseX zeroes eT*, the TmunuBaseX_AddToTmunu group adds into it
    # (AsterX_Tmunu when update_tmunu, AsterX/schedule.ccl:253-266), then
    # eT* is SYNCed
    def tmunu(self):
        p = self.p
        for n in abi.TMUNU:
            if n not in p:
                p.alloc(n)
            p[n] = 0.0
        self.b.tmunu(p.struct(), self.par)
        self.sync(abi.TMUNU)

    # ODESolvers "RK4"
    def rk4_step(self, dt):
        p, b = self.p, self.b
        ps = p.struct()
        state = p.state(abi.STATE)
        rhs = p.state(abi.STATE_RHS)
        old = StateBuf(p)
        ks = [StateBuf(p) for _ in range(4)]
        olds = old.struct()
        kss = [k.struct() for k in ks]
        b.state_lincomb(ps, olds, 0.0, [1.0], [state])
        for stage, w in enumerate([0.5 * dt, 0.5 * dt, dt, None]):
            self.rhs()
            b.state_lincomb(ps, kss[stage], 0.0, [1.0], [rhs])
            if w is not None:
                b.state_lincomb(ps, state, 0.0, [1.0, w], [olds, kss[stage]])
            else:
                b.state_lincomb(ps, state, 0.0,
                                [1.0, dt / 6, dt / 3, dt / 3, dt / 6],
                                [olds] + kss)
            self.poststep()

    # ODESolvers "RK2" (midpoint)
    def rk2_step(self, dt):
        p, b = self.p, self.b
        ps = p.struct()
        state = p.state(abi.STATE)
        rhs = p.state(abi.STATE_RHS)
        old = StateBuf(p)
        k1 = StateBuf(p)
        olds, k1s = old.struct(), k1.struct()
        b.state_lincomb(ps, olds, 0.0, [1.0], [state])
        self.rhs()
        b.state_lincomb(ps, k1s, 0.0, [1.0], [rhs])
        b.state_lincomb(ps, state, 0.0, [1.0, 0.5 * dt], [olds, k1s])
        self.poststep()
        self.rhs()
        b.state_lincomb(ps, state, 0.0, [1.0, dt], [olds, rhs])
        self.poststep()


class Backend:
    """Thin Python wrapper over a ctypes library exposing the asx_* ABI
    (prefix 'asx_' for the HIP library, 'orc_' for the CPU oracle)."""

    def __init__(self, lib, prefix, stream=None):
        self.lib = lib
        self.prefix = prefix
        self.stream = stream

    def _f(self, name):
        return getattr(self.lib, self.prefix + name)

    def _call(self, name, *args):
        abi.check(self.prefix + name, self._f(name)(*args))

    def _s(self):
        return [self.stream] if self.prefix == "asx_" else []

    def fluxes(self, ps, par, halo=0):
        if self.prefix == "asx_":
            self._call("fluxes", ps_ref(ps), ref(par), halo, None, self.stream)
        else:
            self._call("fluxes", ps_ref(ps), ref(par), halo)

    def source_terms(self, ps, par):
        self._call("source_terms", ps_ref(ps), ref(par), *self._s())

    def rhs(self, ps, par):
        self._call("rhs", ps_ref(ps), ref(par), *self._s())

    def substep_bc(self, ps, u0, out, c0, c1, c2, par, kind):
        """out = c0 u0 + c1 u + c2 rhs(u), face data SYNCed by the boundary
        fill `kind` (asx_substep_bc / orc_substep_bc)."""
        u0r = ref(u0) if u0 is not None else None
        self._call("substep_bc", ps_ref(ps), u0r, ref(out), c0, c1, c2, ref(par), kind,
                   *self._s())

    def tmunu(self, ps, par):
        """eT_{mu nu} += T_{mu nu} at the interior vertices (AsterX_Tmunu)."""
        self._call("tmunu", ps_ref(ps), ref(par), *self._s())

    def bstag_from_A(self, ps):
        self._call("bstag_from_A", ps_ref(ps), *self._s())

    def dB_from_bstag(self, ps):
        self._call("dB_from_bstag", ps_ref(ps), *self._s())

    def B_from_dB(self, ps):
        self._call("B_from_dB", ps_ref(ps), *self._s())

    def prim2con(self, ps):
        self._call("prim2con", ps_ref(ps), *self._s())

    def con2prim(self, ps, par, cnt=None):
        if self.prefix == "asx_":
            self._call("con2prim", ps_ref(ps), ref(par), cnt, self.stream)
        else:
            self._call("con2prim", ps_ref(ps), ref(par))

    def state_lincomb(self, ps, dst, scale, fac, srcs):
        import ctypes as C
        facs = (C.c_double * max(1, len(fac)))(*fac)
        arr = ptr_array(srcs)
        self._call("state_lincomb", ps_ref(ps), ref(dst), float(scale),
                   len(fac), facs, arr, *self._s())

    def bc_fill(self, ps, slot, kind):
        self._call("bc_fill", ps_ref(ps), slot, kind, *self._s())

    def bc_fill_many(self, ps, slots, kind):
        if self.prefix != "asx_":  # the oracle fills one grid function at a time
            for s in slots:
                self.bc_fill(ps, s, kind)
            return
        import ctypes as C
        for i in range(0, len(slots), 64):
            chunk = slots[i:i + 64]
            arr = (C.c_int32 * len(chunk))(*chunk)
            self._call("bc_fill_many", ps_ref(ps), len(chunk), arr, kind,
                       *self._s())


def ref(x):
    import ctypes as C
    return C.byref(x)


def ps_ref(ps):
    import ctypes as C
    return C.byref(ps)


def check_prims_ideal(patch, par):
    """AsterX_CheckPrims (AsterX/src/check_prims.cxx:24-228) on interior
    cells, ideal gas, Minkowski-compatible v2c metric averages, numpy.
    Initial data only (not on the measured path)."""
    sl = patch.interior("rho")
    g = {}
    for c in ("gxx", "gxy", "gxz", "gyy", "gyz", "gzz"):
        a = patch.numpy(c)
        s = np.zeros(patch.shape("rho"))
        for dk in (0, 1):
            for dj in (0, 1):
                for di in (0, 1):
                    s = s + a[dk:dk + s.shape[0], dj:dj + s.shape[1],
                              di:di + s.shape[2]]
        g[c] = (s / 8.0)[sl]
    gam = par.gl_gamma
    rho = patch.numpy("rho")[sl].copy()
    eps = patch.numpy("eps")[sl].copy()
    press = patch.numpy("press")[sl].copy()
    Ye = patch.numpy("Ye")[sl].copy()
    temp = patch.numpy("temperature")[sl].copy()
    v = [patch.numpy(n)[sl].copy() for n in ("velx", "vely", "velz")]
    ent = (gam - 1.0) * eps * np.power(rho, 1.0 - gam)
    gm1 = gam - 1.0
    eps_max = par.eps_max
    if gam > 2:
        eps_max = min(eps_max, 1 / (gam * (gam - 2)))
    # atmosphere (r <= r_atmo assumed: ungraded)
    rho_atm = max(par.rho_min, par.rho_abs_min)
    t_atm = max(gm1 * par.particle_mass * par.eps_min, par.t_atmo)
    eps_atm = t_atm / (gm1 * par.particle_mass)
    press_atm = gm1 * rho_atm * eps_atm
    if par.use_press_atmo:
        press_atm = max(gm1 * rho_atm * par.eps_min, par.p_atmo)
        eps_atm = press_atm / (rho_atm * gm1)
    rho_cut = rho_atm * (1 + par.atmo_tol)
    Ye = np.clip(Ye, par.ye_min, par.ye_max)

    def low(vv):
        return [g["gxx"] * vv[0] + g["gxy"] * vv[1] + g["gxz"] * vv[2],
                g["gxy"] * vv[0] + g["gyy"] * vv[1] + g["gyz"] * vv[2],
                g["gxz"] * vv[0] + g["gyz"] * vv[1] + g["gzz"] * vv[2]]

    vl = low(v)
    w_lim = np.sqrt(1.0 + par.vw_lim ** 2)
    v_lim = par.vw_lim / w_lim
    sol_v = np.sqrt(vl[0] * v[0] + vl[1] * v[1] + vl[2] * v[2])
    fast = sol_v > v_lim
    if fast.any():
        s = np.where(fast, v_lim / np.where(fast, sol_v, 1.0), 1.0)
        v = [vi * s for vi in v]
        vl = [vi * s for vi in vl]
    hi = rho > par.rho_max
    rho = np.where(hi, par.rho_max, rho)
    eps = np.where(hi, press / (rho * gm1), eps)
    lo = rho < rho_cut
    rho = np.where(lo, rho_atm, rho)
    eps = np.where(lo, press / (rho * gm1), eps)
    ent = np.where(hi | lo, gm1 * eps * np.power(rho, 1.0 - gam), ent)
    emax = eps > eps_max
    eps = np.where(emax, eps_max, eps)
    temp = np.where(emax, gm1 * par.particle_mass * eps, temp)
    press = np.where(emax, gm1 * rho * eps, press)
    if par.use_press_atmo:
        low_p = press < press_atm
        press = np.where(low_p, press_atm, press)
        eps = np.where(low_p, press / (rho * gm1), eps)
        fix = low_p | emax
    else:
        low_e = eps < eps_atm
        eps = np.where(low_e, eps_atm, eps)
        press = np.where(low_e, gm1 * rho * eps, press)
        fix = low_e | emax
    temp = np.where(fix, gm1 * par.particle_mass * eps, temp)
    ent = np.where(fix, gm1 * eps * np.power(rho, 1.0 - gam), ent)
    out = {"rho": rho, "eps": eps, "press": press, "entropy": ent, "Ye": Ye,
           "temperature": temp, "velx": v[0], "vely": v[1], "velz": v[2],
           "saved_rho": rho, "saved_velx": v[0], "saved_vely": v[1],
           "saved_velz": v[2], "saved_eps": eps, "saved_Ye": Ye}
    wlor = 1.0 / np.sqrt(1.0 - (vl[0] * v[0] + vl[1] * v[1] + vl[2] * v[2]))
    for i, c in enumerate("xyz"):
        out["zvec_" + c] = wlor * v[i]
        out["svec_" + c] = (rho + rho * eps + press) * wlor * wlor * v[i]
    for name, val in out.items():
        if name not in patch:
            patch.alloc(name)
        full = patch.numpy(name).copy()
        full[sl] = val
        patch[name] = full
