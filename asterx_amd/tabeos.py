"""Synthetic tabulated EOS and its descriptor for the C ABI.

The table is 234 x 180 x 50 nodes in log10 rho [g/cc] in [3, 15.5], log10 T
[MeV] in [-2, 2.4] and Ye in [0.035, 0.55].  Per node, in CGS:

  P   = S(Ye) P_c(rho) + rho k T / m_u
  eps = S(Ye) eps_c(rho) + (3/2) k T / m_u
  S(Ye) = 1 + (Ye - 1/2)^2 / 2                 (smooth Ye dependence)

with a piecewise polytrope P_c = K_i rho^Gamma_i (Gamma 1.36 below
rho_t = 5e14 g/cc, 3.0 above, continuous), eps_c its cold energy
(d eps_c = P_c / rho^2 d rho, eps_c(0) = 0), an ideal thermal part, and

  c_s^2 = [dP/drho|_eps + (P/rho^2) dP/deps|_rho] / h,
  h = 1 + (eps + P/rho)/c^2,
  s = 45 + (3/2) ln(T / MeV) - ln(rho / (g/cc))   (entropy per baryon of
                                                   the thermal gas, > 0)

Stored like a stellar-collapse table: log10 P, log10(eps + energy_shift)
(shift 2e19 erg/g), c_s^2, and the entropy, which AsterX's fluxes read
(their low-order fallback triggers on a reconstructed entropy <= 0) and
its con2prim writes.  `code_units` converts it the way EOSX reads a
table (EOSX/src/eos_3p_tabulated3d/eos_3p_tabulated3d.hxx:151-165, the
constants of EOSX/src/utils/eos_constants.hxx:7-12): natural logs, code
units, variable index fastest.
"""
import ctypes as C
import math
import os

import numpy as np

from . import abi

# EOSX/src/utils/eos_constants.hxx:7-12
LENGTHGF = 6.77269222552442e-06
TIMEGF = 2.03040204956746e05
RHOGF = 1.61887093132742e-18
PRESSGF = 1.80123683248503e-39
EPSGF = 1.11265005605362e-21
C2_CGS = 8.9875517873681764e+20
MEV_TO_ERG = 1.60217733e-6
MNUC_CGS = 1.660539040e-24

NRHO, NTEMP, NYE = 234, 180, 50
LOGRHO = (3.0, 15.5)
LOGTEMP = (-2.0, 2.4)
YE = (0.035, 0.55)
RHO_T = 5e14
GAMMA1, GAMMA2 = 1.36, 3.0
K1 = 1e13
SHIFT_CGS = 2e19


def cold(rho):
    """P_c, eps_c, and the adiabatic index Gamma_i of the cold polytrope (CGS)."""
    rho = np.asarray(rho, dtype=np.float64)
    K2 = K1 * RHO_T ** (GAMMA1 - GAMMA2)
    eps_t = K1 * RHO_T ** (GAMMA1 - 1) / (GAMMA1 - 1)
    lo = rho <= RHO_T
    P = np.where(lo, K1 * rho ** GAMMA1, K2 * rho ** GAMMA2)
    eps = np.where(lo, K1 * rho ** (GAMMA1 - 1) / (GAMMA1 - 1),
                   eps_t + K2 * (rho ** (GAMMA2 - 1) - RHO_T ** (GAMMA2 - 1)) / (GAMMA2 - 1))
    G = np.where(lo, GAMMA1, GAMMA2)
    return P, eps, G


def ye_factor(ye):
    return 1.0 + 0.5 * (np.asarray(ye) - 0.5) ** 2


def thermal_eps(temp):
    """(3/2) k T / m_u in erg/g, T in MeV."""
    return 1.5 * np.asarray(temp) * MEV_TO_ERG / MNUC_CGS


def entropy(rho, temp):
    """Entropy per baryon (k_B) of the thermal gas, offset to stay positive
    on the table."""
    return 45.0 + 1.5 * np.log(np.asarray(temp)) - np.log(np.asarray(rho))


def state_cgs(rho, temp, ye):
    """P, eps, c_s^2 (CGS) at (rho [g/cc], T [MeV], Ye)."""
    Pc, ec, G = cold(rho)
    S = ye_factor(ye)
    eth = thermal_eps(temp)
    P = S * Pc + (2.0 / 3.0) * rho * eth
    eps = S * ec + eth
    dPdrho = S * G * Pc / rho + (2.0 / 3.0) * eth - (2.0 / 3.0) * S * Pc / rho
    dPdeps = (2.0 / 3.0) * rho
    h = 1.0 + (eps + P / rho) / C2_CGS
    cs2 = (dPdrho + P / rho ** 2 * dPdeps) / h
    return P, eps, cs2


def make_cgs(nrho=NRHO, ntemp=NTEMP, nye=NYE):
    """Axes (log10 rho, log10 T, Ye) and the CGS tables log10 P,
    log10(eps + shift), c_s^2 and entropy, each [nye][ntemp][nrho]."""
    lr = np.linspace(LOGRHO[0], LOGRHO[1], nrho)
    lt = np.linspace(LOGTEMP[0], LOGTEMP[1], ntemp)
    ye = np.linspace(YE[0], YE[1], nye)
    Y, T, R = np.meshgrid(ye, 10.0 ** lt, 10.0 ** lr, indexing="ij")
    P, eps, cs2 = state_cgs(R, T, Y)
    return lr, lt, ye, np.log10(P), np.log10(eps + SHIFT_CGS), cs2, entropy(R, T)


def code_units(lr10, lt10, ye, logp10, loge10, cs2, ent, shift_cgs=SHIFT_CGS):
    """The table as EOSX holds it after reading (eos_3p_tabulated3d.hxx:
    151-165): natural-log axes in code units, log p and log(eps + shift)
    in code units, c_s^2 in code units, the entropy as it is, interleaved
    [.., 4] with the variable fastest; plus the energy shift and the eps
    range."""
    ln10 = math.log(10.0)
    shift = shift_cgs * EPSGF
    ax0 = lr10 * ln10 + math.log(RHOGF)
    ax1 = lt10 * ln10
    ax2 = np.array(ye, dtype=np.float64)
    inv_time2 = 1.0 / (TIMEGF * TIMEGF)
    lp = logp10 * ln10 + math.log(PRESSGF)
    le = loge10 * ln10 + math.log(EPSGF)
    c2 = cs2 * LENGTHGF * LENGTHGF * inv_time2
    data = np.stack([lp, le, c2, np.asarray(ent, dtype=np.float64)], axis=-1)
    eps = np.exp(le) - shift
    # EOSX's rgeps: min / max of eps over the nodes, the minimum floored
    # (eos_3p_tabulated3d.hxx:341-357)
    eps_min = max(float(eps.min()), 1e-15)
    eps_max = float(eps.max())
    return dict(data=np.ascontiguousarray(data), axes=(ax0, ax1, ax2), shift=shift,
                eps_min=eps_min, eps_max=eps_max)


def build(nrho=NRHO, ntemp=NTEMP, nye=NYE):
    """The synthetic table in code units (host arrays)."""
    return code_units(*make_cgs(nrho, ntemp, nye))


def read_stellarcollapse(path):
    """A StellarCollapse HDF5 table read by the library's reader
    (asx_eos_read_stellarcollapse: EOSX's reader and unit conversion), in
    the form build() returns."""
    from . import lib
    L = lib.eosio()
    t = lib.asx_eos_host_table()
    st = L.asx_eos_read_stellarcollapse(os.fsencode(path), C.byref(t))
    if st:
        raise RuntimeError("asx_eos_read_stellarcollapse(%s): %d" % (path, st))
    try:
        n = [t.n[k] for k in range(3)]
        data = np.ctypeslib.as_array(t.data, shape=(n[2], n[1], n[0], 4)).copy()
        axes = tuple(np.ctypeslib.as_array(t.axis[k], shape=(n[k],)).copy() for k in range(3))
        return dict(data=data, axes=axes, shift=t.energy_shift, eps_min=t.eps_min,
                    eps_max=t.eps_max)
    finally:
        L.asx_eos_host_table_free(C.byref(t))


class Table:
    """A table in device (torch) or host (numpy) memory and its
    asx_eos_table descriptor.  Keeps the arrays alive."""

    def __init__(self, tab, device="cuda"):
        self.host = tab
        if device is None:
            self.data = tab["data"]
            self.axes = [np.ascontiguousarray(a) for a in tab["axes"]]
            ptr = lambda a: a.ctypes.data
        else:
            import torch
            self.data = torch.from_numpy(tab["data"]).to(device)
            self.axes = [torch.from_numpy(np.ascontiguousarray(a)).to(device)
                         for a in tab["axes"]]
            ptr = lambda a: a.data_ptr()
        d = abi.asx_eos_table()
        d.data = ptr(self.data)
        for k in range(3):
            d.axis[k] = ptr(self.axes[k])
            d.n[k] = len(tab["axes"][k])
        d.nvars = 4
        d.energy_shift = tab["shift"]
        d.eps_min = tab["eps_min"]
        d.eps_max = tab["eps_max"]
        for k in range(3):
            a = tab["axes"][k]
            d.x_lo[k] = float(a[0])
            d.x_step[k] = float(a[1] - a[0])
            d.x_hi[k] = float(a[-1])
        self.desc = d

    def attach(self, par):
        """Point asx_params at this table (evolution_eos = Tabulated3d)."""
        par.evolution_eos = abi.EOS["Tabulated3d"]
        par.eos_table = C.addressof(self.desc)
        par._table = self  # the table's arrays live as long as the parameters
        return par

    def ranges(self):
        """(rho_min, rho_max), (T_min, T_max), (Ye_min, Ye_max) in code units."""
        a = self.host["axes"]
        return ((math.exp(a[0][0]), math.exp(a[0][-1])),
                (math.exp(a[1][0]), math.exp(a[1][-1])), (a[2][0], a[2][-1]))
