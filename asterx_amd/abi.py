"""ctypes mirror of include/asterx.h (the C ABI of the substep library).

The structs here must match the header field for field; tests compare
``ctypes.sizeof`` and the slot table against the compiled library.
"""
import ctypes as C

# ---- slot names, in the order of the enum in include/asterx.h -------------
ADM = ["alp", "betax", "betay", "betaz", "gxx", "gxy", "gxz", "gyy", "gyz",
       "gzz", "kxx", "kxy", "kxz", "kyy", "kyz", "kzz"]
PRIMS = ["rho", "velx", "vely", "velz", "eps", "press", "temperature",
         "entropy", "Ye", "Bvecx", "Bvecy", "Bvecz"]
CONS = ["dens", "momx", "momy", "momz", "tau", "DEnt", "DYe"]
DB = ["dBx", "dBy", "dBz"]
SAVED = ["saved_rho", "saved_velx", "saved_vely", "saved_velz", "saved_eps",
         "saved_Ye"]
ZS = ["zvec_x", "zvec_y", "zvec_z", "svec_x", "svec_y", "svec_z"]
FLAGS = ["con2prim_flag", "aster_mask_cc", "aster_mask_vc"]
DBSTAG = ["dBx_stag", "dBy_stag", "dBz_stag"]
AVEC = ["Avec_x", "Avec_y", "Avec_z", "Psi"]
FLUXVARS = ["dens", "momx", "momy", "momz", "tau", "DYe", "DEnt", "Bx", "By",
            "Bz"]
FLUX = ["f%s%s" % (d, v) for d in "xyz" for v in FLUXVARS]
UCT = ["vtilde_y_xface", "vtilde_z_xface", "vtilde_x_yface", "vtilde_z_yface",
       "vtilde_x_zface", "vtilde_y_zface", "amax_xface", "amin_xface",
       "amax_yface", "amin_yface", "amax_zface", "amin_zface"]
AUX = ["Fx", "Fy", "Fz", "Fbetax", "Fbetay", "Fbetaz", "G"]
RHS = ["densrhs", "momxrhs", "momyrhs", "momzrhs", "taurhs", "DEntrhs",
       "DYe_rhs", "Avec_x_rhs", "Avec_y_rhs", "Avec_z_rhs", "Psi_rhs"]
TMUNU = ["eTtt", "eTtx", "eTty", "eTtz", "eTxx", "eTxy", "eTxz", "eTyy",
         "eTyz", "eTzz"]

SLOTS = ADM + PRIMS + CONS + DB + SAVED + ZS + FLAGS + DBSTAG + AVEC + FLUX + \
    UCT + AUX + RHS + TMUNU + ["regrid_error"]
NUM_GF = len(SLOTS)
SLOT = {name: i for i, name in enumerate(SLOTS)}

# evolved state in asx_state order
STATE = ["dens", "momx", "momy", "momz", "tau", "DEnt", "DYe", "Avec_x",
         "Avec_y", "Avec_z", "Psi"]
STATE_RHS = ["densrhs", "momxrhs", "momyrhs", "momzrhs", "taurhs", "DEntrhs",
             "DYe_rhs", "Avec_x_rhs", "Avec_y_rhs", "Avec_z_rhs", "Psi_rhs"]


def centering(name):
    """Bit d set = vertex-centred along axis d (AsterX/interface.ccl)."""
    if name in ADM or name in ("aster_mask_vc", "Psi", "Psi_rhs") or name in AUX \
            or name in TMUNU:
        return 7
    if name in ("dBx_stag",) or name.startswith("fx") or name.endswith("_xface"):
        return 1
    if name in ("dBy_stag",) or name.startswith("fy") or name.endswith("_yface"):
        return 2
    if name in ("dBz_stag",) or name.startswith("fz") or name.endswith("_zface"):
        return 4
    if name in ("Avec_x", "Avec_x_rhs"):
        return 6
    if name in ("Avec_y", "Avec_y_rhs"):
        return 5
    if name in ("Avec_z", "Avec_z_rhs"):
        return 3
    return 0


# Cactus group -> slots, for the regrid_groups parameter (AsterX/param.ccl:52)
GROUPS = {
    "HydroBaseX::rho": ["rho"], "HydroBaseX::vel": ["velx", "vely", "velz"],
    "HydroBaseX::eps": ["eps"], "HydroBaseX::press": ["press"],
    "HydroBaseX::temperature": ["temperature"], "HydroBaseX::entropy": ["entropy"],
    "HydroBaseX::Ye": ["Ye"], "HydroBaseX::Bvec": ["Bvecx", "Bvecy", "Bvecz"],
    "AsterX::dens": ["dens"], "AsterX::mom": ["momx", "momy", "momz"],
    "AsterX::tau": ["tau"], "AsterX::DEnt": ["DEnt"], "AsterX::DYe": ["DYe"],
    "AsterX::dB": ["dBx", "dBy", "dBz"], "AsterX::Avec_x": ["Avec_x"],
    "AsterX::Avec_y": ["Avec_y"], "AsterX::Avec_z": ["Avec_z"], "AsterX::Psi": ["Psi"],
    "ADMBaseX::lapse": ["alp"], "ADMBaseX::shift": ["betax", "betay", "betaz"],
    "ADMBaseX::metric": ["gxx", "gxy", "gxz", "gyy", "gyz", "gzz"],
    "ADMBaseX::curv": ["kxx", "kxy", "kxz", "kyy", "kyz", "kzz"],
}


def regrid_slots(groups="HydroBaseX::rho"):
    """Slot indices of the variables of a whitespace-separated list of full
    group names, in order (AsterX_EstimateError_Setup,
    AsterX/src/estimate_error.cxx:26-70, which asserts on an unknown group:
    KeyError here)."""
    out = []
    for g in groups.split():
        out += [SLOT[n] for n in GROUPS[g]]
    return out


def extents(n, ng, cent):
    """(ex, ey, ez) of a centering on a patch."""
    return tuple(n[d] + 2 * ng + ((cent >> d) & 1) for d in range(3))


# ---- enums -----------------------------------------------------------------
RECON = {"Godunov": 0, "minmod": 1, "monocentral": 2, "ppm": 3, "eppm": 4,
         "wenoz": 5, "mp5": 6}
FLUXTYPE = {"LxF": 0, "HLLE": 1}
RECVAR = {"v_vec": 0, "z_vec": 1, "s_vec": 2}
GAUGE = {"algebraic": 0, "generalized Lorentz": 1}
EOS = {"IdealGas": 0, "Hybrid": 1, "Tabulated3d": 2}
C2P = {"None": 0, "Noble": 1, "Palenzuela": 2, "Entropy": 3}
REGRID = {"first derivative": 0, "second derivative": 1,
          "second derivative norm": 2, "first gradient": 3}
BC = {"linear": 0, "periodic": 1, "copy": 2}
# kernel classes of asx_timing_read
KCLASS = {"bstag": 0, "face_x": 1, "face_y": 2, "face_z": 3, "aux": 4,
          "update": 5, "c2p": 6, "bc": 7, "rhs": 8, "other": 9,
          "c2p_rest": 10, "mlp": 11, "graph": 12, "tmunu": 13, "regrid": 14,
          "persist": 15, "amr": 16}
C2P_FLAG = {"INIT": 0, "PRIME": 1, "SECOND": 2, "ENTROPY": 3, "ATMO": 4,
            "AVG": 5, "FAIL": 6}

STATUS = {0: "ASX_OK", 1: "ASX_ERR_NULL", 2: "ASX_ERR_LAYOUT",
          3: "ASX_ERR_PARAM", 4: "ASX_ERR_UNSUPPORTED", 5: "ASX_ERR_GHOSTS",
          6: "ASX_ERR_DEVICE", 7: "ASX_ERR_SIZE", 8: "ASX_ERR_POISON"}
POISON_BITS = 0x7ff4dead0badf00d


class AsxError(RuntimeError):
    """Raised when an entry point returns a status other than ASX_OK."""

    def __init__(self, fn, code):
        super().__init__("%s returned %s" % (fn, STATUS.get(code, code)))
        self.code = code


def check(fn, code):
    if code != 0:
        raise AsxError(fn, code)


# ---- structs ---------------------------------------------------------------
class asx_gf(C.Structure):
    _fields_ = [("data", C.c_void_p), ("sj", C.c_int64), ("sk", C.c_int64)]


class asx_patch(C.Structure):
    _fields_ = [("n", C.c_int32 * 3), ("ng", C.c_int32),
                ("x0", C.c_double * 3), ("dx", C.c_double * 3),
                ("gf", asx_gf * NUM_GF)]


class asx_state(C.Structure):
    _fields_ = [("gf", asx_gf * len(STATE))]


class asx_eos_table(C.Structure):
    _fields_ = [("data", C.c_void_p), ("axis", C.c_void_p * 3), ("n", C.c_int32 * 3),
                ("nvars", C.c_int32), ("energy_shift", C.c_double),
                ("eps_min", C.c_double), ("eps_max", C.c_double),
                ("x_lo", C.c_double * 3), ("x_step", C.c_double * 3),
                ("x_hi", C.c_double * 3)]


class asx_box(C.Structure):
    _fields_ = [("lo", C.c_int32 * 3), ("hi", C.c_int32 * 3)]


class asx_counters(C.Structure):
    _fields_ = [("nan_faces", C.c_int64), ("lo_faces", C.c_int64),
                ("c2p_flags", C.c_int64 * 7), ("c2p_iters", C.c_int64 * 64)]


_PARAM_FIELDS = [
    ("reconstruction_method", C.c_int32, "Godunov"),
    ("loworder_method", C.c_int32, "Godunov"),
    ("reconstruct_with_temperature", C.c_int32, 0),
    ("ppm_shock_detection", C.c_int32, 1),
    ("ppm_zone_flattening", C.c_int32, 0),
    ("pad0", C.c_int32, 0),
    ("ppm_eta1", C.c_double, 20.0), ("ppm_eta2", C.c_double, 0.05),
    ("ppm_eps", C.c_double, 0.33), ("ppm_eps_shock", C.c_double, 0.01),
    ("ppm_small", C.c_double, 1e-7),
    ("ppm_omega1", C.c_double, 0.75), ("ppm_omega2", C.c_double, 10.0),
    ("enhanced_ppm_C2", C.c_double, 1.25),
    ("weno_eps", C.c_double, 1e-26), ("mp5_alpha", C.c_double, 4.0),
    ("vector_potential_gauge", C.c_int32, "generalized Lorentz"),
    ("use_uct", C.c_int32, 0), ("use_entropy_fix", C.c_int32, 0),
    ("flux_type", C.c_int32, "LxF"), ("recon_type", C.c_int32, "v_vec"),
    ("local_spatial_order", C.c_int32, 2), ("excise", C.c_int32, 0),
    ("use_mask", C.c_int32, 0),
    ("lorenz_damp_fac", C.c_double, 0.0),
    ("evolution_eos", C.c_int32, "IdealGas"), ("pad1", C.c_int32, 0),
    ("poly_gamma", C.c_double, 2.0), ("poly_k", C.c_double, 100.0),
    ("gl_gamma", C.c_double, 2.0),
    ("rho_max", C.c_double, 1.0), ("rho_min", C.c_double, 1e-40),
    ("eps_max", C.c_double, 2.0), ("eps_min", C.c_double, 0.0),
    ("ye_max", C.c_double, 1.0), ("ye_min", C.c_double, 0.0),
    ("particle_mass", C.c_double, 1.0),
    ("c2p_prime", C.c_int32, "Noble"), ("c2p_second", C.c_int32, "Palenzuela"),
    ("max_iter", C.c_int32, 30), ("use_press_atmo", C.c_int32, 0),
    ("thermal_eos_atmo", C.c_int32, 1), ("Ye_lenient", C.c_int32, 0),
    ("use_z", C.c_int32, 0), ("use_temperature", C.c_int32, 0),
    ("rho_abs_min", C.c_double, 1e-13), ("atmo_tol", C.c_double, 1e-3),
    ("eps_atmo", C.c_double, 1e-10), ("p_atmo", C.c_double, 1e-10),
    ("t_atmo", C.c_double, 1e-2), ("Ye_atmo", C.c_double, 0.5),
    ("B_lim", C.c_double, 100.0), ("vw_lim", C.c_double, 10.0),
    ("c2p_tol", C.c_double, 1e-10), ("r_atmo", C.c_double, 1e100),
    ("n_rho_atmo", C.c_double, 0.0), ("n_press_atmo", C.c_double, 0.0),
    ("n_temp_atmo", C.c_double, 0.0), ("alp_thresh", C.c_double, 0.0),
    ("rho_BH", C.c_double, 1e100), ("eps_BH", C.c_double, 1e100),
    ("vwlim_BH", C.c_double, 1e100), ("cons_error_limit", C.c_double, -1.0),
    ("eos_table", C.c_void_p, None),
    ("tmunu_interp_order", C.c_int32, 4), ("update_tmunu", C.c_int32, 1),
    ("local_estimate_error", C.c_int32, 1), ("regrid_method", C.c_int32, "first gradient"),
    ("epsilon_err", C.c_double, 0.01),
]

_KEYWORDS = {"reconstruction_method": RECON, "loworder_method": RECON,
             "flux_type": FLUXTYPE, "recon_type": RECVAR,
             "vector_potential_gauge": GAUGE, "evolution_eos": EOS,
             "c2p_prime": C2P, "c2p_second": C2P, "regrid_method": REGRID}


class asx_params(C.Structure):
    _fields_ = [(n, t) for n, t, _ in _PARAM_FIELDS]


def params(**kw):
    """asx_params with the param.ccl defaults, overridden by keyword.

    Keyword parameters accept the reference's strings ("ppm", "HLLE", ...)
    or the integer enum value; booleans accept True/False/"yes"/"no".
    """
    p = asx_params()
    vals = {n: d for n, _, d in _PARAM_FIELDS}
    for k, v in kw.items():
        if k not in vals:
            raise KeyError("unknown parameter %r" % k)
        vals[k] = v
    for n, t, _ in _PARAM_FIELDS:
        v = vals[n]
        if n in _KEYWORDS and isinstance(v, str):
            v = _KEYWORDS[n][v]
        if isinstance(v, str):
            v = {"yes": 1, "no": 0}[v.lower()]
        if t is C.c_void_p:
            setattr(p, n, v)
        else:
            setattr(p, n, t(v).value if t is not C.c_double else float(v))
    return p
