/*
 * asterx.h -- C ABI of the MI355X-native GRMHD method-of-lines substep.
 *
 * One CarpetX-style patch: N[3] interior cells, ng ghost zones on every side.
 * Every grid function is a caller-owned, dense, x-fastest FP64 array over its
 * centering's full extent including ghosts (cell axis: N+2ng points, vertex
 * axis: N+1+2ng points), the layout AMReX/CarpetX allocate for a box.  The
 * index (i,j,k) of a point counts from the first ghost point, so interior cells
 * are ng..ng+N-1 and interior vertices ng..ng+N on each axis.  Cell i spans
 * [x0+(i-ng)dx, x0+(i-ng+1)dx]; vertex i sits at x0+(i-ng)dx.
 *
 * Entry points mirror the reference's scheduled functions one to one
 * (names in the comment of each declaration, file:line in the reference
 * tree), take plain pointers and sizes, return an asx_status and never abort.
 * Numerical failures are data (con2prim flags, counters), as in the reference.
 * All work is stream-ordered on the given HIP stream (NULL = default stream).
 * Device buffers the library keeps (face data of asx_substep, the queue of
 * asx_con2prim*) are held per device and stream, so calls on different
 * streams do not share them; the counters buffer is per device, so calls that
 * pass asx_counters must not overlap on one device.
 */
#ifndef ASTERX_H
#define ASTERX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------ */
/* Status codes                                                             */
/* ------------------------------------------------------------------------ */
typedef enum {
  ASX_OK = 0,
  ASX_ERR_NULL = 1,        /* required pointer missing                      */
  ASX_ERR_LAYOUT = 2,      /* strides / extents do not match the patch      */
  ASX_ERR_PARAM = 3,       /* invalid parameter value (reference: CCTK_ERROR) */
  ASX_ERR_UNSUPPORTED = 4, /* valid in the reference, not built here        */
  ASX_ERR_GHOSTS = 5,      /* too few ghost zones for the stencil           */
  ASX_ERR_DEVICE = 6,      /* HIP runtime error                             */
  ASX_ERR_SIZE = 7,        /* size out of range                             */
  ASX_ERR_POISON = 8       /* debug poison: a declared output point was not
                              written (asx_set_poison)                      */
} asx_status;

/* ------------------------------------------------------------------------ */
/* Keyword parameters (reference param.ccl keywords as enums)               */
/* ------------------------------------------------------------------------ */
/* ReconX/param.ccl:5-21 reconstruction_method / loworder_method */
enum {
  ASX_RECON_GODUNOV = 0,
  ASX_RECON_MINMOD = 1,
  ASX_RECON_MONOCENTRAL = 2,
  ASX_RECON_PPM = 3,
  ASX_RECON_EPPM = 4, /* accepted by the reference; not built here */
  ASX_RECON_WENOZ = 5,
  ASX_RECON_MP5 = 6
};
/* AsterX/param.ccl:22-26 */
enum { ASX_FLUX_LXF = 0, ASX_FLUX_HLLE = 1 };
/* AsterX/param.ccl:28-33 */
enum { ASX_RECVAR_V = 0, ASX_RECVAR_Z = 1, ASX_RECVAR_S = 2 };
/* AsterX/param.ccl:3-7 */
enum { ASX_GAUGE_ALGEBRAIC = 0, ASX_GAUGE_GENERALIZED_LORENZ = 1 };
/* EOSX/param.ccl:14-20 (Hybrid is not built here) */
enum { ASX_EOS_IDEALGAS = 0, ASX_EOS_HYBRID = 1, ASX_EOS_TABULATED = 2 };
/* Con2PrimFactory/param.ccl:8-22 (Entropy: ideal gas only, as in the reference) */
enum {
  ASX_C2P_NONE = 0,
  ASX_C2P_NOBLE = 1,
  ASX_C2P_PALENZUELA = 2,
  ASX_C2P_ENTROPY = 3
};

/* AsterX/param.ccl:39-45 regrid_method */
enum {
  ASX_REGRID_FIRST_DERIV = 0,       /* "first derivative"       */
  ASX_REGRID_SECOND_DERIV = 1,      /* "second derivative"      */
  ASX_REGRID_SECOND_DERIV_NORM = 2, /* "second derivative norm" */
  ASX_REGRID_FIRST_GRAD = 3         /* "first gradient"         */
};

/* AsterX/src/con2prim.cxx:27-35 -- written to con2prim_flag as FP64 */
enum {
  ASX_C2P_FLAG_INIT = 0,
  ASX_C2P_FLAG_PRIME = 1,
  ASX_C2P_FLAG_SECOND = 2,
  ASX_C2P_FLAG_ENTROPY = 3,
  ASX_C2P_FLAG_ATMO = 4,
  ASX_C2P_FLAG_AVG = 5,
  ASX_C2P_FLAG_FAIL = 6
};

/*
 * Parameters: same names, meaning and defaults as the reference's param.ccl
 * files (AsterX/param.ccl, ReconX/param.ccl, EOSX/param.ccl,
 * Con2PrimFactory/param.ccl).  asx_params_default() sets the defaults.
 */
/* A tabulated EOS as EOSX holds it after reading its table
 * (EOSX/src/eos_3p_tabulated3d/eos_3p_tabulated3d.hxx:60-183): uniform axes
 * log rho and log T (code units, T in MeV) and Ye, and per node the
 * variables log p, log(eps + energy_shift), c_s^2 (code units) and the
 * entropy per baryon, the variable index fastest:
 * data[v + 4 (i + n[0] (j + n[1] k))].  Device
 * memory; the library reads it in place.  eps_min / eps_max are the
 * table's eps range (EOSX rgeps): min and max of exp(log(eps + shift)) -
 * shift over the nodes, the minimum floored at 1e-15. */
typedef struct asx_eos_table_s {
  const double *data;
  const double *axis[3]; /* log rho (n[0] values), log T (n[1]), Ye (n[2]) */
  int32_t n[3];          /* >= 2 each */
  int32_t nvars;         /* 4 */
  double energy_shift;
  double eps_min, eps_max;
  /* host copies of axis[d][0], axis[d][1] - axis[d][0], axis[d][n[d]-1]
   * (the library does not read the device table from the host) */
  double x_lo[3], x_step[3], x_hi[3];
} asx_eos_table;

typedef struct {
  /* ReconX */
  int32_t reconstruction_method;        /* "Godunov"       */
  int32_t loworder_method;              /* "Godunov"       */
  int32_t reconstruct_with_temperature; /* no              */
  int32_t ppm_shock_detection;          /* yes             */
  int32_t ppm_zone_flattening;          /* no              */
  int32_t pad0;
  double ppm_eta1, ppm_eta2;            /* 20, 0.05        */
  double ppm_eps, ppm_eps_shock;        /* 0.33, 0.01      */
  double ppm_small;                     /* 1e-7            */
  double ppm_omega1, ppm_omega2;        /* 0.75, 10        */
  double enhanced_ppm_C2;               /* 1.25            */
  double weno_eps;                      /* 1e-26           */
  double mp5_alpha;                     /* 4               */
  /* AsterX */
  int32_t vector_potential_gauge; /* "generalized Lorentz" */
  int32_t use_uct;                /* no                    */
  int32_t use_entropy_fix;        /* no                    */
  int32_t flux_type;              /* "LxF"                 */
  int32_t recon_type;             /* "v_vec"               */
  int32_t local_spatial_order;    /* 2                     */
  int32_t excise;                 /* no                    */
  int32_t use_mask;               /* no                    */
  double lorenz_damp_fac;         /* 0                     */
  /* EOSX */
  int32_t evolution_eos; /* "IdealGas" */
  int32_t pad1;
  double poly_gamma, poly_k;   /* 2, 100 (initial-data polytrope) */
  double gl_gamma;             /* 2                              */
  double rho_max, rho_min;     /* 1, 1e-40                        */
  double eps_max, eps_min;     /* 2, 0                            */
  double ye_max, ye_min;       /* 1, 0                            */
  double particle_mass;        /* 1                               */
  /* Con2PrimFactory */
  int32_t c2p_prime;        /* "Noble"      */
  int32_t c2p_second;       /* "Palenzuela" */
  int32_t max_iter;         /* 30           */
  int32_t use_press_atmo;   /* no           */
  int32_t thermal_eos_atmo; /* yes          */
  int32_t Ye_lenient;       /* no           */
  int32_t use_z;            /* no           */
  int32_t use_temperature;  /* no           */
  double rho_abs_min;       /* 1e-13        */
  double atmo_tol;          /* 1e-3         */
  double eps_atmo, p_atmo;  /* 1e-10, 1e-10 */
  double t_atmo;            /* 1e-2         */
  double Ye_atmo;           /* 0.5          */
  double B_lim, vw_lim;     /* 100, 10      */
  double c2p_tol;           /* 1e-10        */
  double r_atmo;            /* 1e100        */
  double n_rho_atmo, n_press_atmo, n_temp_atmo; /* 0, 0, 0 */
  double alp_thresh;        /* 0            */
  double rho_BH, eps_BH, vwlim_BH; /* 1e100 each */
  double cons_error_limit;  /* -1           */
  /* EOSX (evolution_eos = "Tabulated3d") */
  const struct asx_eos_table_s *eos_table; /* NULL */
  /* AsterX Tmunu */
  int32_t tmunu_interp_order; /* 4 (2 or 4) */
  int32_t update_tmunu;       /* yes        */
  /* AsterX regrid error estimate */
  int32_t local_estimate_error; /* yes              */
  int32_t regrid_method;        /* "first gradient" */
  double epsilon_err;           /* 0.01             */
} asx_params;

/* ------------------------------------------------------------------------ */
/* Grid functions                                                           */
/* ------------------------------------------------------------------------ */
/* One grid function: x-fastest, element strides for j and k (stride of i is
 * 1).  The strides must equal the dense strides of the grid function's
 * centering on the patch (checked; ASX_ERR_LAYOUT otherwise). */
typedef struct {
  double *data;
  int64_t sj, sk;
} asx_gf;

/* Grid-function slots.  Names and groups follow AsterX/interface.ccl:16-165,
 * HydroBaseX and ADMBaseX.  The centering of each slot is given by
 * asx_gf_centering(). */
enum {
  /* ADMBaseX, vertex {vvv} */
  ASX_ALP = 0, ASX_BETAX, ASX_BETAY, ASX_BETAZ,
  ASX_GXX, ASX_GXY, ASX_GXZ, ASX_GYY, ASX_GYZ, ASX_GZZ,
  ASX_KXX, ASX_KXY, ASX_KXZ, ASX_KYY, ASX_KYZ, ASX_KZZ,
  /* HydroBaseX, cell {ccc} */
  ASX_RHO, ASX_VELX, ASX_VELY, ASX_VELZ, ASX_EPS, ASX_PRESS,
  ASX_TEMPERATURE, ASX_ENTROPY, ASX_YE, ASX_BVECX, ASX_BVECY, ASX_BVECZ,
  /* AsterX conserved state and densitized B, cell */
  ASX_DENS, ASX_MOMX, ASX_MOMY, ASX_MOMZ, ASX_TAU, ASX_DENT, ASX_DYE,
  ASX_DBX, ASX_DBY, ASX_DBZ,
  /* saved_prims, cell */
  ASX_SAVED_RHO, ASX_SAVED_VELX, ASX_SAVED_VELY, ASX_SAVED_VELZ,
  ASX_SAVED_EPS, ASX_SAVED_YE,
  /* zvec, svec, cell */
  ASX_ZVEC_X, ASX_ZVEC_Y, ASX_ZVEC_Z, ASX_SVEC_X, ASX_SVEC_Y, ASX_SVEC_Z,
  /* flags and masks */
  ASX_CON2PRIM_FLAG, /* cell   */
  ASX_MASK_CC,       /* cell   (AsterMasks::aster_mask_cc) */
  ASX_MASK_VC,       /* vertex (AsterMasks::aster_mask_vc) */
  /* staggered densitized B: {vcc}, {cvc}, {ccv} */
  ASX_DBX_STAG, ASX_DBY_STAG, ASX_DBZ_STAG,
  /* vector potential: {cvv}, {vcv}, {vvc}; Psi {vvv} */
  ASX_AVEC_X, ASX_AVEC_Y, ASX_AVEC_Z, ASX_PSI,
  /* face fluxes flux_x {vcc}: order as AsterX/interface.ccl:39-47 */
  ASX_FXDENS, ASX_FXMOMX, ASX_FXMOMY, ASX_FXMOMZ, ASX_FXTAU, ASX_FXDYE,
  ASX_FXDENT, ASX_FXBX, ASX_FXBY, ASX_FXBZ,
  /* flux_y {cvc} */
  ASX_FYDENS, ASX_FYMOMX, ASX_FYMOMY, ASX_FYMOMZ, ASX_FYTAU, ASX_FYDYE,
  ASX_FYDENT, ASX_FYBX, ASX_FYBY, ASX_FYBZ,
  /* flux_z {ccv} */
  ASX_FZDENS, ASX_FZMOMX, ASX_FZMOMY, ASX_FZMOMZ, ASX_FZTAU, ASX_FZDYE,
  ASX_FZDENT, ASX_FZBX, ASX_FZBY, ASX_FZBZ,
  /* upwind-CT face data (AsterX/interface.ccl:137-165) */
  ASX_VTILDE_Y_XFACE, ASX_VTILDE_Z_XFACE, /* {vcc} */
  ASX_VTILDE_X_YFACE, ASX_VTILDE_Z_YFACE, /* {cvc} */
  ASX_VTILDE_X_ZFACE, ASX_VTILDE_Y_ZFACE, /* {ccv} */
  ASX_AMAX_XFACE, ASX_AMIN_XFACE,         /* {vcc} */
  ASX_AMAX_YFACE, ASX_AMIN_YFACE,         /* {cvc} */
  ASX_AMAX_ZFACE, ASX_AMIN_ZFACE,         /* {ccv} */
  /* Aux_in_RHSof_A_Psi, vertex */
  ASX_FX, ASX_FY, ASX_FZ, ASX_FBETAX, ASX_FBETAY, ASX_FBETAZ, ASX_G,
  /* right-hand sides */
  ASX_DENSRHS, ASX_MOMXRHS, ASX_MOMYRHS, ASX_MOMZRHS, ASX_TAURHS,
  ASX_DENTRHS, ASX_DYE_RHS,                       /* cell            */
  ASX_AVEC_X_RHS, ASX_AVEC_Y_RHS, ASX_AVEC_Z_RHS, /* {cvv},{vcv},{vvc} */
  ASX_PSI_RHS,                                    /* vertex          */
  /* TmunuBaseX stress-energy tensor, vertex (accumulated by asx_tmunu) */
  ASX_ETTT, ASX_ETTX, ASX_ETTY, ASX_ETTZ,
  ASX_ETXX, ASX_ETXY, ASX_ETXZ, ASX_ETYY, ASX_ETYZ, ASX_ETZZ,
  /* CarpetXRegrid::regrid_error, cell (written by asx_estimate_error) */
  ASX_REGRID_ERROR,
  ASX_NUM_GF
};

/* A patch: geometry plus one asx_gf per slot.  Unused slots may be NULL;
 * each entry point lists the slots it reads and writes. */
typedef struct {
  int32_t n[3];  /* interior cells per axis                       */
  int32_t ng;    /* ghost zones per side (same on every axis)     */
  double x0[3];  /* coordinate of interior vertex ng on each axis */
  double dx[3];  /* cell widths                                    */
  asx_gf gf[ASX_NUM_GF];
} asx_patch;

/* The evolved state (the groups tagged rhs= in AsterX/interface.ccl:16-102),
 * in this order: dens, momx, momy, momz, tau, DEnt, DYe, Avec_x, Avec_y,
 * Avec_z, Psi.  Used for RK combinations and the fused substep. */
enum { ASX_NUM_STATE = 11 };
typedef struct {
  asx_gf gf[ASX_NUM_STATE];
} asx_state;

/* Optional counters filled by the entry points that take them (NULL = skip).
 * They replace the reference's device asserts / printf diagnostics
 * (AsterX/src/fluxes.cxx:680-770, AsterX/src/rhs.cxx:199-206). */
typedef struct {
  int64_t nan_faces;         /* faces with NaN state/flux or negative rho, p  */
  int64_t lo_faces;          /* face sides that fell back to low order        */
  int64_t c2p_flags[7];      /* histogram of con2prim_flag over interior cells */
  int64_t c2p_iters[64];     /* histogram of primary-solver iterations (0..63, last bin = overflow) */
} asx_counters;

/* ------------------------------------------------------------------------ */
/* Setup                                                                    */
/* ------------------------------------------------------------------------ */
/* Defaults of every param.ccl entry used by the path. */
void asx_params_default(asx_params *p);
/* Centering of a slot: bit d set = vertex-centred along axis d. */
int asx_gf_centering(int slot);
/* Dense strides (sj, sk) of a slot on a patch with n[3], ng. */
void asx_gf_strides(const int32_t n[3], int32_t ng, int slot, int64_t *sj,
                    int64_t *sk);
/* Library version string. */
const char *asx_version(void);

/* Kernel timing (diagnostics; no counterpart in the reference).  While
 * enabled, every kernel launch of the library is bracketed by HIP events on
 * the caller's stream.  asx_timing_read returns, for one kernel class, the
 * number of launches and their summed durations in milliseconds since the
 * last asx_timing_enable call, waiting for launches still in flight.
 * Enabling or disabling resets the sums. */
enum {
  ASX_KC_BSTAG = 0,  /* dB*_stag from A */
  ASX_KC_FACE_X = 1, /* face reconstruction + numerical flux, x faces */
  ASX_KC_FACE_Y = 2,
  ASX_KC_FACE_Z = 3,
  ASX_KC_AUX = 4,    /* vertex auxiliaries of the A/Psi gauge (batch plans;
                        on one patch they ride on the z-face launch) */
  ASX_KC_UPDATE = 5, /* fused RHS + RK combination */
  ASX_KC_C2P = 6,    /* con2prim: every zone, primary solver (patch or batch) */
  ASX_KC_BC = 7,     /* boundary fills */
  ASX_KC_RHS = 8,    /* unfused source terms / RHS */
  ASX_KC_OTHER = 9,
  ASX_KC_C2P_REST = 10, /* con2prim: zones the primary solver failed */
  ASX_KC_MLP = 11,      /* learned con2prim guess: inference and training */
  ASX_KC_GRAPH = 12,    /* replays of a batch plan's captured step */
  ASX_KC_TMUNU = 13,    /* stress-energy tensor at the vertices */
  ASX_KC_REGRID = 14,   /* regrid error estimate */
  ASX_KC_PERSIST = 15,  /* a batch plan's stage as one persistent launch */
  ASX_KC_AMR = 16,      /* prolongation, restriction and regrid tags */
  ASX_KC_COUNT = 17
};
asx_status asx_timing_enable(int32_t on);
asx_status asx_timing_read(int32_t kclass, int64_t *launches, double *ms);

/* Debug poison, the device form of CarpetX::poison_undefined_values = yes
 * (FishboneMoncriefIDX/par/FMdisk_unigrid.par:26).  While enabled, the
 * entry points named below first fill every grid function they only write
 * (not those they also read) with ASX_POISON_BITS on every point, ghosts
 * included, then run, then count the points of the region they declare
 * written that still hold the pattern; any such point makes the call return
 * ASX_ERR_POISON, and asx_poison_report gives the first such slot and its
 * count.  Points outside the declared region keep the pattern, so a later
 * entry point that reads them computes NaN.  Poisoned: asx_fluxes (fluxes,
 * UCT face data, aux), asx_source_terms (hydro RHS), asx_rhs (A and Psi
 * RHS), asx_bstag_from_A, asx_dB_from_bstag, asx_B_from_dB, asx_prim2con,
 * asx_estimate_error, asx_substep and asx_substep_bc (the out state, when it
 * shares no array with the inputs; the library's face scratch is poisoned
 * too).  Each checked call waits for its stream; the count buffer is per
 * device, so checked calls must not overlap on one device (as for
 * asx_counters).  asx_poison_check counts the pattern in any box of one
 * slot (lo, cnt in point indices of that slot's layout). */
#define ASX_POISON_BITS 0x7ff4dead0badf00dULL /* a signalling NaN */
asx_status asx_set_poison(int32_t on);
asx_status asx_poison_report(int32_t *slot, int64_t *points);
asx_status asx_poison_check(const asx_patch *patch, int32_t slot, const int32_t lo[3],
                            const int32_t cnt[3], int64_t *points, void *stream);

/* Launch-shape variants and autotuning (no counterpart in the reference).
 * The face kernels (ASX_KC_FACE_*) and the fused update (ASX_KC_UPDATE) are
 * built in several workgroup tile shapes; results do not depend on the
 * variant.  Face variants 5-7 also stage the tile's vertex metric in LDS
 * (up to 52 KB of LDS per workgroup with the face hand-over, three
 * workgroups per CU).  ASX_KC_MLP has two forms of asx_c2p_batch_mlp (network fused
 * into the solver kernel, or a separate forward pass), whose guesses differ
 * by the fp32 summation order only.  The choice is per device (default 0).  asx_autotune_substep
 * runs asx_substep with the given arguments once untimed and `reps` times
 * timed for every variant, keeps the fastest per kernel class, and leaves the
 * caller's timing state as it was.  It overwrites `out`. */
int32_t asx_variant_count(int32_t kclass);
asx_status asx_set_variant(int32_t kclass, int32_t v);
asx_status asx_get_variant(int32_t kclass, int32_t *v);
asx_status asx_autotune_substep(const asx_patch *stage, const asx_state *u0,
                                const asx_state *out, double c0, double c1,
                                double c2, const asx_params *par, int32_t reps,
                                void *stream);

/* ------------------------------------------------------------------------ */
/* Scheduled functions (drop-in for the reference's extern "C" entry points) */
/* ------------------------------------------------------------------------ */
/* AsterX_Fluxes (AsterX/src/fluxes.cxx:828-878): CalcFlux<0,1,2>
 * (fluxes.cxx:38-796) + CalcAuxForAvecPsi (fluxes.cxx:798-826).
 * Reads prims, dB*_stag, ADM lapse/shift/metric, Avec, Psi.
 * Writes flux_x/y/z, UCT face data (when use_uct) and Aux on interior points.
 * face_halo > 0 additionally computes faces and aux vertices that many layers
 * into the ghost region (every axis), which is what a neighbour patch's SYNC
 * delivers there; face_halo = 0 is the reference's interior-only write. */
asx_status asx_fluxes(const asx_patch *patch, const asx_params *par,
                      int32_t face_halo, asx_counters *cnt, void *stream);

/* AsterX_SourceTerms (AsterX/src/source.cxx:170-185, SourceTerms<2|4>
 * source.cxx:20-168). Writes densrhs..DYe_rhs on interior cells. */
asx_status asx_source_terms(const asx_patch *patch, const asx_params *par,
                            void *stream);

/* AsterX_Tmunu (AsterX/src/tmunu.cxx:19-176, scheduled in
 * TmunuBaseX_AddToTmunu when update_tmunu, AsterX/schedule.ccl:253-266):
 * interpolates rho, eps, press, Bvec and vel (recon_type v_vec) or zvec
 * (otherwise) from cells to the interior vertices with calc_avg_c2v of order
 * tmunu_interp_order (AsterUtils/src/aster_interp.hxx:97-148) and ADDS the
 * lower-index T_{mu nu} of the fluid and the field into eTtt..eTzz (the
 * caller zeroes them first, as TmunuBaseX does).  Reads alp, beta^i, g_ij at
 * the vertices and the cell primitives on ghosts (ng >= order/2, else
 * ASX_ERR_GHOSTS).  tmunu_interp_order other than 2 or 4: ASX_ERR_PARAM (the
 * reference's CCTK_VERROR).  update_tmunu = no: nothing is done.  Grid functions
 * of 4 GiB or more (vertex extent): ASX_ERR_SIZE. */
asx_status asx_tmunu(const asx_patch *patch, const asx_params *par, void *stream);

/* AsterX_EstimateError (AsterX/src/estimate_error.cxx:72-142, estimators
 * estimate_error.hxx:14-100), scheduled IN ODESolvers_EstimateError when
 * local_estimate_error: writes regrid_error on the interior cells as the
 * maximum over the nvars grid functions slots[] of the regrid_method
 * estimator (differences to the +-1 neighbours on each axis; epsilon_err for
 * "second derivative norm").  slots[] replaces the regrid_groups string: the
 * variables of those groups, each read at the cell's index (i,j,k) in its own
 * centering's layout, as the reference does.  nvars = 0 writes 0; nvars >= 10
 * (the reference's assert) or a slot out of range: ASX_ERR_SIZE / ASX_ERR_PARAM;
 * an unknown regrid_method: ASX_ERR_PARAM; local_estimate_error = no: nothing
 * is done. */
asx_status asx_estimate_error(const asx_patch *patch, const asx_params *par, int32_t nvars,
                              const int32_t *slots, void *stream);

/* AsterX_RHS (AsterX/src/rhs.cxx:18-253): adds the flux divergence to the
 * hydro RHS, writes Avec_*_rhs (flux-CT or UCT EMF + gauge) and Psi_rhs. */
asx_status asx_rhs(const asx_patch *patch, const asx_params *par, void *stream);

/* AsterX_ComputedBstagFromA (AsterX/src/computeBfromA.cxx:58-65). */
asx_status asx_bstag_from_A(const asx_patch *patch, void *stream);
/* AsterX_ComputedBFromdBstag (AsterX/src/computeBfromA.cxx:67-85). */
asx_status asx_dB_from_bstag(const asx_patch *patch, void *stream);
/* AsterX_ComputeBFromdB (AsterX/src/computeBfromA.cxx:87-113). */
asx_status asx_B_from_dB(const asx_patch *patch, void *stream);
/* AsterX_Prim2Con_Initial (AsterX/src/prim2con.cxx:12-52). */
asx_status asx_prim2con(const asx_patch *patch, void *stream);

/* AsterX_Con2Prim (AsterX/src/con2prim.cxx:457-502 -> :37-455) for the
 * ideal-gas EOS: atmosphere, seeds, cons floors, Noble / Palenzuela chain,
 * flags, write-back of prims, zvec, svec, cons, saved_prims, mask. */
asx_status asx_con2prim(const asx_patch *patch, const asx_params *par,
                        asx_counters *cnt, void *stream);

/* ------------------------------------------------------------------------ */
/* Driver pieces the reference takes from CarpetX (not vendored there)      */
/* ------------------------------------------------------------------------ */
/* ODESolvers state update: dst = scale*dst + sum_k fac[k]*src[k] on every
 * point of every state grid function (ghosts included). nsrc <= 8. */
asx_status asx_state_lincomb(const asx_patch *patch, const asx_state *dst,
                             double scale, int32_t nsrc, const double *fac,
                             const asx_state *const *src, void *stream);

/* Single-patch physical boundary fill of one grid function's ghost zones
 * (CarpetX "linear extrapolation": ghost = nearest interior point plus
 * distance times the difference of the two nearest interior points; or
 * periodic). kind: 0 = linear extrapolation, 1 = periodic, 2 = copy
 * (zeroth order). */
enum { ASX_BC_LINEAR = 0, ASX_BC_PERIODIC = 1, ASX_BC_COPY = 2 };
asx_status asx_bc_fill(const asx_patch *patch, int32_t slot, int32_t kind,
                       void *stream);
/* The same fill for a list of grid functions (any centerings, at most 64) in
 * one launch per axis: the physical-boundary half of a SYNC statement with
 * several groups (e.g. AsterX/schedule.ccl:78-81).  Same arguments and errors
 * as asx_bc_fill for each listed slot; ASX_ERR_PARAM for nslot > 64. */
asx_status asx_bc_fill_many(const asx_patch *patch, int32_t nslot,
                            const int32_t *slots, int32_t kind, void *stream);

/* Mesh refinement by a factor 2 between two patches of adjacent levels, the
 * CarpetX operators the AsterX regression runs select with
 * CarpetX::prolongation_type = "ddf" (e.g. AsterX/test/Balsara1_shocktube_
 * xdir_ppm.par:59) and CarpetX::max_num_levels > 1.  The fine patch's dx is
 * half the coarse one's, and its interior starts on a coarse cell face:
 * (fine.x0 - coarse.x0) / fine.dx is an even integer on every axis and every
 * fine n[d] is even (ASX_ERR_LAYOUT otherwise).
 *
 * asx_prolongate: coarse -> fine for each listed grid function (at most 64),
 * per axis by its centering ("discrete differential forms"): along a vertex
 * axis a fine point on a coarse point takes its value, a fine point between
 * two takes the centred Lagrange interpolant of degree `order` (order + 1
 * coarse points); along a cell axis each of the two children of a coarse
 * cell takes the average over its half of the polynomial of degree
 * order - 1 whose cell averages match the order coarse cells centred on the
 * parent (so the two children average to the parent).  3-D by tensor
 * product, x innermost.  region 0 writes every point of the fine patch (a
 * new patch after regridding), region 1 only its ghost points (the points
 * outside the patch's own cells and boundary vertices: the coarse-fine
 * boundary fill before a substep).  order 1, 3 or 5 (others:
 * ASX_ERR_UNSUPPORTED); a stencil reaching past the coarse patch's points:
 * ASX_ERR_GHOSTS.  CarpetX's source is not part of the reference tree; the
 * operators are restated from the discrete-differential-form rule above.
 *
 * asx_restrict: fine -> coarse on the coarse patch's own points covered by
 * the fine patch's own points: along a vertex axis the coinciding fine point,
 * along a cell axis the mean of the two children; tensor product. */
asx_status asx_prolongate(const asx_patch *coarse, const asx_patch *fine, int32_t nslot,
                          const int32_t *slots, int32_t order, int32_t region,
                          void *stream);
asx_status asx_restrict(const asx_patch *fine, const asx_patch *coarse, int32_t nslot,
                        const int32_t *slots, void *stream);

/* Refined boxes for the next level from regrid_error (the box table AMReX
 * builds for CarpetX at a regrid, CarpetX::regrid_error_threshold,
 * CarpetX::blocking_factor_*, CarpetX::max_grid_size_*; AsterX/test/
 * Balsara1_shocktube_xdir_ppm.par:52-57).  A cell is tagged when its
 * regrid_error exceeds `threshold`; tags grow by n_error_buf cells (inside
 * the interior), are coarsened to blocks of blocking_factor/2 cells and
 * clustered by Berger & Rigoutsos (1991): a box whose tagged fraction is at
 * least grid_eff (or one block) is kept, otherwise it is cut at the empty
 * plane nearest its middle, else at the strongest inflection of the tag
 * signature, else in half along its longest axis.  The boxes are returned
 * in FINE cell indices (interior cell i of the patch covers fine cells 2i,
 * 2i+1), inclusive, aligned to blocking_factor and cut to at most
 * max_grid_size fine cells per axis.  Each blocking_factor[d] must be a
 * power of two dividing max_grid_size[d], each n[d] a multiple of
 * max(1, blocking_factor[d] / 2), n_error_buf >= 0 and 0 < grid_eff <= 1
 * (ASX_ERR_PARAM otherwise).  At most max_boxes are written; *nboxes is the
 * number found (ASX_ERR_SIZE when it exceeds max_boxes).  Waits for the
 * stream. */
typedef struct {
  int32_t lo[3], hi[3];
} asx_box;
asx_status asx_regrid_boxes(const asx_patch *patch, double threshold, int32_t n_error_buf,
                            const int32_t blocking_factor[3],
                            const int32_t max_grid_size[3], double grid_eff,
                            int32_t max_boxes, asx_box *boxes, int32_t *nboxes,
                            void *stream);

/* ------------------------------------------------------------------------ */
/* Fused substep (the performance path; no reference counterpart as one    */
/* call: it is AsterX_Fluxes + AsterX_SourceTerms + AsterX_RHS + the        */
/* ODESolvers update + AsterX_Con2PrimGroup on one patch)                   */
/* ------------------------------------------------------------------------ */
/* Stencil half of a substep.  Requires valid ghost data for prims, Avec and
 * Psi on `stage` (as after a SYNC); the faces take dB*_stag as the curl of
 * Avec at the face, exactly as AsterX_ComputedBstagFromA computes it, so
 * dB*_stag is neither read nor written.  Computes
 * out = c0*u0 + c1*u + c2*rhs(u) on interior points, where u is the state
 * held in `stage` and rhs is SourceTerms + RHS of the reference.  u0 is not
 * read when c0 == 0.  Face scratch is allocated on first use per patch shape
 * and kept (one buffer set per device and stream).  use_uct is not supported
 * here (ASX_ERR_UNSUPPORTED): the upwind CT reads face data three rows into
 * the ghost region, which needs a SYNC of the faces (asx_substep_bc). */
asx_status asx_substep(const asx_patch *stage, const asx_state *u0,
                       const asx_state *out, double c0, double c1, double c2,
                       const asx_params *par, void *stream);

/* asx_substep for a patch whose ghost zones are filled by a physical
 * boundary condition rather than by neighbour patches: the face fluxes, the
 * vertex aux and, with use_uct, the upwind-CT face data and dB*_stag
 * (AsterX_ComputedBstagFromA) are computed on the interior only and their
 * ghost points filled with bc_kind (ASX_BC_*), which is AsterX_Fluxes
 * followed by the SYNC of AsterX/schedule.ccl:191-193 on a single patch;
 * then the same RHS and combination as asx_substep.  Supports use_uct (the
 * upwind CT reconstructs the SYNCed face data, AsterX/src/rhs.cxx:124-168).
 * bc_kind out of range: ASX_ERR_PARAM; too few interior points for the
 * fill: ASX_ERR_SIZE. */
asx_status asx_substep_bc(const asx_patch *stage, const asx_state *u0,
                          const asx_state *out, double c0, double c1, double c2,
                          const asx_params *par, int32_t bc_kind, void *stream);

/* PostStep half of a substep, fused: AsterX_ComputedBstagFromA +
 * AsterX_ComputedBFromdBstag + AsterX_Con2Prim on interior cells
 * (AsterX/schedule.ccl:119-165).  Writes dB*_stag only if those slots are
 * set. */
asx_status asx_con2prim_group(const asx_patch *patch, const asx_params *par,
                              asx_counters *cnt, void *stream);

/* Frees the device buffers the library keeps for `stream` on the current
 * device -- the face scratch of asx_substep / asx_substep_bc for every patch
 * shape used on that stream and the con2prim queue -- after waiting for the
 * stream (no counterpart in the reference).  For drivers whose patch shapes
 * change (after a regrid) or that retire a stream; the next call on the
 * stream allocates again.  Must not overlap with calls on that stream. */
asx_status asx_release_scratch(void *stream);

/* asx_con2prim_group that allocates nothing: no counterpart in the
 * reference.  For callers that cannot allocate device memory (code under HIP
 * graph capture, serving loops, harnesses that forbid hipMalloc).  The same
 * kernels and results as asx_con2prim_group; the queue of zones the primary
 * solver failed, its count and the counters live in a fixed device array
 * of the library (64 MiB + 2 KiB per device, held until the process ends)
 * instead of library-allocated buffers, and the launches are not timed
 * (asx_timing_*).  Patches of more than 2^24 interior zones: ASX_ERR_SIZE.
 * Every call on a device shares that device's array: calls that overlap on
 * one device (two streams, or two host threads) race; calls in one stream
 * order are safe.  The count and the counters are cleared by kernels on the
 * caller's stream at the start of each call. */
asx_status asx_con2prim_group_noalloc(const asx_patch *patch, const asx_params *par,
                                      asx_counters *cnt, void *stream);

/* Con2PrimFactory solve on a batch of independent zones (no grid): the
 * c2p_2DNoble / c2p_1DPalenzuela call of con2prim.cxx:138-306 with the
 * atmosphere and floors of con2prim.cxx:82-133 for r <= r_atmo.
 * Inputs (length nz each): cons[8][nz] = dens, momx, momy, momz, tau, dBx,
 * dBy, dBz (densitized); g[6][nz] = gxx, gxy, gxz, gyy, gyz, gzz;
 * seed[5][nz] = saved rho, eps, velx, vely, velz.
 * Outputs: prim[8][nz] = rho, eps, press, velx, vely, velz, w_lor, temperature;
 * flag[nz] (C2PFlag) and iters[nz] (primary-solver iterations). */
asx_status asx_c2p_batch(int64_t nz, const double *cons, const double *g,
                         const double *seed, double *prim, double *flag,
                         int32_t *iters, const asx_params *par, void *stream);

/* ------------------------------------------------------------------------ */
/* Independent zones and the learned con2prim guess (no counterpart in the  */
/* reference: workload generation and the network around the Noble solver)  */
/* ------------------------------------------------------------------------ */
/* Zones of the learned-guess workload: rho log-uniform in [1e-10, 1e-2],
 * p/rho log-uniform in [1e-6, 1e2], W uniform in [1, 10] along a random
 * direction, b^2/rho log-uniform in [1e-4, 1e2] along a random direction, a
 * symmetric 3-metric within 10% of the identity; ideal gas `gamma`.
 * prim[8][nz] = rho, eps, v^i, B^i; g[6][nz]; seed_out[5][nz] (optional) =
 * rho, eps, v^i of the truth times (1 + seed_rel u), u ~ U[-1, 1];
 * target[2][nz] (optional) = log W, log(rho h W^2 / (rho W)).  Zone z
 * depends only on (seed, z). */
asx_status asx_zone_sample(int64_t nz, uint64_t seed, double gamma, double seed_rel,
                           double *prim, double *g, double *seed_out, double *target,
                           void *stream);
/* Valencia prim -> cons of independent zones (Con2PrimFactory/src/cons.hxx:
 * 49-91, W from v): prim[8][nz] as above, cons[8][nz] = dens, S_i, tau, dB^i
 * (densitized, the asx_c2p_batch layout); cold_seed[5][nz] (optional) = the
 * cold-start seed rho = D, eps = tau/D, v = 0. */
asx_status asx_prim2con_batch(int64_t nz, const double *prim, const double *g, double *cons,
                              double *cold_seed, const asx_params *par, void *stream);

/* The network: 5 inputs (log D, log S^2, log tau, log B^2 in units of
 * E = D + tau, and cos(S,B); the ideal-gas con2prim is invariant under
 * D, S, tau -> l D, l S, l tau, B -> sqrt(l) B, so the guess needs only these
 * ratios) -> 64 ReLU -> 64 ReLU -> 2 outputs (log W, log(Z/D),
 * Z = rho h W^2), on a
 * float32 blob of ASX_MLP_NBLOB values: W1[64][5] b1[64] W2[64][64] b2[64]
 * W3[2][64] b3[2] (ASX_MLP_NPARAM parameters), then the input means and
 * inverse standard deviations (5 + 5) and output means and standard
 * deviations (2 + 2).  Products take bf16-rounded operands with float32
 * accumulation; the 64x64 layer runs on the matrix cores. */
#define ASX_MLP_NPARAM 4674
#define ASX_MLP_NBLOB 4688
/* x[5][nz] (float32): the raw inputs (before normalisation). */
asx_status asx_mlp_features(int64_t nz, const double *cons, const double *g, float *x,
                            void *stream);
/* out[2][nz]: the network's outputs in target units (log W, log(Z/D)). */
asx_status asx_mlp_forward(int64_t nz, const double *cons, const double *g, const float *blob,
                           double *out, void *stream);
/* One training step on a batch of nz zones: gradient of the mean squared
 * error of the normalised outputs against target[2][nz], then an Adam step
 * (beta1 0.9, beta2 0.999, eps 1e-8, bias-corrected with `step` >= 1) on the
 * parameters in blob with moments m, v (ASX_MLP_NPARAM floats each).
 * lr = 0: gradient only (m, v untouched).  partial: caller scratch of
 * ceil(nz/64) * (ASX_MLP_NPARAM + 1) floats.  loss (1 float) and grad
 * (ASX_MLP_NPARAM floats) are optional device outputs.  Deterministic. */
asx_status asx_mlp_train_step(int64_t nz, const double *cons, const double *g,
                              const double *target, float *blob, float *m, float *v, float lr,
                              int64_t step, float *partial, float *loss, float *grad,
                              void *stream);
/* asx_c2p_batch with the network's guess: Noble (c2p_prime must be Noble)
 * from the network's (Z, v^2) for at most kguess iterations; a zone not
 * converged by then runs the full chain from the plain seed, as
 * asx_c2p_batch would (fallback[z] = 1, optional).  iters counts both.
 * nz < 2^31.  Variant 0 of ASX_KC_MLP (default) evaluates the network on
 * the matrix cores inside the primary-solver kernel, so the guesses never
 * reach memory; variant 1 runs the forward pass as a kernel of its own and
 * keeps the guesses in library scratch.  The queue of unconverged zones
 * lives in library scratch per device and stream, grown on demand. */
asx_status asx_c2p_batch_mlp(int64_t nz, const double *cons, const double *g,
                             const double *seed, const float *blob, int32_t kguess,
                             double *prim, double *flag, int32_t *iters, int32_t *fallback,
                             const asx_params *par, void *stream);

/* ------------------------------------------------------------------------ */
/* Independent zones with composition, any 3-parameter EOS (evolution_eos  */
/* IdealGas or Tabulated3d with eos_table)                                  */
/* ------------------------------------------------------------------------ */
/* The EOSX interface on n points (eos_3p.hxx; eos_3p_tabulated3d.hxx:
 * 186-336 for a table).  From (rho, T, Ye): out[4][n] = press, eps, c_s,
 * entropy (press_/eps_/csnd_/entropy_from_valid_rho_temp_ye).  From (rho, eps, Ye):
 * out[3][n] = T, press and eps after the EOS's clamping of eps
 * (temp_from_valid_rho_eps_ye, then press_from_valid_rho_eps_ye). */
asx_status asx_eos_from_rho_temp(int64_t n, const double *rho, const double *temp,
                                 const double *ye, double *out, const asx_params *par,
                                 void *stream);
asx_status asx_eos_from_rho_eps(int64_t n, const double *rho, const double *eps,
                                const double *ye, double *out, const asx_params *par,
                                void *stream);
/* Valencia prim -> cons with composition (Con2PrimFactory/src/cons.hxx:
 * 49-91): prim[9][nz] = rho, T, Ye, v^i, B^i (undensitized); eps and press
 * from the EOS at (rho, T, Ye); cons[9][nz] = dens, S_i, tau, dB^i, DYe;
 * thermo[2][nz] (optional) = eps, press. */
asx_status asx_prim2con_batch_ye(int64_t nz, const double *prim, const double *g, double *cons,
                                 double *thermo, const asx_params *par, void *stream);
/* asx_c2p_batch with composition: cons[9][nz] (asx_c2p_batch rows, then
 * DYe), seed[7][nz] = saved rho, eps, v^i, Ye and the temperature;
 * prim[9][nz] = the asx_c2p_batch rows, then Ye.  With a table: Palenzuela
 * (c2p_prime), use_temperature = yes and use_press_atmo = no, as the
 * reference's table requires (eps_from_valid_rho_press_ye is refused there).
 * Two passes as asx_con2prim (primary solver, then the queue of the zones
 * it failed through the full chain); nz < 2^31. */
asx_status asx_c2p_batch_ye(int64_t nz, const double *cons, const double *g,
                            const double *seed, double *prim, double *flag, int32_t *iters,
                            const asx_params *par, void *stream);

/* The learned con2prim guess on a table (no counterpart in the reference:
 * the network around the Palenzuela solver of
 * Con2PrimFactory/src/c2p_1DPalenzuela.hxx:306-556).  A table sets a density
 * scale, so the network takes the 5 inputs of asx_mlp_features, log D
 * (undensitized) and Ye = DYe/D -> 64 tanh -> 64 tanh -> 2 outputs (log W,
 * log x, x = h W: Palenzuela's unknown), on a float32 blob of
 * ASX_MLP_YE_NBLOB values laid out as for ASX_MLP_NBLOB with W1[64][7] and
 * 7 input statistics.  cons are the asx_c2p_batch_ye rows (9). */
#define ASX_MLP_YE_NPARAM 4802
#define ASX_MLP_YE_NBLOB 4820
/* Zones with composition drawn on the device for any 3-parameter EOS:
 * box[7] = rho lo, hi; T lo, hi (log-uniform); Ye lo, hi (uniform); W max
 * (W uniform in [1, W max] along a random direction); b^2/rho log-uniform in
 * [1e-4, 1e2] along a random direction; a symmetric 3-metric within 10% of
 * the identity.  prim[9][nz] (asx_prim2con_batch_ye layout), g[6][nz];
 * seed_out[7][nz] (optional, the asx_c2p_batch_ye seed) = rho, eps, v^i,
 * Ye, T with rho, eps, v^i and T times (1 + seed_rel u), u ~ U[-1, 1];
 * target[2][nz] (optional) = log W, log(h W) from the EOS.  Zone z depends
 * only on (seed, z). */
asx_status asx_zone_sample_ye(int64_t nz, uint64_t seed, const double *box, double seed_rel,
                              double *prim, double *g, double *seed_out, double *target,
                              const asx_params *par, void *stream);
/* x[7][nz] (float32): the table network's raw inputs. */
asx_status asx_mlp_ye_features(int64_t nz, const double *cons, const double *g, float *x,
                               void *stream);
/* out[2][nz]: the table network's outputs (log W, log x). */
asx_status asx_mlp_ye_forward(int64_t nz, const double *cons, const double *g,
                              const float *blob, double *out, void *stream);
/* asx_mlp_train_step for the table network (ASX_MLP_YE_NPARAM parameters;
 * partial: ceil(nz/64) * (ASX_MLP_YE_NPARAM + 1) floats). */
asx_status asx_mlp_ye_train_step(int64_t nz, const double *cons, const double *g,
                                 const double *target, float *blob, float *m, float *v,
                                 float lr, int64_t step, float *partial, float *loss,
                                 float *grad, void *stream);
/* asx_c2p_batch_ye on a table with the table network's guess: the
 * Palenzuela primary brackets x in [x_g (1 - bracket), x_g (1 + bracket)]
 * (inside the widest bracket the reference may take) when the residual
 * changes sign there, and takes the reference's bracket otherwise; the
 * root, hence the result, is the same to the solver tolerance.
 * 0 < bracket < 1.  The guesses and the queue live in library scratch per
 * device and stream.  An ideal gas returns ASX_ERR_UNSUPPORTED (use
 * asx_c2p_batch_mlp). */
asx_status asx_c2p_batch_ye_mlp(int64_t nz, const double *cons, const double *g,
                                const double *seed, const float *blob, double bracket,
                                double *prim, double *flag, int32_t *iters,
                                const asx_params *par, void *stream);

/* ------------------------------------------------------------------------ */
/* Batched patches (replaces the driver loop that calls the scheduled        */
/* functions once per local box and RK stage: AsterX/schedule.ccl:119-232)  */
/* ------------------------------------------------------------------------ */
/* A plan runs whole SSPRK3 steps over npatch independent patches of any
 * sizes: every kernel class is one launch over all patches, driven by
 * device-side descriptor tables built once here.  Per stage s = 0, 1, 2 (with
 * (c0, c1, c2) = (0, 1, dt), (3/4, 1/4, dt/4), (1/3, 2/3, 2dt/3)):
 *   asx_substep from state buffer b_in = s into b_out = (1, 2, 0)[s] with
 *   u0 = buffer 0, boundary fill of the 11 state fields of b_out,
 *   asx_con2prim_group on b_out, boundary fill of the 12 synced primitives
 *   (rho, vel^i, eps, press, temperature, entropy, Ye, Bvec^i),
 * so a step leaves the new state in buffer 0.  Buffer 0 is the state slots
 * of patches[p]; bufs[2p] and bufs[2p+1] are buffers 1 and 2 of patch p.
 * Requirements per patch are those of asx_substep and asx_con2prim_group;
 * ghost zones are filled with `bc_kind` on every face (there is no
 * inter-patch exchange).  The plan owns the face/aux scratch of all
 * patches and the con2prim queue; it holds the descriptors, not the data. */
typedef struct asx_batch_s asx_batch;
asx_status asx_batch_create(int32_t npatch, const asx_patch *patches, const asx_state *bufs,
                            const asx_params *par, double dt, int32_t bc_kind,
                            asx_batch **plan);
/* One stage (0, 1 or 2) with plain launches on `stream`. */
asx_status asx_batch_stage(asx_batch *plan, int32_t stage, void *stream);
/* nsteps whole steps.  use_graph = 1: the three stages are captured once
 * into a HIP graph (re-captured when a launch-shape variant changes) and the
 * graph is replayed on `stream`; 0: plain launches, one per kernel class
 * and phase; 2: each stage is ONE persistent cooperative launch whose
 * workgroups walk every phase over every patch (faces and aux, update,
 * state fill, con2prim, queued zones, primitive fill) with grid-wide
 * barriers between phases, every workgroup resident for the whole stage;
 * it uses fixed launch shapes (face variant 0, update variant 1) and is
 * built for PPM, WENO-Z and MP5 with Noble as the primary solver and
 * without caller dB_stag (ASX_ERR_UNSUPPORTED otherwise).  All three modes
 * give the same results bit for bit.  Other values: ASX_ERR_PARAM. */
asx_status asx_batch_step(asx_batch *plan, int32_t nsteps, int32_t use_graph, void *stream);
/* Launch shapes: a plan starts with the device's current variants
 * (asx_set_variant) and keeps its own from then on.  asx_batch_autotune
 * runs, for every face and update variant, the stage-0 stencil work
 * (faces, aux, update into buffer 1) once untimed and `reps` times
 * timed, and keeps the fastest variant per class in the plan.  It writes
 * only buffer 1 and the plan's scratch. */
asx_status asx_batch_autotune(asx_batch *plan, int32_t reps, void *stream);
asx_status asx_batch_get_variant(const asx_batch *plan, int32_t kclass, int32_t *v);
/* Interior zones of all patches and device bytes the plan allocated. */
asx_status asx_batch_info(const asx_batch *plan, int64_t *zones, int64_t *device_bytes);
asx_status asx_batch_destroy(asx_batch *plan);

#ifdef __cplusplus
}
#endif

#endif /* ASTERX_H */
