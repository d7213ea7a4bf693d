#!/usr/bin/env python3
"""Throughput of the GRMHD method-of-lines substep on one MI355X.

Workload (default): the curved-spacetime torus, one 192^3 patch over
[-40,40]^3 M with 3 ghost zones, FP64, Kerr-Schild a = 0.9375, Fishbone-
Moncrief disk (r_in 6, r_max 12, Gamma 4/3, K 1e-3, 2% pressure
perturbations, min beta 100), WENO-Z + HLLE, Noble 2D con2prim with
Palenzuela backup, floors rho 1e-8 / p 1e-10.

One "step" is one RK substep of the reference's schedule on the whole patch:
  RHS (fluxes, sources, A/Psi RHS) and the RK combination  (asx_substep)
  SYNC of the state                                         (asx_bc_fill_many)
  B from A + con2prim                                       (asx_con2prim_group)
  SYNC of the primitives the next RHS reads                 (asx_bc_fill_many)
cycling the three SSPRK3 stages (stage 1 reads no previous-stage state).

Prints one JSON line: zone-updates/s (interior zones x substeps / s), the
roofline of the dominant kernel class (the library brackets each launch
with HIP events on the caller's stream, asx_timing_*), and the CPU oracle
timed on the same inputs on the host cores.
"""
import argparse
import ctypes as C
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E, GB/s
FP64_PEAK_TFS = 78.6           # MI355X FP64, TFLOP/s
BF16_PEAK_TFS = 2500.0         # MI355X dense BF16 matrix, TFLOP/s
# Algorithmic FLOPs of the ideal-gas con2prim (Noble, 1e-10): per Newton
# iteration and for the rest of a zone, pinned by the flop-counting oracle
# build (tests/test_c2p_flops.py).
C2P_FLOPS_PER_ITER = 95
C2P_FLOPS_SETUP = 445
# The guess network 5 -> 64 -> 64 -> 2: 2 (5 64 + 64 64 + 64 2) per zone
MLP_FLOPS_FORWARD = 9088

# The primitives the reference SYNCs after con2prim (AsterX/schedule.ccl:78,
# with Bvec): what the next substep may read in the ghost zones.
SYNC_PRIMS = ["rho", "velx", "vely", "velz", "eps", "press", "temperature", "entropy",
              "Ye", "Bvecx", "Bvecy", "Bvecz"]


def parse():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=6)
    ap.add_argument("--n", type=int, default=192, help="cells per axis")
    ap.add_argument("--recon", default="wenoz")
    ap.add_argument("--flux", default="HLLE")
    ap.add_argument("--cpu", choices=["on", "off"], default="on",
                    help="time the CPU oracle on the same inputs")
    ap.add_argument("--verbose", action="store_true")
    ap.add_argument("--tune", choices=["on", "off"], default="on",
                    help="pick launch-shape variants with asx_autotune_substep "
                         "before the warm-up")
    ap.add_argument("--batch", choices=["on", "off"], default="on",
                    help="also time the batched-patch plan: 400 patches with "
                         "edges from {16,24,32,48,64} cut from the torus")
    ap.add_argument("--batch-patches", type=int, default=400)
    ap.add_argument("--batch-steps", type=int, default=10,
                    help="timed SSPRK3 steps (3 substeps each) of the batch")
    ap.add_argument("--c2p-guess", choices=["on", "off"], default="on",
                    help="also time con2prim with the learned guess against "
                         "the plain seed on 2^24 independent zones")
    ap.add_argument("--c2p-zones", type=int, default=1 << 24)
    ap.add_argument("--amr", choices=["on", "off"], default="on",
                    help="prolongation, restriction and regrid boxes on the torus")
    ap.add_argument("--tab", choices=["on", "off"], default="on",
                    help="also run the tabulated-EOS workload (torus on the table and "
                         "con2prim with composition)")
    ap.add_argument("--tab-n", type=int, default=128)
    ap.add_argument("--tab-steps", type=int, default=12)
    ap.add_argument("--tab-c2p-zones", type=int, default=1 << 22)
    ap.add_argument("--tab-bracket", type=float, default=0.03,
                    help="relative half-width of the guided Palenzuela bracket")
    ap.add_argument("--variants", choices=["on", "off"], default="on",
                    help="also time WENO-Z + LxF, MP5 + HLLE and MP5 + LxF on the torus")
    ap.add_argument("--n256", choices=["on", "off"], default="on",
                    help="also time the torus workload on a 256^3 patch")
    ap.add_argument("--plain64", choices=["on", "off"], default="on",
                    help="also time the plain 64^3 Balsara-1 workload (flat metric, PPM)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a device-to-device copy of 1 GiB five times "
                         "(reference kernel for the HBM byte counters)")
    return ap.parse_args()


def copy_calibration():
    """Five copies of a 1 GiB FP64 array: each reads and writes exactly
    2^30 bytes, so FETCH_SIZE / WRITE_SIZE per copy dispatch give the
    counters' scale for known traffic."""
    import torch
    n = 1 << 27
    src = torch.ones(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / 5
    print(json.dumps({"calibration": "copy", "bytes_read": n * 8,
                      "bytes_written": n * 8, "ms": round(dt * 1e3, 4),
                      "GBs": round(2 * n * 8 / dt / 1e9, 1)}), flush=True)


def log(args, *a):
    if args.verbose:
        print(*a, file=sys.stderr, flush=True)


def state_struct(patch, arrays):
    from asterx_amd import abi
    s = abi.asx_state()
    for q, name in enumerate(abi.STATE):
        a = arrays[name]
        ex, ey, _ = abi.extents(patch.n, patch.ng, abi.centering(name))
        s.gf[q].data = patch._ptr(a)
        s.gf[q].sj = ex
        s.gf[q].sk = ex * ey
    return s


SSPRK3 = [(0.0, 1.0, 1.0), (0.75, 0.25, 0.25), (1.0 / 3.0, 2.0 / 3.0, 2.0 / 3.0)]


class Stepper:
    """Runs substeps on one patch with three rotating state buffers."""

    def __init__(self, patch, par, backend, dt, stream_handle):
        import torch
        from asterx_amd import abi
        self.p, self.par, self.b, self.dt = patch, par, backend, dt
        self.stream = stream_handle
        self.bufs = [{n: patch.arr[n] for n in abi.STATE}]
        for _ in range(2):
            self.bufs.append({n: torch.empty_like(patch.arr[n])
                              for n in abi.STATE})
        self.cur = 0
        self.base = 0
        self.stage = 0
        self.L = backend.lib
        self.bc = abi.BC["copy"]
        st = [abi.SLOT[n] for n in abi.STATE]
        pr = [abi.SLOT[n] for n in SYNC_PRIMS]
        self.slots_state = (C.c_int32 * len(st))(*st)
        self.slots_prims = (C.c_int32 * len(pr))(*pr)

    def _bc(self, ps, slots):
        r = self.L.asx_bc_fill_many(C.byref(ps), len(slots), slots, self.bc,
                                    self.stream)
        if r:
            raise RuntimeError("asx_bc_fill_many: %d" % r)

    def substep(self):
        from asterx_amd import abi
        c0, c1, cf = SSPRK3[self.stage]
        if self.stage == 0:
            self.base = self.cur
        out = [i for i in range(3) if i not in (self.cur, self.base)][0]
        p = self.p
        ps = p.struct()
        u0s = state_struct(p, self.bufs[self.base])
        outs = state_struct(p, self.bufs[out])
        r = self.L.asx_substep(C.byref(ps), C.byref(u0s), C.byref(outs), c0, c1,
                               cf * self.dt, C.byref(self.par), self.stream)
        if r:
            raise RuntimeError("asx_substep: %d" % r)
        for n in abi.STATE:
            p.arr[n] = self.bufs[out][n]
        self.cur = out
        ps = p.struct()
        self._bc(ps, self.slots_state)
        r = self.L.asx_con2prim_group(C.byref(ps), C.byref(self.par), None,
                                      self.stream)
        if r:
            raise RuntimeError("asx_con2prim_group: %d" % r)
        self._bc(ps, self.slots_prims)
        self.stage = (self.stage + 1) % 3


def face_bytes(par):
    """Algorithmic HBM bytes per face of one face-flux launch: the 9
    reconstructed cell fields (rho, s, Ye, p or T, two transverse B, the
    velocity-like vector) and the normal staggered B read once, the 10
    vertex metric fields (alpha, beta^i, gamma_ij) read once, the 10 face
    fluxes written (4 more upwind-CT fields with use_uct)."""
    return 8 * (9 + 1 + 10 + 10 + (4 if par.use_uct else 0))


def update_bytes(par, first_stage):
    """Algorithmic bytes per interior zone of the fused update: 21 hydro
    face fluxes (7 variables x 3 directions), 6 induction fluxes for the
    EMF, 7 vertex auxiliaries, 16 vertex ADM fields, 9 cell primitives
    (rho, eps, p, v^i, B^i; 3 more with z/s reconstruction vectors), the 11
    state fields and the 11 previous-stage fields (not in stage 1) read
    once; 11 new state fields written."""
    reads = 21 + 6 + 7 + 16 + 9 + 11 + (0 if first_stage else 11)
    if par.recon_type in (1, 2):
        reads += 3
    return 8 * (reads + 11)


def stencil_bytes_per_zone(par, first_stage):
    """Algorithmic HBM bytes of the RHS + RK combination per interior zone:
    each input grid function read once, each output written once.
    Reads: 11 cell primitives (rho, v^i, eps, p, s, Ye, B^i) [+T], 4 EM
    state (A_i, Psi), 16 vertex ADM, 7 hydro conserved, 11 previous-stage
    state (not for the first stage); writes: 11 new state."""
    reads = 11 + 4 + 16 + 7 + (0 if first_stage else 11)
    if par.reconstruct_with_temperature:
        reads += 1
    if par.recon_type in (1, 2):
        reads += 3
    return 8 * (reads + 11)


def c2p_bytes_per_zone(with_mask):
    """Algorithmic bytes of B-from-A + con2prim per interior zone.
    Reads: 7 conserved, 3 A_i, 7 vertex metric (alpha, gamma_ij), 6 saved
    primitives, temperature; writes: 12 primitives, 7 conserved, 3 dB,
    6 saved primitives, 6 z/s vectors, flag [+ cell mask]."""
    return 8 * (7 + 3 + 7 + 6 + 1 + 12 + 7 + 3 + 6 + 6 + 1 + (1 if with_mask else 0))


def read_traffic(kernel, workload):
    """Measured HBM bytes per launch (FETCH_SIZE + WRITE_SIZE, calibrated)
    from profiles/traffic.txt, lines: kernel workload fetch_B write_B."""
    path = os.path.join(ROOT, "profiles", "traffic.txt")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            t = line.split()
            if len(t) >= 4 and not line.startswith("#") and t[0] == kernel \
                    and t[1] == workload:
                return float(t[2]) + float(t[3])
    return None


def bench_batched(args, src, par, dt, L, sh, stream):
    """Batched uneven patches: `--batch-patches` patches with per-axis edges
    drawn from {16,24,32,48,64}, cut at random offsets from the torus patch
    (3 ghost zones each), whole SSPRK3 steps of the plan replayed as one HIP
    graph, launch shapes from asx_batch_autotune.  zone-updates/s = interior
    zones x 3 substeps x steps / wall time."""
    import torch
    from asterx_amd import abi, batch
    boxes = batch.sample_boxes(src.n, args.batch_patches, seed=2024)
    with torch.cuda.stream(stream):
        patches = [batch.cut_patch(src, n, o) for n, o in boxes]
    torch.cuda.synchronize()
    plan = batch.Plan(L, patches, par, dt)
    zones, dbytes = plan.info()
    plan.autotune(reps=4, stream=sh)
    variants = plan.variants()
    plan.step(1, graph=True, stream=sh)  # records the graph
    torch.cuda.synchronize()
    K = args.batch_steps
    abi.check("asx_timing_enable", L.asx_timing_enable(1))
    t0 = time.perf_counter()
    plan.step(K, graph=True, stream=sh)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    nl, gms = C.c_int64(), C.c_double()
    abi.check("asx_timing_read", L.asx_timing_read(abi.KCLASS["graph"], C.byref(nl),
                                                   C.byref(gms)))
    # the same steps as one persistent cooperative launch per stage (fixed
    # launch shapes: face variant 0, update variant 1)
    persist = None
    try:
        plan.step(1, persistent=True, stream=sh)
        torch.cuda.synchronize()
        abi.check("asx_timing_enable", L.asx_timing_enable(1))
        t0 = time.perf_counter()
        plan.step(K, persistent=True, stream=sh)
        torch.cuda.synchronize()
        pwall = time.perf_counter() - t0
        nlp, pms = C.c_int64(), C.c_double()
        abi.check("asx_timing_read", L.asx_timing_read(abi.KCLASS["persist"], C.byref(nlp),
                                                       C.byref(pms)))
        persist = {"value": round(3 * zones * K / pwall, 1), "unit": "zone-updates/s",
                   "ms_per_rk3_step": round(1e3 * pwall / K, 4),
                   "launch_ms_per_stage": round(pms.value / max(1, nlp.value), 4),
                   "vs_graph": round(wall / pwall, 3)}
    except abi.AsxError as e:
        persist = {"skipped": str(e)}
    # per-kernel split from one step of plain launches
    abi.check("asx_timing_enable", L.asx_timing_enable(1))
    plan.step(1, graph=False, stream=sh)
    torch.cuda.synchronize()
    kt = {}
    for name, kc in abi.KCLASS.items():
        n_, ms = C.c_int64(), C.c_double()
        abi.check("asx_timing_read", L.asx_timing_read(kc, C.byref(n_), C.byref(ms)))
        if n_.value:
            kt[name] = round(ms.value, 4)
    abi.check("asx_timing_enable", L.asx_timing_enable(0))
    for p in patches[:8]:
        for name in abi.CONS:
            if not bool(torch.isfinite(p.arr[name]).all()):
                raise RuntimeError("non-finite %s in the batch" % name)
    edges = sorted({e for n, _ in boxes for e in n})
    out = {"workload": "%d patches, edges per axis from %s, 3 ghosts, cut from the "
                       "%d^3 torus; SSPRK3 steps as one HIP graph" %
                       (len(patches), edges, src.n[0]),
           "zones": zones, "steps": K,
           "value": round(3 * zones * K / wall, 1), "unit": "zone-updates/s",
           "ms_per_rk3_step": round(1e3 * wall / K, 4),
           "graph_ms_per_step": round(gms.value / max(1, nl.value), 4),
           "plain_launch_kernels_ms_per_step": kt,
           "variants": variants, "plan_device_bytes": dbytes,
           "persistent": persist}
    plan.close()
    del patches
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    return out


def bench_c2p_guess(args, L):
    """Con2prim on 2^24 independent zones of the learned-guess workload
    (ideal gas Gamma 5/3, tolerance 1e-10): plain seed against the trained
    network's guess (8 guided iterations, then the plain chain), for the
    cold-start seed and the previous-step seed (truth x (1 + 1e-2 u))."""
    from asterx_amd import abi, mlp
    if not os.path.exists(mlp.DEFAULT_WEIGHTS):
        return {"skipped": "no trained weights at %s" % mlp.DEFAULT_WEIGHTS}
    blob = mlp.load()
    par = abi.params(gl_gamma=5.0 / 3.0, eps_max=1e8, rho_max=1e8, rho_abs_min=1e-12,
                     vw_lim=1e3, B_lim=1e3, c2p_tol=1e-10)
    ev = mlp.evaluate(L, par, 5.0 / 3.0, blob, nz=args.c2p_zones, reps=3, separate=True)
    out = {"zones": args.c2p_zones, "kguess": 8, "tolerance": 1e-10,
           "mlp": "network on the matrix cores inside the solver kernel",
           "mlp_separate": "forward-pass kernel, guesses through memory, then the solver"}
    for seed in ("cold", "perturbed"):
        a, b = ev["%s_plain" % seed], ev["%s_mlp" % seed]
        c = ev["%s_mlp_separate" % seed]
        out[seed] = {"plain_zones_per_s": round(a["zones_per_s"], 1),
                     "mlp_zones_per_s": round(b["zones_per_s"], 1),
                     "speedup": round(b["zones_per_s"] / a["zones_per_s"], 3),
                     "mlp_separate_zones_per_s": round(c["zones_per_s"], 1),
                     "speedup_separate": round(c["zones_per_s"] / a["zones_per_s"], 3),
                     "iters_mean_plain": round(a["iters_mean"], 3),
                     "iters_mean_mlp": round(b["iters_mean"], 3),
                     "fallback_rate": round(b["fallback_rate"], 5),
                     "flags_plain": a["flags"], "flags_mlp": b["flags"]}
    out["guess_abs_err"] = ev["guess_abs_err"]
    # FP64 roofline of the solve: pinned algorithmic FLOPs per zone
    # (C2P_FLOPS_SETUP + C2P_FLOPS_PER_ITER x the mean primary-solver
    # iterations; tests/test_c2p_flops.py) x zones/s, and the HBM fraction
    # of its 228 B per zone (cons 8, metric 6, seed 5 read; 8 primitives and
    # the flag written, 8 B each; 4 B of iterations)
    for seed in ("cold", "perturbed"):
        for mode in ("plain", "mlp"):
            r = ev["%s_%s" % (seed, mode)]
            fl = C2P_FLOPS_SETUP + C2P_FLOPS_PER_ITER * r["iters_mean"]
            tf = r["zones_per_s"] * fl / 1e12
            out[seed]["%s_fp64" % mode] = {
                "flops_per_zone": round(fl, 1), "achieved_tflops": round(tf, 3),
                "peak_tflops": FP64_PEAK_TFS, "frac": round(tf / FP64_PEAK_TFS, 4),
                "hbm_frac": round(r["zones_per_s"] * 228 / 1e9 / HBM_PEAK_GBS, 4)}
    fwd = ev["forward_ms"]
    tf = MLP_FLOPS_FORWARD * args.c2p_zones / (fwd * 1e-3) / 1e12
    out["mlp_forward"] = {"avg_ms": round(fwd, 4), "flops_per_zone": MLP_FLOPS_FORWARD,
                          "achieved_tflops": round(tf, 2), "peak_tflops": BF16_PEAK_TFS,
                          "frac": round(tf / BF16_PEAK_TFS, 4)}
    return out


def table_params(table, **kw):
    """asx_params on the table: Palenzuela on the temperature, atmosphere
    T = 0.01 MeV at rho 1e-14 (code), WENO-Z + HLLE reconstructing T."""
    from asterx_amd import abi
    base = dict(c2p_prime="Palenzuela", c2p_second="None", use_temperature=1,
                use_press_atmo=0, thermal_eos_atmo=1, rho_abs_min=1e-14, t_atmo=0.01,
                Ye_atmo=0.5, c2p_tol=1e-10, max_iter=100, vw_lim=1e3, B_lim=1e3,
                ye_min=0.035, ye_max=0.55, reconstruction_method="wenoz", flux_type="HLLE",
                reconstruct_with_temperature=1)
    base.update(kw)
    return table.attach(abi.params(**base))


def bench_guided_ye(args, L, sh, par, nz, cd, gd, seed, od, fd, plain_s):
    """The same zones through asx_c2p_batch_ye_mlp: the table network's
    guess of x = h W brackets the Palenzuela root (relative half-width
    --tab-bracket); zones whose root lies outside go to the full chain.
    Against the plain solve: throughput, Brent iterations, identical flags,
    largest relative difference of rho and T."""
    import numpy as np
    import torch
    from asterx_amd import abi, mlp_ye
    if not os.path.exists(mlp_ye.DEFAULT_WEIGHTS):
        return {"skipped": "no trained table weights at %s" % mlp_ye.DEFAULT_WEIGHTS}
    bt = torch.from_numpy(mlp_ye.load()).cuda()
    og = torch.empty_like(od)
    fg = torch.empty_like(fd)
    ig = torch.empty(nz, dtype=torch.int32, device="cuda")
    ptr = lambda t: C.c_void_p(t.data_ptr())
    a = (nz, ptr(cd), ptr(gd), ptr(seed), ptr(bt), args.tab_bracket, ptr(og), ptr(fg), ptr(ig),
         C.byref(par), sh)
    abi.check("asx_c2p_batch_ye_mlp", L.asx_c2p_batch_ye_mlp(*a))
    torch.cuda.synchronize()
    reps = 5
    t0 = time.perf_counter()
    for _ in range(reps):
        abi.check("asx_c2p_batch_ye_mlp", L.asx_c2p_batch_ye_mlp(*a))
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) / reps
    ok = (fd == 1) & (fg == 1)
    d = lambda r: float(((og[r] - od[r]).abs() / od[r].abs())[ok].max())
    return {"bracket": args.tab_bracket, "zones_per_s": round(nz / wall, 1),
            "ms_per_batch": round(1e3 * wall, 4), "speedup": round(plain_s / wall, 3),
            "iters_mean": round(float(ig[fg == 1].double().mean()), 3),
            "flags_equal": bool(torch.equal(fg, fd)),
            "max_rel_diff_rho": d(0), "max_rel_diff_T": d(7)}


def bench_tabulated(args, L, sh, stream):
    """The tabulated-EOS workload: the 234x180x50 synthetic table; the
    torus geometry at `--tab-n`^3 on the table (SSPRK3 substeps, as the
    headline), and Palenzuela con2prim with composition on 2^22 independent
    zones (rho 1e-12..1e-4 code, T 0.05..80 MeV, Ye 0.06..0.5, W up to 10,
    b^2/rho 1e-4..1e2, previous-step seed = truth x (1 + 1e-2 u))."""
    import numpy as np
    import torch
    from asterx_amd import abi, tabeos, torus, zones
    from asterx_amd.patch import Patch
    host_tab = tabeos.build()
    table = tabeos.Table(host_tab, device="cuda")
    par = table_params(table)
    from asterx_amd import lib
    be = lib.backend(sh)
    n, ng, x0, dx = torus.torus_patch_geometry(n=args.tab_n)
    slots = abi.ADM + abi.PRIMS + abi.CONS + abi.DB + abi.SAVED + abi.ZS + \
        ["con2prim_flag", "aster_mask_cc"] + abi.DBSTAG + abi.AVEC
    with torch.cuda.stream(stream):
        p = Patch(n, ng, x0, dx, slots, device="cuda:0")
        torus.build_torus_tab(p, be, par, table.ranges())
    torch.cuda.synchronize()
    for s in abi.DBSTAG:
        del p.arr[s]
    zn = n[0] * n[1] * n[2]
    dt = 0.2 * dx[0]
    st = Stepper(p, par, be, dt, sh)
    # launch shapes for the table's face kernels on this patch
    variants = tune_variants(be, p, st, par, dt, sh) if args.tune == "on" else {}
    for _ in range(args.warmup):
        st.substep()
    torch.cuda.synchronize()
    K = args.tab_steps
    abi.check("asx_timing_enable", L.asx_timing_enable(1))
    t0 = time.perf_counter()
    for _ in range(K):
        st.substep()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kt = {}
    for name, kc in abi.KCLASS.items():
        nl, ms = C.c_int64(), C.c_double()
        abi.check("asx_timing_read", L.asx_timing_read(kc, C.byref(nl), C.byref(ms)))
        if nl.value:
            kt[name] = round(ms.value / K, 4)
    abi.check("asx_timing_enable", L.asx_timing_enable(0))
    cnt = abi.asx_counters()
    ps = p.struct()
    abi.check("asx_con2prim_group", L.asx_con2prim_group(C.byref(ps), C.byref(par),
                                                         C.byref(cnt), sh))
    torch.cuda.synchronize()
    for name in abi.CONS:
        if not bool(torch.isfinite(p.arr[name][p.interior(name)]).all()):
            raise RuntimeError("non-finite %s on the table" % name)
    patch_out = {"workload": "torus %d^3 +3 ghosts on the table, WENO-Z + HLLE on (rho, T, "
                             "Ye), Palenzuela on T" % args.tab_n,
                 "zones": zn, "steps": K, "ms_per_step": round(1e3 * wall / K, 4),
                 "value": round(zn * K / wall, 1), "unit": "zone-updates/s",
                 "kernels_ms_per_step": kt, "c2p_flags": list(cnt.c2p_flags),
                 "variants": variants}
    del st, p
    torch.cuda.synchronize()
    torch.cuda.empty_cache()

    # con2prim with composition on independent zones
    nz = args.tab_c2p_zones
    prim, g = zones.make_batch_ye(nz, 5, (1e-12, 1e-4), (0.05, 80.0), (0.06, 0.5))
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()
    pd, gd = dev(prim), dev(g)
    del prim
    cd = torch.empty((9, nz), dtype=torch.float64, device="cuda")
    td = torch.empty((2, nz), dtype=torch.float64, device="cuda")
    ptr = lambda t: C.c_void_p(t.data_ptr())
    abi.check("asx_prim2con_batch_ye", L.asx_prim2con_batch_ye(
        nz, ptr(pd), ptr(gd), ptr(cd), ptr(td), C.byref(par), sh))
    torch.cuda.synchronize()
    gen = torch.Generator(device="cuda").manual_seed(6)
    u = lambda: 1 + 1e-2 * (2 * torch.rand(nz, device="cuda", dtype=torch.float64,
                                           generator=gen) - 1)
    seed = torch.stack([pd[0] * u(), td[0] * u(), pd[3] * u(), pd[4] * u(), pd[5] * u(),
                        pd[2], pd[1]]).contiguous()
    od = torch.empty((9, nz), dtype=torch.float64, device="cuda")
    fd = torch.empty(nz, dtype=torch.float64, device="cuda")
    idd = torch.empty(nz, dtype=torch.int32, device="cuda")
    args_c = (nz, ptr(cd), ptr(gd), ptr(seed), ptr(od), ptr(fd), ptr(idd), C.byref(par), sh)
    abi.check("asx_c2p_batch_ye", L.asx_c2p_batch_ye(*args_c))   # warm-up
    torch.cuda.synchronize()
    reps = 5
    t0 = time.perf_counter()
    for _ in range(reps):
        abi.check("asx_c2p_batch_ye", L.asx_c2p_batch_ye(*args_c))
    torch.cuda.synchronize()
    cwall = (time.perf_counter() - t0) / reps
    flags = torch.bincount(fd.to(torch.int64).clamp(0, 6), minlength=7).tolist()
    ok = fd == 1
    rel_rho = float(((od[0] - pd[0]).abs() / pd[0])[ok].max())
    rel_t = float(((od[7] - pd[1]).abs() / pd[1])[ok].max())
    c2p_out = {"zones": nz, "zones_per_s": round(nz / cwall, 1),
               "ms_per_batch": round(1e3 * cwall, 4),
               "iters_mean": round(float(idd[ok].double().mean()), 3), "flags": flags,
               "max_rel_err_rho": rel_rho, "max_rel_err_T": rel_t}
    c2p_out["guided"] = bench_guided_ye(args, L, sh, par, nz, cd, gd, seed, od, fd, cwall)
    log(args, "tabulated patch:", patch_out, "\ntabulated c2p batch:", c2p_out)
    # the CPU oracle on the first 2^15 zones of the same batch
    m = min(nz, 1 << 15)
    hc = np.ascontiguousarray(cd[:, :m].cpu().numpy())
    hg = np.ascontiguousarray(gd[:, :m].cpu().numpy())
    hs = np.ascontiguousarray(seed[:, :m].cpu().numpy())
    htable = tabeos.Table(host_tab, device=None)  # holds the arrays hpar points into
    hpar = table_params(htable)
    import oracle
    ho, hf, hi = np.zeros((9, m)), np.zeros(m), np.zeros(m, dtype=np.int32)
    t0 = time.perf_counter()
    abi.check("orc_c2p_batch_ye", oracle.lib().orc_c2p_batch_ye(
        m, hc.ctypes.data, hg.ctypes.data, hs.ctypes.data, ho.ctypes.data, hf.ctypes.data,
        hi.ctypes.data, C.byref(hpar)))
    c2p_out["cpu_oracle_zones_per_s"] = round(m / (time.perf_counter() - t0), 1)
    c2p_out["cpu_sample_zones"] = m
    c2p_out["flags_equal_to_oracle_on_sample"] = bool(
        np.array_equal(hf, fd[:m].cpu().numpy()))
    del htable, pd, gd, cd, td, seed, od, fd, idd
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    return {"table": "234x180x50, 4 variables (log p, log(eps+shift), c_s^2, entropy)",
            "patch": patch_out, "c2p_batch": c2p_out}


def tmunu_bytes_per_vertex():
    """Algorithmic bytes of AsterX_Tmunu per interior vertex: 9 cell
    primitives (rho, eps, p, B^i, v^i) and 10 vertex metric fields (alpha,
    beta^i, gamma_ij) read once, the 10 eT components read and written."""
    return 8 * (9 + 10 + 10 + 10)


def bench_tmunu(args, p, par, L, sh, workload):
    """AsterX_Tmunu on the evolved torus patch (order-4 interpolation):
    average launch time from HIP events, HBM roofline, and the CPU oracle
    on a 48^3-cell window of the same data."""
    import torch
    from asterx_amd import abi
    from asterx_amd.patch import Patch
    for nm in abi.TMUNU:
        p.alloc(nm)
    ps = p.struct()
    for _ in range(2):
        abi.check("asx_tmunu", L.asx_tmunu(C.byref(ps), C.byref(par), sh))
    torch.cuda.synchronize()
    reps = 20
    abi.check("asx_timing_enable", L.asx_timing_enable(1))
    for _ in range(reps):
        abi.check("asx_tmunu", L.asx_tmunu(C.byref(ps), C.byref(par), sh))
    torch.cuda.synchronize()
    nl, ms = C.c_int64(), C.c_double()
    abi.check("asx_timing_read", L.asx_timing_read(abi.KCLASS["tmunu"], C.byref(nl),
                                                  C.byref(ms)))
    abi.check("asx_timing_enable", L.asx_timing_enable(0))
    avg = ms.value / nl.value
    nv = (p.n[0] + 1) * (p.n[1] + 1) * (p.n[2] + 1)
    per = tmunu_bytes_per_vertex() * nv
    ach = per / (avg * 1e-3) / 1e9
    ett = p.arr["eTtt"][p.interior("eTtt")]
    if not bool(torch.isfinite(ett).all()):
        raise RuntimeError("non-finite eTtt")
    out = {"vertices": nv, "order": int(par.tmunu_interp_order),
           "avg_ms": round(avg, 4), "vertices_per_s": round(nv / (avg * 1e-3), 1),
           "roofline": {"bound": "hbm", "achieved": round(ach, 2), "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4),
                        "traffic": read_traffic("tmunu", workload),
                        "bytes_per_launch": int(per)}}
    for nm in abi.TMUNU:
        del p.arr[nm]
    if args.cpu == "on":
        import oracle
        ng = p.ng
        m = min(48, *p.n)
        hp = Patch((m,) * 3, ng, (0.0,) * 3, p.dx)
        # window origin (first ghost point) per axis, centred in the patch;
        # arrays are indexed (z, y, x)
        o = [(p.n[d] - m) // 2 for d in range(3)]
        for nm in abi.ADM[:10] + ["rho", "eps", "press", "velx", "vely", "velz",
                                  "Bvecx", "Bvecy", "Bvecz", "zvec_x", "zvec_y",
                                  "zvec_z"]:
            e = m + 2 * ng + (abi.centering(nm) & 1)
            a = p.arr[nm][o[2]:o[2] + e, o[1]:o[1] + e, o[0]:o[0] + e]
            hp[nm] = a.cpu().numpy()
        for nm in abi.TMUNU:
            hp.alloc(nm)
        hs = hp.struct()
        ts = []
        for _ in range(3):
            t0 = time.perf_counter()
            abi.check("orc_tmunu", oracle.lib().orc_tmunu(C.byref(hs), C.byref(par)))
            ts.append(time.perf_counter() - t0)
        out["cpu_oracle_vertices_per_s"] = round((m + 1) ** 3 / min(ts), 1)
        out["cpu_sample"] = "oracle (OpenMP) on a %d^3-cell window of the torus, best of 3" % m
    return out


def bench_regrid(p, par, L, sh):
    """AsterX_EstimateError (first gradient of rho, the reference's default)
    on the evolved torus: average launch time from HIP events and the HBM
    roofline at 16 B per cell (rho read once, regrid_error written)."""
    import torch
    from asterx_amd import abi
    p.alloc("regrid_error")
    ps = p.struct()
    slots = (C.c_int32 * 1)(*abi.regrid_slots("HydroBaseX::rho"))
    for _ in range(2):
        abi.check("asx_estimate_error", L.asx_estimate_error(C.byref(ps), C.byref(par), 1,
                                                             slots, sh))
    torch.cuda.synchronize()
    abi.check("asx_timing_enable", L.asx_timing_enable(1))
    for _ in range(20):
        abi.check("asx_estimate_error", L.asx_estimate_error(C.byref(ps), C.byref(par), 1,
                                                             slots, sh))
    torch.cuda.synchronize()
    nl, ms = C.c_int64(), C.c_double()
    abi.check("asx_timing_read", L.asx_timing_read(abi.KCLASS["regrid"], C.byref(nl),
                                                  C.byref(ms)))
    abi.check("asx_timing_enable", L.asx_timing_enable(0))
    del p.arr["regrid_error"]
    avg = ms.value / nl.value
    cells = p.n[0] * p.n[1] * p.n[2]
    ach = 16 * cells / (avg * 1e-3) / 1e9
    return {"method": "first gradient", "vars": 1, "avg_ms": round(avg, 4),
            "achieved_GBs": round(ach, 1), "frac": round(ach / HBM_PEAK_GBS, 4)}


def bench_amr(p, par, L, sh, reps=10):
    """Refinement by 2 on the evolved torus: a fine patch of 2m^3 cells
    over the central m^3 coarse cells (m = n/2, 3 ghosts), the 11 evolved
    grid functions (cell, edge and vertex centred).  Prolongation of order 3
    (every fine point: a new patch; ghost points only: the coarse-fine
    boundary fill of a substep), restriction of the fine patch back, and the
    regrid box table from the first gradient of rho.  Algorithmic bytes:
    prolongation 8 B per fine point written plus the coarse points of the
    region read once (1/8 per fine point); restriction 8 B per coarse point
    written plus its fine points read once.  HIP events on the library's
    stream; the box table (device tags, host clustering) by wall clock."""
    import numpy as np
    import torch
    from asterx_amd import abi
    from asterx_amd.patch import Patch
    m = (p.n[0] // 4) * 2
    lo = [(p.n[d] - m) // 2 for d in range(3)]
    fdx = tuple(d / 2 for d in p.dx)
    names = list(abi.STATE)
    with torch.cuda.device(0):
        f = Patch((2 * m,) * 3, p.ng, tuple(p.x0[d] + lo[d] * p.dx[d] for d in range(3)), fdx,
                  slots=names, device="cuda:0")
    slots = (C.c_int32 * len(names))(*[abi.SLOT[n] for n in names])
    ps, fs = p.struct(), f.struct()

    def timed(fn):
        for _ in range(2):
            abi.check("amr", fn())
        torch.cuda.synchronize()
        abi.check("asx_timing_enable", L.asx_timing_enable(1))
        for _ in range(reps):
            abi.check("amr", fn())
        torch.cuda.synchronize()
        nl, ms = C.c_int64(), C.c_double()
        abi.check("asx_timing_read", L.asx_timing_read(abi.KCLASS["amr"], C.byref(nl),
                                                      C.byref(ms)))
        abi.check("asx_timing_enable", L.asx_timing_enable(0))
        return ms.value / nl.value

    def points(n, ng, name, own):
        e = abi.extents(n, ng, abi.centering(name))
        if own:
            return (e[0] - 2 * ng) * (e[1] - 2 * ng) * (e[2] - 2 * ng)
        return e[0] * e[1] * e[2]

    out = {"workload": "fine %d^3 over the central %d^3 cells of the %d^3 torus, 11 evolved "
                       "grid functions" % (2 * m, m, p.n[0])}
    for region, key in ((0, "prolongate_patch"), (1, "prolongate_ghosts")):
        ms = timed(lambda: L.asx_prolongate(C.byref(ps), C.byref(fs), len(names), slots, 3,
                                            region, sh))
        pts = sum(points(f.n, f.ng, n, False) - (points(f.n, f.ng, n, True) if region else 0)
                  for n in names)
        ach = 9 * pts / (ms * 1e-3) / 1e9
        out[key] = {"order": 3, "fine_points": pts, "avg_ms": round(ms, 4),
                    "achieved_GBs": round(ach, 1), "frac": round(ach / HBM_PEAK_GBS, 4)}
    ms = timed(lambda: L.asx_restrict(C.byref(fs), C.byref(ps), len(names), slots, sh))
    cpts, byt = 0, 0
    for n in names:
        c = abi.centering(n)
        cp = 1
        for d in range(3):
            cp *= m + ((c >> d) & 1)
        kids = 1
        for d in range(3):
            kids *= 1 if (c >> d) & 1 else 2
        cpts += cp
        byt += 8 * cp * (1 + kids)
    ach = byt / (ms * 1e-3) / 1e9
    out["restrict"] = {"coarse_points": cpts, "avg_ms": round(ms, 4),
                       "achieved_GBs": round(ach, 1), "frac": round(ach / HBM_PEAK_GBS, 4)}
    # box table from the first gradient of rho
    p.alloc("regrid_error")
    ps = p.struct()
    rs = (C.c_int32 * 1)(*abi.regrid_slots("HydroBaseX::rho"))
    abi.check("asx_estimate_error", L.asx_estimate_error(C.byref(ps), C.byref(par), 1, rs, sh))
    torch.cuda.synchronize()
    err = p.numpy("regrid_error")[p.ng:-p.ng, p.ng:-p.ng, p.ng:-p.ng]
    thr = float(np.quantile(err, 0.9))
    bf, mgs = (C.c_int32 * 3)(8, 8, 8), (C.c_int32 * 3)(64, 64, 64)
    cap = 1 << 16
    boxes = (abi.asx_box * cap)()
    nb = C.c_int32()
    t = []
    for _ in range(6):
        t0 = time.perf_counter()
        abi.check("asx_regrid_boxes", L.asx_regrid_boxes(C.byref(ps), thr, 1, bf, mgs, 0.7,
                                                        cap, boxes, C.byref(nb), sh))
        t.append(time.perf_counter() - t0)
    fine_cells = 0
    for q in range(nb.value):
        b = boxes[q]
        fine_cells += (b.hi[0] - b.lo[0] + 1) * (b.hi[1] - b.lo[1] + 1) * (b.hi[2] - b.lo[2] + 1)
    tagged = float((err > thr).mean())
    out["regrid_boxes"] = {"threshold": "90th percentile of the first gradient of rho",
                           "n_error_buf": 1, "blocking_factor": 8, "max_grid_size": 64,
                           "grid_eff": 0.7, "boxes": nb.value, "tagged_frac": round(tagged, 4),
                           "covered_frac": round(fine_cells / (8.0 * err.size), 4),
                           "median_ms": round(1e3 * float(np.median(t[1:])), 2)}
    del p.arr["regrid_error"]
    return out


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def oracle_substep_once(L, p, out, par, dt):
    """One stage-1 substep of the reference's schedule on the CPU oracle:
    RHS + forward-Euler combination, SYNC of the state, B from A,
    con2prim, SYNC of the synced primitives."""
    from asterx_amd import abi
    bc = abi.BC["copy"]
    ps = p.struct()
    os_ = out.struct()
    abi.check("orc_substep", L.orc_substep(C.byref(ps), None, C.byref(os_),
                                          0.0, 1.0, dt, C.byref(par)))
    for n in abi.STATE:
        p.arr[n], out.arr[n] = out.arr[n], p.arr[n]
    ps = p.struct()
    for n in abi.STATE:
        abi.check("orc_bc_fill", L.orc_bc_fill(C.byref(ps), abi.SLOT[n], bc))
    abi.check("orc_bstag_from_A", L.orc_bstag_from_A(C.byref(ps)))
    abi.check("orc_dB_from_bstag", L.orc_dB_from_bstag(C.byref(ps)))
    abi.check("orc_con2prim", L.orc_con2prim(C.byref(ps), C.byref(par)))
    for n in SYNC_PRIMS:
        abi.check("orc_bc_fill", L.orc_bc_fill(C.byref(ps), abi.SLOT[n], bc))


def time_oracle(host_patch, par, dt, threads, reps=5):
    """Median of `reps` stage-1 substeps (after one untimed) of the CPU oracle
    on host_patch with `threads` OpenMP threads; seconds."""
    import numpy as np
    import oracle
    from asterx_amd.patch import StateBuf
    L = oracle.lib()
    old = L.orc_set_threads(threads)
    out = StateBuf(host_patch)
    ts = []
    try:
        for r in range(reps + 1):
            t0 = time.perf_counter()
            oracle_substep_once(L, host_patch, out, par, dt)
            if r:
                ts.append(time.perf_counter() - t0)
            assert np.isfinite(host_patch.arr["dens"]).all()
    finally:
        L.orc_set_threads(old)
    return sorted(ts)[len(ts) // 2]


def host_window(host, m):
    """An m^3-cell window (with ghosts) at the centre of a host patch, every
    slot of the oracle's schedule allocated."""
    from asterx_amd import batch
    from asterx_amd.driver import all_slots
    o = tuple((host.n[d] - m) // 2 for d in range(3))
    names = [s for s in all_slots() if s in host.arr]
    w = batch.cut_patch(host, (m, m, m), o, names)
    for s in all_slots():
        if s not in w:
            w.alloc(s)
    return w


def cpu_baseline(host, par, dt, label, window_1core=64):
    """The CPU oracle (g++ -O3 -march=x86-64-v3, OpenMP over k-planes and
    zones, strict IEEE) on the same inputs: one stage-1 substep incl. SYNC,
    B from A and con2prim, median of 5 after one warm-up, on all host cores
    (the whole patch) and on one core (a window of window_1core^3 cells)."""
    import oracle
    cores = oracle.lib().orc_set_threads(0)
    zones = host.n[0] * host.n[1] * host.n[2]
    t_all = time_oracle(host, par, dt, cores)
    m = min(window_1core, *host.n)
    w = host_window(host, m)
    t_one = time_oracle(w, par, dt, 1)
    return {"value": round(zones / t_all, 1), "unit": "zone-updates/s", "cores": cores,
            "kind": "port",
            "one_core_value": round(m ** 3 / t_one, 1),
            "cpu_model": cpu_model(),
            "sample": "CPU oracle (g++ -O3 -march=x86-64-v3 -fopenmp, strict IEEE) on the "
                      "same %s inputs: one stage-1 substep incl. SYNC, B from A and con2prim; "
                      "median of 5 after a warm-up: %d threads on the whole patch (%.2f s "
                      "each), 1 thread on a %d^3 window (%.2f s each)" %
                      (label, cores, t_all, m, t_one)}


def substep_bytes_per_zone(par, first_stage):
    """Algorithmic HBM bytes of a whole substep per interior zone: the
    stencil half (stencil_bytes_per_zone) plus con2prim's saved-primitive
    reads (6) and its writes of 12 primitives, 6 saved primitives, 3 z, 3 s,
    the flag and the cell mask (32 fields): 736 B, 648 B in a first stage."""
    return stencil_bytes_per_zone(par, first_stage) + 8 * 32


SUBSTEP_KERNELS = ("bstag", "face_x", "face_y", "face_z", "aux", "update", "c2p",
                   "c2p_rest", "bc")


def read_kernel_times(L, K):
    from asterx_amd import abi
    kt = {}
    for name, kc in abi.KCLASS.items():
        nl, ms = C.c_int64(), C.c_double()
        abi.check("asx_timing_read", L.asx_timing_read(kc, C.byref(nl), C.byref(ms)))
        if nl.value:
            kt[name] = (nl.value, ms.value)
    return kt


def build_torus_patch(n, recon, flux, stream, be):
    import torch
    from asterx_amd import abi, torus
    from asterx_amd.patch import Patch
    nn, ng, x0, dx = torus.torus_patch_geometry(n=n)
    spec = torus.TorusSpec()
    par = torus.eos_params(spec, reconstruction_method=recon, flux_type=flux)
    slots = abi.ADM + abi.PRIMS + abi.CONS + abi.DB + abi.SAVED + abi.ZS + \
        ["con2prim_flag", "aster_mask_cc"] + abi.DBSTAG + abi.AVEC
    with torch.cuda.stream(stream):
        p = Patch(nn, ng, x0, dx, slots, device="cuda:0")
        torus.build_torus(p, be, par, spec)
    torch.cuda.synchronize()
    # the fused substep keeps its face data in library scratch; con2prim
    # derives dB from A directly, so the staggered field is not kept.
    for s in abi.DBSTAG:
        del p.arr[s]
    return p, par, 0.2 * dx[0]


def tune_variants(be, p, st, par, dt, sh):
    """Launch shapes from asx_autotune_substep (SSPRK3 stage-2 coefficients
    into a scratch state; setup, untimed)."""
    import torch
    from asterx_amd import abi
    from asterx_amd.patch import StateBuf
    scratch = StateBuf(p)
    ps, u0s, outs = p.struct(), state_struct(p, st.bufs[0]), scratch.struct()
    # one untimed substep first (clocks and caches as in the timed run), then
    # ten substeps per variant: the faces' variants differ by a few percent
    abi.check("asx_substep", be.lib.asx_substep(
        C.byref(ps), C.byref(u0s), C.byref(outs), 0.75, 0.25, 0.25 * dt, C.byref(par), sh))
    abi.check("asx_autotune_substep", be.lib.asx_autotune_substep(
        C.byref(ps), C.byref(u0s), C.byref(outs), 0.75, 0.25, 0.25 * dt,
        C.byref(par), 10, sh))
    torch.cuda.synchronize()
    del scratch
    variants = {}
    for name in ("face_x", "face_y", "face_z", "update"):
        v = C.c_int32()
        abi.check("asx_get_variant", be.lib.asx_get_variant(abi.KCLASS[name], C.byref(v)))
        variants[name] = v.value
    return variants


def time_substeps(st, L, K, warmup):
    """warmup untimed substeps, then K timed with per-kernel HIP events:
    (ms per substep of wall time, kernel times, first stages among the K)."""
    import torch
    from asterx_amd import abi
    for _ in range(warmup):
        st.substep()
    torch.cuda.synchronize()
    first = sum(1 for k in range(K) if (k + st.stage) % 3 == 0)
    abi.check("asx_timing_enable", L.asx_timing_enable(1))
    t0 = time.perf_counter()
    for _ in range(K):
        st.substep()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kt = read_kernel_times(L, K)
    abi.check("asx_timing_enable", L.asx_timing_enable(0))
    return 1e3 * wall / K, kt, first


def time_graph(st, stream, K):
    """The same substeps replayed from a HIP graph (stream capture of the
    library calls): six substeps per graph, two SSPRK3 steps after which the
    three state buffers are back in their places; ceil(K / 6) timed replays
    after one untimed.  Wall ms per substep; the launches of one graph cost
    one host call, where the substep-by-substep form pays ~20 launches."""
    import torch
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=stream):
        for _ in range(6):
            st.substep()
    g.replay()
    torch.cuda.synchronize()
    nrep = max(1, (K + 5) // 6)
    t0 = time.perf_counter()
    for _ in range(nrep):
        g.replay()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    del g
    return 1e3 * wall / (6 * nrep)


def graph_timing(st, stream, K):
    """(ms per substep from HIP-graph replays, how it was measured), or
    (None, why not) -- then the substep-by-substep time stands."""
    if K % 6:
        return None, "substep by substep (%d substeps is not a whole number of graphs)" % K
    try:
        return time_graph(st, stream, K), "HIP graph of 6 substeps, %d timed replays" % (K // 6)
    except Exception as e:  # the launch-by-launch time is still a valid measurement
        return None, "substep by substep (graph capture failed: %s)" % e


def substep_roofline(par, zones, K, first, kt, ms_step, workload):
    """HBM roofline of the whole substep: its algorithmic bytes (736 B per
    zone, 648 in a first stage, averaged over the timed stages) over the GPU
    time of its launches per substep (HIP events), beside the same bytes
    over the wall time per substep."""
    sub_b = (first * substep_bytes_per_zone(par, True) +
             (K - first) * substep_bytes_per_zone(par, False)) / K
    gpu_ms = sum(kt[k][1] for k in SUBSTEP_KERNELS if k in kt) / K
    per = sub_b * zones
    ach = per / (gpu_ms * 1e-3) / 1e9
    traffic = None
    parts = [read_traffic(k, workload) for k in ("face_x", "face_y", "face_z", "aux",
                                                 "update", "c2p") if k in kt]
    if all(t is not None for t in parts):
        traffic = sum(parts)
    return {"bound": "hbm", "scope": "whole substep: RHS + RK combination, SYNC, B from A "
                                     "+ con2prim, SYNC",
            "achieved": round(ach, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": traffic,
            "bytes_per_substep": int(per), "bytes_per_zone": round(sub_b, 1),
            "gpu_ms_per_substep": round(gpu_ms, 4),
            "wall_frac": round(per / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}


def dominant_kernel(par, n, zones, K, first, kt, workload):
    """The kernel class with the largest summed time, its algorithmic bytes
    per launch (each grid function it reads or writes once per point) over
    its average HIP-event duration."""
    nf = [(n[d] + 1) * (n[(d + 1) % 3] + 2) * (n[(d + 2) % 3] + 2) for d in range(3)]
    upd_b = (first * update_bytes(par, True) + (K - first) * update_bytes(par, False)) / K
    per_launch = {"face_x": face_bytes(par) * nf[0], "face_y": face_bytes(par) * nf[1],
                  "face_z": face_bytes(par) * nf[2], "update": upd_b * zones,
                  "c2p": c2p_bytes_per_zone(True) * zones}
    dom = max((k for k in kt if k in per_launch), key=lambda k: kt[k][1])
    nl, ms = kt[dom]
    avg = ms / nl
    ach = per_launch[dom] / (avg * 1e-3) / 1e9
    return {"kernel": dom, "avg_ms": round(avg, 4), "bytes_per_launch": int(per_launch[dom]),
            "achieved": round(ach, 2), "frac": round(ach / HBM_PEAK_GBS, 4),
            "traffic": read_traffic(dom, workload),
            "bytes_definition": "face: 9 reconstructed cell fields + normal B + 10 vertex "
                                "metric fields read, 10 fluxes written per face; update: "
                                "27 face fluxes, 7 aux, 16 ADM, 9 prims, 11 (+11) state "
                                "read, 11 written per zone; c2p: 60 fields per zone"}


def bench_torus_variant(args, n, recon, flux, steps, warmup, tune, sh, stream, be):
    """The torus workload at n^3 with one reconstruction / Riemann solver:
    zone-updates/s and the substep roofline (the patch is freed after)."""
    import torch
    p, par, dt = build_torus_patch(n, recon, flux, stream, be)
    st = Stepper(p, par, be, dt, sh)
    if tune:
        tune_variants(be, p, st, par, dt, sh)
    ms_eager, kt, first = time_substeps(st, be.lib, steps, warmup)
    ms_graph, launch = graph_timing(st, stream, steps)
    ms_step = ms_graph if ms_graph is not None else ms_eager
    zones = n ** 3
    out = {"recon": recon, "flux": flux, "n": n, "steps": steps,
           "value": round(zones / (ms_step * 1e-3), 1), "unit": "zone-updates/s",
           "ms_per_step": round(ms_step, 4), "eager_ms_per_step": round(ms_eager, 4),
           "launch": launch,
           "roofline": substep_roofline(par, zones, steps, first, kt, ms_step,
                                        "torus%d_%s_%s" % (n, recon, flux)),
           "kernels_ms_per_step": {k: round(v[1] / steps, 4) for k, v in kt.items()}}
    del st, p
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    return out


def bench_plain64(args, sh, stream, be):
    """The plain small workload: one 64^3 patch with 3 ghosts over
    [-0.5, 0.5]^3, flat metric, ideal gas Gamma 5/3, PPM + HLLE, Noble with
    Palenzuela backup, the Balsara-1 shock tube along x with every zone's
    rho and p multiplied by (1 + 1e-3 u), u ~ U[-1, 1] seeded; B from a
    staggered A.  Substeps cycled through the SSPRK3 stages at dt = 0.25 dx;
    the CPU oracle on the same inputs beside it."""
    import torch
    from asterx_amd import abi, seeds
    from asterx_amd.driver import Evolution, all_slots, check_prims_ideal
    from asterx_amd.patch import Patch
    n, ng = 64, 3
    dx = 1.0 / n
    par = abi.params(reconstruction_method="ppm", flux_type="HLLE", gl_gamma=5.0 / 3.0,
                     c2p_prime="Noble", c2p_second="Palenzuela", eps_max=1e8, rho_max=1e8,
                     rho_abs_min=1e-8)
    slots = abi.ADM + abi.PRIMS + abi.CONS + abi.DB + abi.SAVED + abi.ZS + \
        ["con2prim_flag", "aster_mask_cc"] + abi.DBSTAG + abi.AVEC
    with torch.cuda.stream(stream):
        p = Patch((n, n, n), ng, (-0.5, -0.5, -0.5), (dx, dx, dx), slots, device="cuda:0")
        seeds.minkowski(p)
        seeds.balsara(p, "Balsara1", "x", gamma=5.0 / 3.0, perturb=1e-3, seed=100)
        seeds.post_initial(p, 5.0 / 3.0)
        Evolution(be, p, par, bc="copy").initial(check_prims_ideal)
    torch.cuda.synchronize()
    for s in abi.DBSTAG:
        del p.arr[s]
    host = None
    if args.cpu == "on":
        host = p.to(None)
        for s in all_slots():
            if s not in host:
                host.alloc(s)
    dt = 0.25 * dx
    st = Stepper(p, par, be, dt, sh)
    # launch shapes for this patch size (the ones tuned on the larger
    # patches are not the best at 64^3)
    variants = tune_variants(be, p, st, par, dt, sh) if args.tune == "on" else {}
    K = 30
    ms_eager, kt, first = time_substeps(st, be.lib, K, 6)
    ms_graph, launch = graph_timing(st, stream, K)
    ms_step = ms_graph if ms_graph is not None else ms_eager
    zones = n ** 3
    out = {"workload": "Balsara-1 along x, 64^3 +3 ghosts, flat metric, rho and p x (1 + "
                       "1e-3 u), PPM + HLLE, Noble/Palenzuela, Gamma 5/3",
           "zones": zones, "steps": K, "value": round(zones / (ms_step * 1e-3), 1),
           "unit": "zone-updates/s", "ms_per_step": round(ms_step, 4),
           "eager_ms_per_step": round(ms_eager, 4), "launch": launch,
           "roofline": substep_roofline(par, zones, K, first, kt, ms_step, "plain64"),
           "kernels_ms_per_step": {k: round(v[1] / K, 4) for k, v in kt.items()},
           "variants": variants}
    for name in abi.CONS:
        if not bool(torch.isfinite(p.arr[name][p.interior(name)]).all()):
            raise RuntimeError("non-finite %s in the 64^3 run" % name)
    del st, p
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    if host is not None:
        import oracle
        cores = oracle.lib().orc_set_threads(0)
        t_all = time_oracle(host, par, dt, cores)
        t_one = time_oracle(host, par, dt, 1, reps=3)
        out["cpu_baseline"] = {"value": round(zones / t_all, 1), "cores": cores,
                               "one_core_value": round(zones / t_one, 1),
                               "sample": "CPU oracle, same inputs, stage-1 substep incl. "
                                         "con2prim: median of 5 on %d threads, of 3 on 1" %
                                         cores}
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU; --gpus %d is not "
              "supported (run with --gpus 1)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        copy_calibration()
        return
    import torch
    from asterx_amd import abi, lib
    from asterx_amd.driver import all_slots

    assert torch.cuda.is_available(), "bench.py needs a HIP device"
    torch.cuda.set_device(0)
    stream = torch.cuda.Stream()
    sh = C.c_void_p(stream.cuda_stream)
    be = lib.backend(sh)
    L = be.lib
    t0 = time.perf_counter()
    p, par, dt = build_torus_patch(args.n, args.recon, args.flux, stream, be)
    n = p.n
    log(args, "initial data: %.1f s" % (time.perf_counter() - t0))
    zones = n[0] * n[1] * n[2]
    host = None
    if args.cpu == "on":
        # same inputs for the CPU oracle, copied before the GPU evolves them
        host = p.to(None)
        for s in all_slots():
            if s not in host:
                host.alloc(s)

    st = Stepper(p, par, be, dt, sh)
    variants = tune_variants(be, p, st, par, dt, sh) if args.tune == "on" else {}
    log(args, "tuned variants:", variants)
    K = args.steps
    ms_eager, kt, first = time_substeps(st, L, K, args.warmup)
    # the value: the same K substeps replayed from HIP graphs (the per-kernel
    # times above come from the launch-by-launch run, with HIP events)
    ms_graph, launch = graph_timing(st, stream, K)
    ms_step = ms_graph if ms_graph is not None else ms_eager
    log(args, "ms per substep: %.4f launch by launch, %s -> %.4f" % (ms_eager, launch, ms_step))
    log(args, "kernel classes (launches, total ms, avg ms):")
    for name, (nl, ms) in sorted(kt.items(), key=lambda x: -x[1][1]):
        log(args, "  %-8s %4d %9.3f %8.4f" % (name, nl, ms, ms / nl))
    # flags histogram of one more substep, outside the timed region
    cnt = abi.asx_counters()
    ps = p.struct()
    abi.check("asx_con2prim_group", L.asx_con2prim_group(
        C.byref(ps), C.byref(par), C.byref(cnt), sh))
    torch.cuda.synchronize()
    log(args, "c2p flags:", list(cnt.c2p_flags))
    for name in abi.CONS:
        a = p.arr[name][p.interior(name)]
        if not bool(torch.isfinite(a).all()):
            raise RuntimeError("non-finite %s after the timed substeps" % name)
    workload = "torus%d_%s_%s" % (args.n, args.recon, args.flux)
    roof = substep_roofline(par, zones, K, first, kt, ms_step, workload)
    roof["dominant_kernel"] = dominant_kernel(par, n, zones, K, first, kt, workload)
    roof["kernels_ms_per_step"] = {k: round(v[1] / K, 4) for k, v in kt.items()}
    roof["variants"] = variants
    roof["c2p_flags"] = list(cnt.c2p_flags)
    sten_ms = sum(kt[k][1] for k in ("bstag", "face_x", "face_y", "face_z", "aux",
                                     "update") if k in kt) / K
    sten_b = (first * stencil_bytes_per_zone(par, True) +
              (K - first) * stencil_bytes_per_zone(par, False)) / K
    roof["stencil_ms"] = round(sten_ms, 4)
    roof["stencil_frac"] = round(sten_b * zones / (sten_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)

    batched = None
    if args.batch == "on":
        t0 = time.perf_counter()
        batched = bench_batched(args, p, par, dt, L, sh, stream)
        log(args, "batched: %.1f s" % (time.perf_counter() - t0), batched)
    t0 = time.perf_counter()
    tmunu = bench_tmunu(args, p, par, L, sh, workload)
    log(args, "tmunu: %.1f s" % (time.perf_counter() - t0), tmunu)
    regrid = bench_regrid(p, par, L, sh)
    log(args, "regrid:", regrid)
    amr = None
    if args.amr == "on":
        t0 = time.perf_counter()
        amr = bench_amr(p, par, L, sh)
        log(args, "amr: %.1f s" % (time.perf_counter() - t0), amr)
    del st, p
    torch.cuda.synchronize()
    torch.cuda.empty_cache()

    extra = {}
    if args.variants == "on":
        t0 = time.perf_counter()
        extra["torus_variants"] = [
            bench_torus_variant(args, args.n, r, f, 12, 3, False, sh, stream, be)
            for r, f in (("wenoz", "LxF"), ("mp5", "HLLE"), ("mp5", "LxF"))]
        log(args, "variants: %.1f s" % (time.perf_counter() - t0), extra["torus_variants"])
    if args.n256 == "on":
        t0 = time.perf_counter()
        extra["torus256"] = bench_torus_variant(args, 256, args.recon, args.flux, 12, 3,
                                                args.tune == "on", sh, stream, be)
        log(args, "256^3: %.1f s" % (time.perf_counter() - t0), extra["torus256"])
    if args.plain64 == "on":
        t0 = time.perf_counter()
        extra["plain64"] = bench_plain64(args, sh, stream, be)
        log(args, "plain 64^3: %.1f s" % (time.perf_counter() - t0), extra["plain64"])

    guess = None
    if args.c2p_guess == "on":
        t0 = time.perf_counter()
        guess = bench_c2p_guess(args, L)
        log(args, "c2p guess: %.1f s" % (time.perf_counter() - t0), guess)
    tab = None
    if args.tab == "on":
        t0 = time.perf_counter()
        tab = bench_tabulated(args, L, sh, stream)
        log(args, "tabulated: %.1f s" % (time.perf_counter() - t0), tab)

    cpu = None
    if host is not None:
        t0 = time.perf_counter()
        cpu = cpu_baseline(host, par, dt, "%d^3 torus" % args.n)
        log(args, "cpu baseline: %.1f s" % (time.perf_counter() - t0))
        del host

    out = {"metric": "zone-updates/s per RK substep",
           "value": round(zones / (ms_step * 1e-3), 1),
           "unit": "zone-updates/s", "n_gpus": 1, "steps": K,
           "warmup": args.warmup, "ms_per_step": round(ms_step, 4),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f64", "data": "synthetic",
           "config": {"workload": "torus %d^3 +3 ghosts, Kerr-Schild a=0.9375, "
                                  "FM disk, %s + %s, Noble/Palenzuela c2p" %
                                  (args.n, args.recon, args.flux),
                      "zones": zones, "substep": "SSPRK3 stages cycled",
                      "launch": launch},
           "eager_ms_per_step": round(ms_eager, 4),
           "roofline": roof, "cpu_baseline": cpu,
           "batched": batched, "c2p_guess": guess, "tabulated": tab,
           "tmunu": tmunu, "regrid_error": regrid, "amr": amr}
    out.update(extra)
    print(json.dumps(out), flush=True)
    reap_children()


def reap_children():
    """End every child process this run left (none is started on purpose),
    so that no process outlives bench.py; each is named on stderr."""
    try:
        import psutil
    except ImportError:
        return
    kids = psutil.Process().children(recursive=True)
    for k in kids:
        try:
            print("bench.py: ending leftover child %d (%s)" % (k.pid, " ".join(k.cmdline())),
                  file=sys.stderr, flush=True)
            k.terminate()
        except psutil.Error:
            pass
    _, alive = psutil.wait_procs(kids, timeout=5)
    for k in alive:
        try:
            k.kill()
        except psutil.Error:
            pass


if __name__ == "__main__":
    main()
