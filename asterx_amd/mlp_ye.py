"""The learned con2prim guess on a tabulated EOS: training and evaluation.

The table network takes the five scale-free inputs of the ideal-gas
network, log D and Ye (a table sets a density scale of its own) and
predicts log W and log x, x = h W: the unknown of the Palenzuela solver
(Con2PrimFactory/src/c2p_1DPalenzuela.hxx:306-556), which
asx_c2p_batch_ye_mlp brackets around the guess.  Zones are drawn on the GPU
(asx_zone_sample_ye) on the table, mapped to conserved variables
(asx_prim2con_batch_ye) and the network is trained with the library's own
forward/backward/Adam kernels (asx_mlp_ye_train_step); torch only holds
the device buffers.  The weights belong to the table they were trained on.
"""
import ctypes as C
import math
import os

import numpy as np

from . import abi

NIN = 7
H = 64
NPARAM = H * NIN + H + H * H + H + 2 * H + 2
NBLOB = NPARAM + 2 * NIN + 4
W1, W2, W3 = 0, H * NIN + H, H * NIN + H + H * H + H
_HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_WEIGHTS = os.path.join(_HERE, "data", "c2p_guess_table_ye.npy")
# the tabulated workload's zones (code units): rho, T [MeV], Ye, W max
BOX = (1e-12, 1e-4, 0.05, 80.0, 0.06, 0.5, 10.0)


def _p(t):
    return C.c_void_p(t.data_ptr())


class Zones:
    """A batch of zones with composition on the device (asx_c2p_batch_ye
    layout): prim[9], g[6], seed[7] (truth x (1 + seed_rel u)), target[2]
    (log W, log(h W)), cons[9], thermo[2] (eps, p)."""

    def __init__(self, L, nz, seed, par, box=BOX, seed_rel=1e-2, stream=None):
        import torch
        f64 = dict(dtype=torch.float64, device="cuda")
        self.nz = nz
        self.prim = torch.empty((9, nz), **f64)
        self.g = torch.empty((6, nz), **f64)
        self.seed = torch.empty((7, nz), **f64)
        self.target = torch.empty((2, nz), **f64)
        self.cons = torch.empty((9, nz), **f64)
        self.thermo = torch.empty((2, nz), **f64)
        b = (C.c_double * 7)(*box)
        abi.check("asx_zone_sample_ye", L.asx_zone_sample_ye(
            nz, seed, b, seed_rel, _p(self.prim), _p(self.g), _p(self.seed), _p(self.target),
            C.byref(par), stream))
        abi.check("asx_prim2con_batch_ye", L.asx_prim2con_batch_ye(
            nz, _p(self.prim), _p(self.g), _p(self.cons), _p(self.thermo), C.byref(par),
            stream))


def init_blob(L, par, seed=1234, nstat=1 << 20, box=BOX, stream=None):
    """Xavier-uniform weights, zero biases, and input/output normalisation
    from the statistics of nstat generated zones."""
    import torch
    rng = np.random.default_rng(seed)
    blob = np.zeros(NBLOB, dtype=np.float32)

    def xavier(n_out, n_in):
        a = math.sqrt(6.0 / (n_in + n_out))
        return rng.uniform(-a, a, n_out * n_in).astype(np.float32)
    blob[W1:W1 + H * NIN] = xavier(H, NIN)
    blob[W2:W2 + H * H] = xavier(H, H)
    blob[W3:W3 + 2 * H] = xavier(2, H)
    z = Zones(L, nstat, seed + 7, par, box, stream=stream)
    x = torch.empty((NIN, nstat), dtype=torch.float32, device="cuda")
    abi.check("asx_mlp_ye_features", L.asx_mlp_ye_features(nstat, _p(z.cons), _p(z.g), _p(x),
                                                           stream))
    torch.cuda.synchronize()
    xh = x.cpu().numpy().astype(np.float64)
    th = z.target.cpu().numpy()
    blob[NPARAM:NPARAM + NIN] = xh.mean(1)
    blob[NPARAM + NIN:NPARAM + 2 * NIN] = 1.0 / np.maximum(xh.std(1), 1e-6)
    blob[NPARAM + 2 * NIN:NPARAM + 2 * NIN + 2] = th.mean(1)
    blob[NPARAM + 2 * NIN + 2:NBLOB] = np.maximum(th.std(1), 1e-6)
    return blob


def train(L, par, samples=int(1e7), batch=8192, lr=2e-3, lr_end=2e-5, seed=1234,
          log_every=200, box=BOX, stream=None, log=print, blob=None):
    """Adam with cosine learning-rate decay over samples/batch steps, a fresh
    batch every step.  Returns (blob, loss history)."""
    import torch
    if blob is None:
        blob = init_blob(L, par, seed=seed, box=box, stream=stream)
    b = torch.from_numpy(blob.copy()).cuda()
    m = torch.zeros(NPARAM, dtype=torch.float32, device="cuda")
    v = torch.zeros_like(m)
    nw = (batch + 63) // 64
    partial = torch.empty(nw * (NPARAM + 1), dtype=torch.float32, device="cuda")
    loss = torch.zeros(1, dtype=torch.float32, device="cuda")
    steps = max(1, samples // batch)
    hist = []
    for s in range(1, steps + 1):
        z = Zones(L, batch, seed * 1000003 + s, par, box, stream=stream)
        lr_s = lr_end + 0.5 * (lr - lr_end) * (1 + math.cos(math.pi * (s - 1) / steps))
        abi.check("asx_mlp_ye_train_step", L.asx_mlp_ye_train_step(
            batch, _p(z.cons), _p(z.g), _p(z.target), _p(b), _p(m), _p(v),
            C.c_float(lr_s), s, _p(partial), _p(loss), None, stream))
        if s == 1 or s % log_every == 0 or s == steps:
            torch.cuda.synchronize()
            hist.append((s, float(loss.item())))
            if log:
                log("table mlp step %d/%d loss %.4e lr %.2e" % (s, steps, hist[-1][1], lr_s))
    torch.cuda.synchronize()
    return b.cpu().numpy(), hist


def load(path=DEFAULT_WEIGHTS):
    blob = np.load(path, allow_pickle=False)
    if blob.shape != (NBLOB,) or blob.dtype != np.float32:
        raise ValueError("not a table con2prim-guess weight blob: %s" % path)
    return blob


def save(blob, path=DEFAULT_WEIGHTS):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    np.save(path, np.asarray(blob, dtype=np.float32), allow_pickle=False)


def evaluate(L, par, blob, nz=1 << 20, seed=99, bracket=0.03, reps=3, box=BOX, stream=None):
    """Plain (asx_c2p_batch_ye) and guided (asx_c2p_batch_ye_mlp) con2prim on
    the same nz fresh zones with the previous-step seed: zones/s (median of
    reps, HIP events), Brent iterations, flags, the largest difference
    between the two and against the truth, the guess's error and the share
    of zones whose root lies inside the guessed bracket."""
    import torch
    z = Zones(L, nz, seed, par, box, stream=stream)
    bt = torch.from_numpy(np.asarray(blob, dtype=np.float32)).cuda()
    f64 = dict(dtype=torch.float64, device="cuda")
    res = {}
    outs = {}
    for mode in ("plain", "mlp"):
        prim = torch.empty((9, nz), **f64)
        flag = torch.empty(nz, **f64)
        it = torch.empty(nz, dtype=torch.int32, device="cuda")

        def run():
            if mode == "plain":
                return L.asx_c2p_batch_ye(nz, _p(z.cons), _p(z.g), _p(z.seed), _p(prim),
                                          _p(flag), _p(it), C.byref(par), stream)
            return L.asx_c2p_batch_ye_mlp(nz, _p(z.cons), _p(z.g), _p(z.seed), _p(bt),
                                          bracket, _p(prim), _p(flag), _p(it), C.byref(par),
                                          stream)
        abi.check("c2p", run())
        ts = []
        for _ in range(reps):
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record()
            abi.check("c2p", run())
            e1.record()
            torch.cuda.synchronize()
            ts.append(e0.elapsed_time(e1))
        ms = sorted(ts)[len(ts) // 2]
        fl = flag.cpu().numpy()
        res[mode] = dict(zones_per_s=nz / (ms * 1e-3), ms=ms,
                         iters_mean=float(it.double().mean()),
                         flags={int(k): int(v) for k, v in zip(*np.unique(fl, return_counts=True))})
        outs[mode] = (prim, fl)
    ok = (outs["plain"][1] == 1) & (outs["mlp"][1] == 1)
    okt = torch.from_numpy(ok).cuda()
    pa, pb = outs["plain"][0], outs["mlp"][0]
    for row, name in ((0, "rho"), (7, "temperature")):
        d = ((pb[row] - pa[row]).abs() / pa[row].abs())[okt]
        res["max_rel_diff_%s" % name] = float(d.max()) if d.numel() else 0.0
    truth = {0: z.prim[0], 7: z.prim[1]}
    for row, name in ((0, "rho"), (7, "temperature")):
        d = ((pb[row] - truth[row]).abs() / truth[row])[okt]
        res["mlp_max_rel_err_%s" % name] = float(d.max()) if d.numel() else 0.0
    guess = torch.empty((2, nz), **f64)
    abi.check("asx_mlp_ye_forward", L.asx_mlp_ye_forward(nz, _p(z.cons), _p(z.g), _p(bt),
                                                         _p(guess), stream))
    torch.cuda.synchronize()
    err = (guess - z.target).abs().cpu().numpy()
    res["guess_abs_err"] = dict(logW_median=float(np.median(err[0])),
                                logx_median=float(np.median(err[1])),
                                logW_p99=float(np.percentile(err[0], 99)),
                                logx_p99=float(np.percentile(err[1], 99)))
    res["root_in_bracket"] = float(np.mean(err[1] < math.log1p(bracket)))
    res["speedup"] = res["mlp"]["zones_per_s"] / res["plain"]["zones_per_s"]
    res["zones"], res["bracket"] = nz, bracket
    return res


def table_params(table, **kw):
    """Palenzuela on the temperature on `table` (tabeos.Table), the
    atmosphere below the workload's densities."""
    base = dict(c2p_prime="Palenzuela", c2p_second="None", use_temperature=1,
                use_press_atmo=0, thermal_eos_atmo=1, rho_abs_min=1e-20, t_atmo=0.01,
                Ye_atmo=0.3, c2p_tol=1e-10, max_iter=100, vw_lim=1e3, B_lim=1e3,
                ye_min=0.035, ye_max=0.55)
    base.update(kw)
    return table.attach(abi.params(**base))


def main():
    import argparse
    import json
    import sys
    import time
    from . import lib, tabeos
    ap = argparse.ArgumentParser(description="train / evaluate the table con2prim guess")
    ap.add_argument("--samples", type=float, default=1e7)
    ap.add_argument("--batch", type=int, default=8192)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--out", default=DEFAULT_WEIGHTS)
    ap.add_argument("--eval-only", action="store_true")
    ap.add_argument("--eval-zones", type=int, default=1 << 20)
    ap.add_argument("--bracket", type=float, default=0.03)
    args = ap.parse_args()
    L = lib.lib()
    table = tabeos.Table(tabeos.build(), device="cuda")
    par = table_params(table)
    if args.eval_only:
        blob = load(args.out)
    else:
        t0 = time.time()
        blob, hist = train(L, par, samples=int(args.samples), batch=args.batch, lr=args.lr,
                           log=lambda s: print(s, file=sys.stderr, flush=True))
        save(blob, args.out)
        print(json.dumps({"trained": args.out, "seconds": round(time.time() - t0, 1),
                          "loss": hist}), flush=True)
    print(json.dumps(evaluate(L, par, blob, nz=args.eval_zones, bracket=args.bracket)),
          flush=True)


if __name__ == "__main__":
    main()
