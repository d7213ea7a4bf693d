"""Training and evaluation of the learned con2prim initial guess.

All computation runs in the HIP library: zones are drawn on the GPU
(asx_zone_sample), mapped to conserved variables (asx_prim2con_batch), and
the network is trained with its own forward/backward/Adam kernels
(asx_mlp_train_step).  torch only holds the device buffers.  Every step
draws a fresh batch, so `samples` counts distinct generated pairs.

The committed weights (DEFAULT_WEIGHTS) are the ReLU network trained by
`python -m asterx_amd.mlp --samples 1.6e10 --batch 65536 --lr 3e-3`
(about 90 s on one MI355X).
"""
import ctypes as C
import math
import os

import numpy as np

from . import abi

NPARAM = 4674
NBLOB = 4688
_HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_WEIGHTS = os.path.join(_HERE, "data", "c2p_guess_gamma5_3.npy")


def _p(t):
    return C.c_void_p(t.data_ptr())


class Zones:
    """A batch of independent zones on the device (asx_c2p_batch layout)."""

    def __init__(self, L, nz, seed, par, gamma, seed_rel=1e-2, stream=None):
        import torch
        dev = "cuda"
        f64 = dict(dtype=torch.float64, device=dev)
        self.nz = nz
        self.prim = torch.empty((8, nz), **f64)
        self.g = torch.empty((6, nz), **f64)
        self.seed = torch.empty((5, nz), **f64)
        self.target = torch.empty((2, nz), **f64)
        self.cons = torch.empty((8, nz), **f64)
        self.cold = torch.empty((5, nz), **f64)
        abi.check("asx_zone_sample", L.asx_zone_sample(
            nz, seed, gamma, seed_rel, _p(self.prim), _p(self.g), _p(self.seed),
            _p(self.target), stream))
        abi.check("asx_prim2con_batch", L.asx_prim2con_batch(
            nz, _p(self.prim), _p(self.g), _p(self.cons), _p(self.cold),
            C.byref(par), stream))


def init_blob(L, par, gamma, seed=1234, nstat=1 << 20, stream=None):
    """He-uniform weights (the hidden layers are ReLU), zero biases, and
    input/output normalisation from the statistics of nstat generated
    zones."""
    import torch
    rng = np.random.default_rng(seed)
    blob = np.zeros(NBLOB, dtype=np.float32)

    def he(n_out, n_in):
        a = math.sqrt(6.0 / n_in)
        return rng.uniform(-a, a, n_out * n_in).astype(np.float32)
    blob[0:320] = he(64, 5)
    blob[384:4480] = he(64, 64)
    blob[4544:4672] = he(2, 64)
    z = Zones(L, nstat, seed + 7, par, gamma, stream=stream)
    x = torch.empty((5, nstat), dtype=torch.float32, device="cuda")
    abi.check("asx_mlp_features", L.asx_mlp_features(nstat, _p(z.cons), _p(z.g), _p(x),
                                                     stream))
    torch.cuda.synchronize()
    xh = x.cpu().numpy().astype(np.float64)
    th = z.target.cpu().numpy()
    mean, std = xh.mean(1), xh.std(1)
    blob[4674:4679] = mean
    blob[4679:4684] = 1.0 / np.maximum(std, 1e-6)
    blob[4684:4686] = th.mean(1)
    blob[4686:4688] = np.maximum(th.std(1), 1e-6)
    return blob


def train(L, par, gamma, samples=int(1e7), batch=8192, lr=2e-3, lr_end=2e-5,
          seed=1234, log_every=200, stream=None, log=print, blob=None):
    """Adam with cosine learning-rate decay over samples/batch steps, a fresh
    batch every step.  Returns (blob, loss history)."""
    import torch
    if blob is None:
        blob = init_blob(L, par, gamma, seed=seed, stream=stream)
    b = torch.from_numpy(blob.copy()).cuda()
    m = torch.zeros(NPARAM, dtype=torch.float32, device="cuda")
    v = torch.zeros_like(m)
    nw = (batch + 63) // 64
    partial = torch.empty(nw * (NPARAM + 1), dtype=torch.float32, device="cuda")
    loss = torch.zeros(1, dtype=torch.float32, device="cuda")
    steps = max(1, samples // batch)
    hist = []
    for s in range(1, steps + 1):
        z = Zones(L, batch, seed * 1000003 + s, par, gamma, stream=stream)
        lr_s = lr_end + 0.5 * (lr - lr_end) * (1 + math.cos(math.pi * (s - 1) / steps))
        abi.check("asx_mlp_train_step", L.asx_mlp_train_step(
            batch, _p(z.cons), _p(z.g), _p(z.target), _p(b), _p(m), _p(v),
            C.c_float(lr_s), s, _p(partial), _p(loss), None, stream))
        if s == 1 or s % log_every == 0 or s == steps:
            torch.cuda.synchronize()
            hist.append((s, float(loss.item())))
            if log:
                log("mlp step %d/%d loss %.4e lr %.2e" % (s, steps, hist[-1][1], lr_s))
    torch.cuda.synchronize()
    return b.cpu().numpy(), hist


def load(path=DEFAULT_WEIGHTS):
    blob = np.load(path, allow_pickle=False)
    if blob.shape != (NBLOB,) or blob.dtype != np.float32:
        raise ValueError("not a con2prim-guess weight blob: %s" % path)
    return blob


def save(blob, path=DEFAULT_WEIGHTS):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    np.save(path, np.asarray(blob, dtype=np.float32), allow_pickle=False)


def evaluate(L, par, gamma, blob, nz=1 << 20, seed=99, kguess=8, reps=3, stream=None,
             separate=False):
    """Plain and learned-guess con2prim on nz fresh zones, for the perturbed
    seed (previous-step primitives) and the cold start: zones/s (median of
    reps timed runs), iteration histograms, fallback rate, accuracy.  The
    learned guess runs with the network inside the solver kernel ("mlp");
    separate=True adds the form with a forward-pass kernel of its own
    ("mlp_separate")."""
    import torch
    z = Zones(L, nz, seed, par, gamma, stream=stream)
    bt = torch.from_numpy(np.asarray(blob, dtype=np.float32)).cuda()
    f64 = dict(dtype=torch.float64, device="cuda")
    prim = torch.empty((8, nz), **f64)
    flag = torch.empty(nz, **f64)
    it = torch.empty(nz, dtype=torch.int32, device="cuda")
    fb = torch.empty(nz, dtype=torch.int32, device="cuda")
    out = {}
    modes = ("plain", "mlp", "mlp_separate") if separate else ("plain", "mlp")
    for sname, sd in (("perturbed", z.seed), ("cold", z.cold)):
        for mode in modes:
            abi.check("asx_set_variant", L.asx_set_variant(
                abi.KCLASS["mlp"], 1 if mode == "mlp_separate" else 0))

            def run():
                if mode == "plain":
                    return L.asx_c2p_batch(nz, _p(z.cons), _p(z.g), _p(sd), _p(prim),
                                           _p(flag), _p(it), C.byref(par), stream)
                return L.asx_c2p_batch_mlp(nz, _p(z.cons), _p(z.g), _p(sd), _p(bt), kguess,
                                           _p(prim), _p(flag), _p(it), _p(fb), C.byref(par),
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
            iters = it.cpu().numpy()
            fl = flag.cpu().numpy()
            rho = prim[0].cpu().numpy()
            rho_true = z.prim[0].cpu().numpy()
            r = dict(zones_per_s=nz / (ms * 1e-3), ms=ms,
                     iters_mean=float(iters.mean()),
                     iters_hist={int(k): int(v) for k, v in
                                 zip(*np.unique(iters, return_counts=True))},
                     flags={int(k): int(v) for k, v in
                            zip(*np.unique(fl, return_counts=True))},
                     rho_rel_err_max=float(np.max(np.abs(rho - rho_true) / rho_true)))
            if mode != "plain":
                r["fallback_rate"] = float(fb.cpu().numpy().mean())
            out["%s_%s" % (sname, mode)] = r
    abi.check("asx_set_variant", L.asx_set_variant(abi.KCLASS["mlp"], 0))
    # accuracy of the guess itself, and the forward pass's launch time
    g2 = torch.empty((2, nz), **f64)
    abi.check("asx_mlp_forward", L.asx_mlp_forward(nz, _p(z.cons), _p(z.g), _p(bt), _p(g2),
                                                   stream))
    torch.cuda.synchronize()
    abi.check("asx_timing_enable", L.asx_timing_enable(1))
    for _ in range(reps):
        abi.check("asx_mlp_forward", L.asx_mlp_forward(nz, _p(z.cons), _p(z.g), _p(bt),
                                                       _p(g2), stream))
    torch.cuda.synchronize()
    nl, ms = C.c_int64(), C.c_double()
    abi.check("asx_timing_read", L.asx_timing_read(abi.KCLASS["mlp"], C.byref(nl),
                                                  C.byref(ms)))
    abi.check("asx_timing_enable", L.asx_timing_enable(0))
    out["forward_ms"] = ms.value / max(1, nl.value)
    t = z.target.cpu().numpy()
    gh = g2.cpu().numpy()
    out["guess_abs_err"] = dict(logW_median=float(np.median(np.abs(gh[0] - t[0]))),
                                logZD_median=float(np.median(np.abs(gh[1] - t[1]))),
                                logW_p99=float(np.percentile(np.abs(gh[0] - t[0]), 99)),
                                logZD_p99=float(np.percentile(np.abs(gh[1] - t[1]), 99)))
    return out


def main():
    import argparse
    import json
    import sys
    import time
    from . import lib
    ap = argparse.ArgumentParser(description="train / evaluate the con2prim guess")
    ap.add_argument("--samples", type=float, default=1e7)
    ap.add_argument("--batch", type=int, default=8192)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--out", default=DEFAULT_WEIGHTS)
    ap.add_argument("--eval-only", action="store_true")
    ap.add_argument("--eval-zones", type=int, default=1 << 20)
    ap.add_argument("--kguess", type=int, default=8)
    args = ap.parse_args()
    L = lib.lib()
    par = abi.params(gl_gamma=5.0 / 3.0, eps_max=1e8, rho_max=1e8, rho_abs_min=1e-12,
                     vw_lim=1e3, B_lim=1e3, c2p_tol=1e-10)
    if args.eval_only:
        blob = load(args.out)
    else:
        t0 = time.time()
        blob, hist = train(L, par, 5.0 / 3.0, samples=int(args.samples), batch=args.batch,
                           lr=args.lr, log=lambda s: print(s, file=sys.stderr, flush=True))
        save(blob, args.out)
        print(json.dumps({"trained": args.out, "seconds": round(time.time() - t0, 1),
                          "loss": hist}), flush=True)
    print(json.dumps(evaluate(L, par, 5.0 / 3.0, blob, nz=args.eval_zones,
                              kguess=args.kguess)), flush=True)


if __name__ == "__main__":
    main()
