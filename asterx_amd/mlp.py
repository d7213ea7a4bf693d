"""Training and evaluation of the learned con2prim initial guess.

All computation runs in the HIP library: zones are drawn on the GPU
(asx_zone_sample), mapped to conserved variables (asx_prim2con_batch), and
the network is trained with its own forward/backward/Adam kernels
(asx_mlp_train_step).  torch only holds the device buffers.  Every step
draws a fresh batch, so `samples` counts distinct generated pairs.
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
    """Xavier-uniform weights, zero biases, and input/output normalisation
    from the statistics of nstat generated zones."""
    import torch
    rng = np.random.default_rng(seed)
    blob = np.zeros(NBLOB, dtype=np.float32)

    def xavier(n_out, n_in):
        a = math.sqrt(6.0 / (n_in + n_out))
        return rng.uniform(-a, a, n_out * n_in).astype(np.float32)
    blob[0:320] = xavier(64, 5)
    blob[384:4480] = xavier(64, 64)
    blob[4544:4672] = xavier(2, 64)
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
        if s % log_every == 0 or s == steps:
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
