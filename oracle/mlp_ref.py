"""CPU reference of the learned con2prim-guess network.

Test infrastructure only: the HIP kernels' inference and gradient are
checked against it.  Same network, parameter layout and arithmetic as the
kernels -- inputs computed in double precision from the densitized conserved
variables, every product with bf16-rounded operands and float32
accumulation, ReLU hidden activations for the ideal-gas network (5 inputs)
and tanh for the table network (7 inputs) -- and the gradient of the mean squared error
of the normalised outputs with bf16 rounding taken as the identity in the
derivatives (bf16=True).  With bf16=False every rounding is dropped and the
arithmetic is float64, so the backpropagation itself can be checked against
finite differences.
"""
import numpy as np

NH, NOUT = 64, 2


def layout(nin):
    """Offsets of the blob of the network with nin inputs (5: the ideal
    gas; 7: a table, with log D and Ye added)."""
    L = dict(NIN=nin, W1=0)
    L["B1"] = L["W1"] + NH * nin
    L["W2"] = L["B1"] + NH
    L["B2"] = L["W2"] + NH * NH
    L["W3"] = L["B2"] + NH
    L["B3"] = L["W3"] + NOUT * NH
    L["NPARAM"] = L["B3"] + NOUT
    L["IN_MEAN"] = L["NPARAM"]
    L["IN_ISTD"] = L["IN_MEAN"] + nin
    L["OUT_MEAN"] = L["IN_ISTD"] + nin
    L["OUT_STD"] = L["OUT_MEAN"] + NOUT
    L["NBLOB"] = L["OUT_STD"] + NOUT
    return L


_L5 = layout(5)
NIN = 5
W1, B1, W2, B2, W3, B3 = (_L5[k] for k in ("W1", "B1", "W2", "B2", "W3", "B3"))
NPARAM, IN_MEAN, IN_ISTD = _L5["NPARAM"], _L5["IN_MEAN"], _L5["IN_ISTD"]
OUT_MEAN, OUT_STD, NBLOB = _L5["OUT_MEAN"], _L5["OUT_STD"], _L5["NBLOB"]


def bf16(x):
    """float32 -> nearest bf16 (ties to even), returned as float32."""
    a = np.asarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    out = r.astype(np.uint32).view(np.float32)
    return np.where(np.isnan(a), np.float32(np.nan), out)


def features(cons, g):
    """x[nz, 5]: log D, log S^2, log tau, log B^2 in units of E = D + tau
    (tau floored at 1e-15 D, S^2/E^2 and B^2/E at 1e-30), and cos(S, B)
    (float64)."""
    gxx, gxy, gxz, gyy, gyz, gzz = g
    det = 2 * gxy * gxz * gyz + gxx * gyy * gzz - gyy * gxz ** 2 - \
        gxx * gyz ** 2 - gzz * gxy ** 2
    sq = np.sqrt(det)
    inv = [(gyy * gzz - gyz ** 2) / det, (gxz * gyz - gxy * gzz) / det,
           (gxy * gyz - gxz * gyy) / det, (gxx * gzz - gxz ** 2) / det,
           (gxy * gxz - gxx * gyz) / det, (gxx * gyy - gxy ** 2) / det]
    D = cons[0] / sq
    S = cons[1:4] / sq
    tau = cons[4] / sq
    B = cons[5:8] / sq

    def quad(m, a, b):
        xx, xy, xz, yy, yz, zz = m
        return (a[0] * (xx * b[0] + xy * b[1] + xz * b[2]) +
                a[1] * (xy * b[0] + yy * b[1] + yz * b[2]) +
                a[2] * (xz * b[0] + yz * b[1] + zz * b[2]))
    S2 = quad(inv, S, S)
    B2 = quad(g, B, B)
    SB = S[0] * B[0] + S[1] * B[1] + S[2] * B[2]
    den = np.sqrt(S2 * B2)
    cosv = np.where(den > 0, SB / np.where(den > 0, den, 1.0), 0.0)
    taup = np.maximum(tau, 1e-15 * D)
    E = D + taup
    return np.stack([np.log(D / E), np.log(np.maximum(S2 / E ** 2, 1e-30)),
                     np.log(taup / E), np.log(np.maximum(B2 / E, 1e-30)), cosv], axis=1)


def features_tab(cons, g):
    """x[nz, 7]: the 5 inputs of features(), log D (undensitized, floored at
    1e-37) and Ye = DYe / D (cons row 8)."""
    gxx, gxy, gxz, gyy, gyz, gzz = g
    det = 2 * gxy * gxz * gyz + gxx * gyy * gzz - gyy * gxz ** 2 - \
        gxx * gyz ** 2 - gzz * gxy ** 2
    D = cons[0] / np.sqrt(det)
    extra = np.stack([np.log(np.maximum(D, 1e-37)), cons[8] / cons[0]], axis=1)
    return np.concatenate([features(cons, g), extra], axis=1)


def _nin(blob):
    for nin in (5, 7):
        if len(blob) == layout(nin)["NBLOB"]:
            return nin
    raise ValueError("not a network blob: %d values" % len(blob))


def _params(blob, bf):
    L = layout(_nin(blob))
    f = bf16 if bf else (lambda a: np.asarray(a, dtype=np.float64))
    dt = np.float32 if bf else np.float64
    p = dict(W1=f(blob[L["W1"]:L["B1"]].reshape(NH, L["NIN"])),
             b1=blob[L["B1"]:L["W2"]].astype(dt),
             W2=f(blob[L["W2"]:L["B2"]].reshape(NH, NH)), b2=blob[L["B2"]:L["W3"]].astype(dt),
             W3=f(blob[L["W3"]:L["B3"]].reshape(NOUT, NH)),
             b3=blob[L["B3"]:L["NPARAM"]].astype(dt))
    return p


def normalized_inputs(blob, cons, g, bf=True):
    """The network's inputs (5 or 7, from the blob's size), normalised.
    In float32 (bf=True) the logarithms are of float32-rounded ratios, as on
    the device, to within float32 rounding."""
    L = layout(_nin(blob))
    x = features(cons, g) if L["NIN"] == 5 else features_tab(cons, g)
    dt = np.float32 if bf else np.float64
    x = x.astype(np.float32) if bf else x
    xn = (x - blob[L["IN_MEAN"]:L["IN_ISTD"]].astype(dt)) * \
        blob[L["IN_ISTD"]:L["OUT_MEAN"]].astype(dt)
    return xn.astype(dt)


def activation(nin):
    """(act, dact) of the hidden layers, dact as a function of the value
    h = act(z): ReLU for the ideal-gas network, tanh for the table's."""
    if nin == 5:
        return (lambda z: np.maximum(z, 0.0)), (lambda h: (h > 0).astype(h.dtype))
    return np.tanh, (lambda h: 1 - h ** 2)


def forward(blob, cons, g, bf=True):
    """Normalised outputs y[nz, 2] and the cache for backward (5 or 7
    inputs, from the blob's size)."""
    r = bf16 if bf else (lambda a: a)
    p = _params(blob, bf)
    act, _ = activation(_nin(blob))
    xb = r(normalized_inputs(blob, cons, g, bf))
    z1 = xb @ p["W1"].T + p["b1"]
    h1 = act(z1)
    h1b = r(h1)
    z2 = h1b @ p["W2"].T + p["b2"]
    h2 = act(z2)
    h2b = r(h2)
    y = h2b @ p["W3"].T + p["b3"]
    return y, dict(p=p, xb=xb, h1=h1, h1b=h1b, h2=h2, h2b=h2b)


def normalized_targets(blob, target, bf=True):
    L = layout(_nin(blob))
    dt = np.float32 if bf else np.float64
    t = np.asarray(target).T.astype(dt)
    return (t - blob[L["OUT_MEAN"]:L["OUT_STD"]].astype(dt)) / \
        blob[L["OUT_STD"]:L["NBLOB"]].astype(dt)


def loss_and_grad(blob, cons, g, target, bf=True):
    """Mean squared error of the normalised outputs and its gradient with
    respect to the NPARAM parameters (flat, blob order)."""
    r = bf16 if bf else (lambda a: a)
    y, c = forward(blob, cons, g, bf)
    tn = normalized_targets(blob, target, bf)
    n = y.shape[0]
    e = y - tn
    loss = float((e * e).sum() / n)
    dy = 2.0 * e / n
    p = c["p"]
    gW3 = dy.T @ c["h2b"]
    gb3 = dy.sum(0)
    _, dact = activation(_nin(blob))
    dh2 = dy @ p["W3"]
    dz2 = dh2 * dact(c["h2"])
    dz2b = r(dz2)
    gW2 = dz2b.T @ c["h1b"]
    gb2 = dz2.sum(0)
    dh1 = dz2b @ p["W2"]
    dz1 = dh1 * dact(c["h1"])
    gW1 = dz1.T @ c["xb"]
    gb1 = dz1.sum(0)
    grad = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3])
    return loss, grad
