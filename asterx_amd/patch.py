"""A CarpetX-style patch: geometry plus one dense FP64 array per grid function.

Arrays are numpy (host) or torch tensors on a device; both are laid out
x-fastest over the centering's full extent including ghosts, shape
(ez, ey, ex), which is what include/asterx.h requires.
"""
import ctypes as C

import numpy as np

from . import abi


class Patch:
    def __init__(self, n, ng, x0, dx, slots=(), device=None):
        self.n = tuple(int(v) for v in n)
        self.ng = int(ng)
        self.x0 = tuple(float(v) for v in x0)
        self.dx = tuple(float(v) for v in dx)
        self.device = device
        self.arr = {}
        for s in slots:
            self.alloc(s)

    # ---- storage ------------------------------------------------------
    def shape(self, name):
        ex, ey, ez = abi.extents(self.n, self.ng, abi.centering(name))
        return (ez, ey, ex)

    def _new(self, shape, fill):
        if self.device is None:
            return np.full(shape, fill, dtype=np.float64)
        import torch
        return torch.full(shape, fill, dtype=torch.float64, device=self.device)

    def alloc(self, name, fill=0.0):
        if name not in abi.SLOT:
            raise KeyError(name)
        self.arr[name] = self._new(self.shape(name), fill)
        return self.arr[name]

    def __getitem__(self, name):
        return self.arr[name]

    def __setitem__(self, name, value):
        if name not in self.arr:
            self.alloc(name)
        if self.device is None:
            self.arr[name][...] = np.asarray(value, dtype=np.float64)
        else:
            import torch
            if isinstance(value, np.ndarray):
                value = torch.from_numpy(np.ascontiguousarray(value))
            self.arr[name].copy_(value if not isinstance(value, (int, float))
                                 else torch.full_like(self.arr[name], value))

    def __contains__(self, name):
        return name in self.arr

    def numpy(self, name):
        a = self.arr[name]
        if self.device is None:
            return a
        return a.detach().cpu().numpy()

    def to(self, device, names=None):
        """Copy (all or the named) arrays to a new Patch on `device`."""
        out = Patch(self.n, self.ng, self.x0, self.dx, device=device)
        for k in (names or list(self.arr)):
            a = self.arr[k]
            if device is None:
                out.arr[k] = self.numpy(k).copy()
            else:
                import torch
                src = torch.from_numpy(np.ascontiguousarray(a)) \
                    if isinstance(a, np.ndarray) else a
                out.arr[k] = src.to(device=device, dtype=torch.float64).contiguous().clone()
        return out

    # ---- index helpers ------------------------------------------------
    def interior(self, name, halo=0):
        """Slice tuple (k, j, i) of the interior (widened by halo)."""
        c = abi.centering(name)
        sl = []
        for d in (2, 1, 0):
            lo = self.ng - halo
            hi = self.ng + self.n[d] + ((c >> d) & 1) + halo
            sl.append(slice(lo, hi))
        return tuple(sl)

    def coords(self, cent):
        """Coordinates (z, y, x) of a centering as broadcastable arrays."""
        out = []
        for d in (2, 1, 0):
            c = (cent >> d) & 1
            m = self.n[d] + 2 * self.ng + c
            idx = np.arange(m, dtype=np.float64) - self.ng
            off = 0.0 if c else 0.5
            v = self.x0[d] + (idx + off) * self.dx[d]
            shp = [1, 1, 1]
            shp[2 - d] = m
            out.append(v.reshape(shp))
        return out  # z, y, x

    # ---- ABI ----------------------------------------------------------
    def _ptr(self, a):
        if self.device is None:
            return a.ctypes.data
        return a.data_ptr()

    def struct(self):
        p = abi.asx_patch()
        for d in range(3):
            p.n[d] = self.n[d]
            p.x0[d] = self.x0[d]
            p.dx[d] = self.dx[d]
        p.ng = self.ng
        for name, a in self.arr.items():
            i = abi.SLOT[name]
            ex, ey, _ = abi.extents(self.n, self.ng, abi.centering(name))
            p.gf[i].data = self._ptr(a)
            p.gf[i].sj = ex
            p.gf[i].sk = ex * ey
        self._keep = p
        return p

    def state(self, names=abi.STATE):
        """asx_state over the given 11 grid functions of this patch."""
        s = abi.asx_state()
        for q, name in enumerate(names):
            a = self.arr[name]
            ex, ey, _ = abi.extents(self.n, self.ng, abi.centering(name))
            s.gf[q].data = self._ptr(a)
            s.gf[q].sj = ex
            s.gf[q].sk = ex * ey
        return s


class StateBuf:
    """Eleven state arrays with the centerings of the evolved state."""

    def __init__(self, patch, names=abi.STATE, fill=0.0):
        self.patch = patch
        self.names = names
        self.arr = {}
        for q, name in enumerate(abi.STATE):
            self.arr[name] = patch._new(patch.shape(name), fill)

    def struct(self):
        s = abi.asx_state()
        for q, name in enumerate(abi.STATE):
            a = self.arr[name]
            ex, ey, _ = abi.extents(self.patch.n, self.patch.ng,
                                    abi.centering(name))
            s.gf[q].data = self.patch._ptr(a)
            s.gf[q].sj = ex
            s.gf[q].sk = ex * ey
        return s


def ptr_array(states):
    """C array of pointers to asx_state structs (kept alive by the caller)."""
    arr = (C.POINTER(abi.asx_state) * len(states))()
    for i, s in enumerate(states):
        arr[i] = C.pointer(s)
    return arr
