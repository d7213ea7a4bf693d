"""CarpetX TSV 1-D cuts of a patch, the output format of the reference's
regression data (AsterX/test/<case>/hydrobasex-<group>.itNNNNNN.<d>.tsv):
a header "# 1:iteration 2:time 3:patch 4:level 5:i 6:j 7:k 8:x 9:y 10:z 11:..."
(tab-separated), then one row per cell along axis d, ghosts included, with
the cell indices counted from the first interior cell (ghosts negative),
the cell-centre coordinates and the group's variables; integers in decimal,
reals as %.16e.  Readers of the reference's TSV output (its test harness,
analysis scripts) read these files unchanged.
"""
import numpy as np

from . import abi

# HydroBaseX groups written by the reference's tests, with their variables
GROUPS = {"rho": ["rho"], "vel": ["velx", "vely", "velz"], "eps": ["eps"],
          "press": ["press"], "bvec": ["Bvecx", "Bvecy", "Bvecz"],
          "temperature": ["temperature"], "entropy": ["entropy"], "ye": ["Ye"]}


def cut_rows(patch, names, d, line):
    """(idx, xyz, vals) of the cut along axis d through the cells whose
    other two indices (counted from the first interior cell) are line[e]."""
    ng = patch.ng
    m = patch.n[d] + 2 * ng
    idx = np.zeros((m, 3), dtype=np.int64)
    for e in range(3):
        idx[:, e] = np.arange(m) - ng if e == d else line[e]
    # cell centres as CarpetX evaluates them: (x0 + dx/2) + i dx (this
    # rounding reproduces the reference's coordinate columns bit for bit)
    xyz = np.stack([(patch.x0[e] + 0.5 * patch.dx[e]) + idx[:, e] * patch.dx[e]
                    for e in range(3)], axis=1)
    vals = []
    for n in names:
        if abi.centering(n) != 0:
            raise ValueError("%s is not cell-centred" % n)
        a = patch.numpy(n)
        vals.append(a[idx[:, 2] + ng, idx[:, 1] + ng, idx[:, 0] + ng])
    return idx, xyz, np.stack(vals, axis=1)


def format_cut(patch, names, d, line, iteration=0, time=0.0, patch_id=0, level=0):
    """The text of one TSV cut file."""
    idx, xyz, vals = cut_rows(patch, names, d, line)
    cols = ["iteration", "time", "patch", "level", "i", "j", "k", "x", "y", "z"] + list(names)
    out = ["# " + "\t".join("%d:%s" % (c + 1, n) for c, n in enumerate(cols))]
    for r in range(idx.shape[0]):
        f = ["%d" % iteration, "%.16e" % time, "%d" % patch_id, "%d" % level]
        f += ["%d" % v for v in idx[r]]
        f += ["%.16e" % v for v in xyz[r]]
        f += ["%.16e" % v for v in vals[r]]
        out.append("\t".join(f))
    return "\n".join(out) + "\n"


def write_cut(path, patch, group, d, line, iteration=0, time=0.0):
    """Write hydrobasex-<group> along axis d ("x", "y", "z" or 0..2)."""
    d = "xyz".index(d) if isinstance(d, str) else d
    with open(path, "w") as f:
        f.write(format_cut(patch, GROUPS[group], d, line, iteration, time))


def read_cut(path):
    """(columns, rows) of a TSV cut: the header's names and a float array."""
    with open(path) as f:
        head = f.readline()
    cols = [c.split(":", 1)[1] for c in head[1:].split()]
    return cols, np.loadtxt(path, ndmin=2)
