"""Convert the reference's regression outputs into compact test fixtures.

Reads the CarpetX TSV 1-D cuts that the reference's test suite holds under
/root/reference/AsterX/test/<case>/hydrobasex-<var>.itNNNNNN.<dir>.tsv and
writes tests/golden/tsv_<case>.npz with, per cut direction d in x, y, z:
  it        iterations present
  time      simulation time per iteration
  idx_d     (npts, 3) grid indices (i, j, k) of the cut, counted from the
            first interior cell as in the TSV files
  <var>_d   (nit, npts, ncomp) values for var in rho, vel, eps, press, bvec
Only data is copied (inputs and expected outputs of the reference's own
tests); run in the build container, where the reference tree exists.
"""
import glob
import os
import re
import sys

import numpy as np

REF = "/root/reference/AsterX/test"
CASES = [
    "Balsara1_shocktube_xdir_ppm",
    "Balsara1_shocktube_xdir_wenoz",
    "Balsara1_shocktube_xdir_mp5",
    "Balsara1_shocktube_xdir_godunov",
    "Balsara1_shocktube_xdir_monocentral",
    "Balsara1_shocktube_xdir_ppm_PalenzuelaC2P",
    "Balsara1_shocktube_ydir_ppm",
    "Balsara1_shocktube_zdir_ppm",
    "Balsara2_shocktube_xdir_ppm",
    "Balsara3_shocktube_xdir_minmod_LxF",
    "Balsara4_shocktube_xdir_ppm",
    "Balsara5_shocktube_xdir_ppm",
    "Alfven_wave",
    "Sound_wave",
    "Magnetic_rotor",
    "Cylindrical_blast",
]
VARS = ["rho", "vel", "eps", "press", "bvec"]


def convert(case, outdir):
    out = {}
    for d in "xyz":
        its = None
        for var in VARS:
            files = sorted(glob.glob(os.path.join(
                REF, case, "hydrobasex-%s.it*.%s.tsv" % (var, d))))
            if not files:
                continue
            cur_its, vals, times = [], [], []
            for f in files:
                it = int(re.search(r"\.it(\d+)\.", f).group(1))
                a = np.loadtxt(f, ndmin=2)
                cur_its.append(it)
                times.append(a[0, 1])
                idx = a[:, 4:7].astype(np.int32)
                vals.append(a[:, 10:])
            order = np.argsort(cur_its)
            cur_its = np.array(cur_its)[order]
            if its is None:
                its = cur_its
                out["it_" + d] = its
                out["time_" + d] = np.array(times)[order]
                out["idx_" + d] = idx
            out["%s_%s" % (var, d)] = np.stack([vals[o] for o in order])
    path = os.path.join(outdir, "tsv_%s.npz" % case)
    np.savez_compressed(path, **out)
    return path


# One cut file kept verbatim, for the text format of the TSV writer.
FORMAT_FILE = ("Balsara1_shocktube_xdir_ppm", "hydrobasex-rho.it000000.x.tsv")


def copy_format_file(outdir):
    case, name = FORMAT_FILE
    with open(os.path.join(REF, case, name)) as f:
        txt = f.read()
    path = os.path.join(outdir, "%s.%s" % (case, name))
    with open(path, "w") as f:
        f.write(txt)
    return path


def main():
    outdir = os.path.dirname(os.path.abspath(__file__))
    print(copy_format_file(outdir))
    cases = sys.argv[1:] or CASES
    for c in cases:
        print(convert(c, outdir))


if __name__ == "__main__":
    main()
