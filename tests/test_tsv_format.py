"""The TSV writer against a cut file of the reference's own regression data
(AsterX/test/Balsara1_shocktube_xdir_ppm/hydrobasex-rho.it000000.x.tsv, kept
verbatim in tests/golden): the initial Balsara-1 state written by
asterx_amd.tsv reproduces the file byte for byte (header, index columns,
coordinates and values), and the reader returns its columns."""
import os
import sys

import numpy as np

from asterx_amd import tsv

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import tsv_cases  # noqa: E402

REF = os.path.join(HERE, "golden",
                   "Balsara1_shocktube_xdir_ppm.hydrobasex-rho.it000000.x.tsv")


def test_balsara_initial_cut_byte_for_byte(tmp_path):
    c = tsv_cases.CASES["Balsara1_shocktube_xdir_ppm"]
    p = tsv_cases.make_patch(c)
    out = tmp_path / "rho.tsv"
    tsv.write_cut(str(out), p, "rho", "x", (0, 0, 0), iteration=0, time=0.0)
    with open(REF) as f:
        want = f.read()
    assert out.read_text() == want


def test_reader_and_multicomponent_groups(tmp_path):
    cols, rows = tsv.read_cut(REF)
    assert cols[:10] == ["iteration", "time", "patch", "level", "i", "j", "k", "x", "y", "z"]
    assert cols[10:] == ["rho"] and rows.shape == (56, 11)
    c = tsv_cases.CASES["Balsara1_shocktube_xdir_ppm"]
    p = tsv_cases.make_patch(c)
    out = tmp_path / "bvec.tsv"
    tsv.write_cut(str(out), p, "bvec", "y", (10, 0, 1))
    cols, rows = tsv.read_cut(str(out))
    assert cols[10:] == ["Bvecx", "Bvecy", "Bvecz"]
    ng = p.ng
    assert np.array_equal(rows[:, 5], np.arange(-ng, p.n[1] + ng))
    assert np.all(rows[:, 4] == 10) and np.all(rows[:, 6] == 1)
    assert np.array_equal(rows[:, 11], p.numpy("Bvecy")[1 + ng, :, 10 + ng])
