"""The CPU oracle reproduces the reference's regression outputs.

Each case restates one AsterX/test/*.par (Balsara 1-5 shock tubes with PPM,
WENO-Z, MP5, minmod, monocentral, Godunov; HLLE and LxF; Noble->Palenzuela
and Palenzuela-only con2prim; shocks along x, y, z; the periodic Alfven wave
with RK2 and the algebraic gauge; the 3-D PPM sound wave; the 2-D magnetic
rotor; the 2-D cylindrical blast with upwind CT, z_vec and the entropy fix)
and is evolved on the single-patch driver; every primitive on the
golden 1-D cuts must agree with the reference's TSV output within the
reference's tolerance (AsterX/test/test.ccl: ABSTOL = RELTOL = 1e-8).
"""
import pytest

import oracle
from tsv_cases import CASES, compare, run_case


@pytest.mark.parametrize("name", sorted(CASES))
def test_tsv_golden(name):
    res, gold = run_case(name, oracle.backend())
    worst, where = compare(res, gold)
    assert worst <= 0.0, (worst, where)
