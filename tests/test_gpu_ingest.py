"""A hierarchy read from level dumps (pmg_level_read) feeds pmg_level_add unchanged: the
GPU solve of the read-back Oseen 16^3 hierarchy equals the solve of the generated one bit
for bit, and both match the oracle (iterations +-1, solution 1e-10)."""
import numpy as np
import pytest

from conftest import params, rel

pytestmark = pytest.mark.gpu


def test_gpu_solve_from_dumps(tmp_path, oseen16):
    import pmg_oracle
    from to_merge_master_parmoon_amd.dump import DumpedHierarchy, write_hierarchy
    from to_merge_master_parmoon_amd.multigrid import build_hierarchy
    write_hierarchy(str(tmp_path), oseen16.descs)
    H = DumpedHierarchy(str(tmp_path))
    f = oseen16.rhs(-1).copy()
    xs = []
    for src in (oseen16, H):
        mg = build_hierarchy(src, params())
        x = np.zeros_like(f)
        it, _, st = mg.fgmres(x, f.copy())
        assert st == 0
        xs.append((it, x))
        mg.close()
    assert xs[0][0] == xs[1][0] and np.array_equal(xs[0][1], xs[1][1])
    orc = pmg_oracle.Oracle(params(), H.descs, refactor=0)
    xo = np.zeros_like(f)
    ito, _, sto = orc.fgmres(xo, f.copy())
    orc.close()
    assert sto == 0 and abs(ito - xs[1][0]) <= 1 and rel(xs[1][1], xo) <= 1e-10
    H.close()
