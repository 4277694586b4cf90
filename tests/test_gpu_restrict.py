"""RestrictFunction and RestrictToAllGrids on the GPU (pmg_restrict_function,
pmg_restrict_to_all_grids; MGComponents3D.C:1509-1755, NSE_MultiGrid.C:391-431) against
the oracle's restatement (pinned in tests/test_oracle_restrict.py), through the C ABI:
1e-14 relative per level; a coarse velocity prolongated and restricted back is itself.
"""
import ctypes as C

import numpy as np
import pytest

from conftest import params, rel

pytestmark = pytest.mark.gpu


def test_restrict_function_matches_oracle(oseen16):
    import pmg_oracle
    from to_merge_master_parmoon_amd.multigrid import build_hierarchy
    mg = build_hierarchy(oseen16, params())
    orc = pmg_oracle.Oracle(params(), oseen16.descs, refactor=0)
    lib, ctx = mg._lib, mg._ctx
    L = oseen16.n_levels
    u = np.random.default_rng(800).standard_normal(mg.n_dof[-1])
    ptr = lambda a: a.ctypes.data_as(C.c_void_p)
    # level by level
    for lvl in range(L - 1, 0, -1):
        uf = np.random.default_rng(801 + lvl).standard_normal(mg.n_dof[lvl])
        uc = np.full(mg.n_dof[lvl - 1], 7.0)
        mg._check(lib.pmg_restrict_function(ctx, lvl, ptr(uf), ptr(uc), 0), "restrict_function")
        want = orc.restrict_function(lvl, uf)
        n3 = 3 * oseen16.descs[lvl - 1].n_udof
        assert rel(uc[:n3], want[:n3]) <= 1e-14
        assert np.all(uc[n3:] == 7.0)   # pressure untouched
        # prolongated coarse velocity restricts back to itself
        vc = np.random.default_rng(810 + lvl).standard_normal(mg.n_dof[lvl - 1])
        vf = mg.prolongate(lvl, vc)
        back = np.zeros_like(vc)
        mg._check(lib.pmg_restrict_function(ctx, lvl, ptr(vf), ptr(back), 0), "restrict_function")
        assert rel(back[:n3], vc[:n3]) <= 1e-13
    # all levels at once, device vectors
    from to_merge_master_parmoon_amd.multigrid import DeviceVector
    dv = [DeviceVector(mg, mg.n_dof[l], np.zeros(mg.n_dof[l])) for l in range(L - 1)]
    top = DeviceVector(mg, mg.n_dof[-1], u)
    arr = (C.c_void_p * (L - 1))(*[v.ptr for v in dv])
    mg._check(lib.pmg_restrict_to_all_grids(ctx, top.ptr, arr, 1), "restrict_to_all_grids")
    cur = u
    for lvl in range(L - 1, 0, -1):
        want = orc.restrict_function(lvl, cur)
        got = dv[lvl - 1].download()
        n3 = 3 * oseen16.descs[lvl - 1].n_udof
        print(f"level {lvl - 1}: {rel(got[:n3], want[:n3]):.2e}", flush=True)
        assert rel(got[:n3], want[:n3]) <= 1e-14
        cur = got
    mg.close()
    orc.close()
