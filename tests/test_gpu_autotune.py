"""The launch-shape autotuner (pmg_autotune), through the C ABI.

The tuner times the candidate launch shapes of each smoothed level (Vanka residual work
items per workgroup, per-XCD groups, operator rows per workgroup) and keeps the fastest.
The shapes distribute the same work differently, so after tuning:
  * every smoothed level reports shapes from the candidate sets, the coarsest none;
  * the FP64 V-cycle equals the untuned one to 1e-13 (only the summation order of the
    pressure means can differ), the BF16 form's to 1e-5 (FP32 means), and FGMRES takes the
    same number of iterations.
"""
import numpy as np
import pytest

from conftest import params, rel

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision,tol", [(0, 1e-13), (2, 1e-5)])
def test_autotune_keeps_results(oseen16, precision, tol):
    from to_merge_master_parmoon_amd.multigrid import build_hierarchy
    mg = build_hierarchy(oseen16, params(precision=precision))
    v = np.random.default_rng(1000 + precision).standard_normal(mg.n_dof[-1])
    z0 = mg.cycle(v)
    f = oseen16.rhs(-1).copy()
    x0 = np.zeros_like(f)
    it0, _, st0 = mg.fgmres(x0, f)
    shapes = mg.autotune(10)
    print(f"precision {precision}: {shapes}", flush=True)
    pi = 0 if precision == 0 else 1
    assert shapes[0]["res_items_per_wg"] == [0, 0] and shapes[0]["op_rows_per_wg"] == [0, 0]
    for s in shapes[1:]:
        assert s["res_items_per_wg"][pi] in (16, 32, 64)
        assert s["res_xcd_groups"][pi] in (0, 1)
        assert s["op_rows_per_wg"][pi] in (32, 64, 128, 256)
    z1 = mg.cycle(v)
    x1 = np.zeros_like(f)
    it1, _, st1 = mg.fgmres(x1, f)
    e = rel(z1, z0)
    print(f"precision {precision}: V-cycle after tuning {e:.2e}, FGMRES {it1} vs {it0}", flush=True)
    assert e <= tol
    assert st0 == 0 and st1 == 0 and it1 == it0
    mg.close()


def test_tuning_round_trips_through_json(oseen16, tmp_path):
    """Shapes saved as JSON and applied to a new hierarchy (pmg_set_level_tuning) read back
    unchanged, and the tuned hierarchy's V-cycle equals the untuned one's."""
    from to_merge_master_parmoon_amd.multigrid import build_hierarchy, load_tuning, save_tuning
    a = build_hierarchy(oseen16, params(precision=2))
    shapes = a.autotune(5)
    path = tmp_path / "tuning.json"
    save_tuning(shapes, str(path))
    b = build_hierarchy(oseen16, params(precision=2))
    v = np.random.default_rng(1010).standard_normal(b.n_dof[-1])
    z0 = b.cycle(v)
    load_tuning(b, str(path))
    import ctypes as C
    for l in range(b.n_levels):
        out = (C.c_int * 6)()
        assert b._lib.pmg_level_tuning(b._ctx, l, out) == 0
        s = shapes[l]
        assert list(out) == s["res_items_per_wg"] + s["res_xcd_groups"] + s["op_rows_per_wg"]
    assert rel(b.cycle(v), z0) <= 1e-5
    bad = (C.c_int * 6)(16, 16, 2, 1, 64, 64)
    assert b._lib.pmg_set_level_tuning(b._ctx, 1, bad) == 1
    a.close()
    b.close()
