"""GPU parity at the benchmark size: the Oseen 32^3 hierarchy (954,947 unknowns, levels
4^3 .. 32^3) against the FP64 oracle, through the C ABI.

* plain FP64 form: one V-cycle <= 1e-12 relative, FGMRES(30) to 1e-10 with the iteration
  count within +-1 and the solution within 1e-10 relative, residual histories within
  1e-9 of the initial residual;
* mixed forms (FP32 smoother with FP32 or BF16 stored patch data, the latter the
  benchmark's default): the FGMRES solution against the oracle's FP64 solution (pressure
  modulo its constant, which the all-Dirichlet operator leaves free), within 1e-9
  relative (the FP64 outer iteration refines to its own 1e-10 tolerance);
* the local solves of the 32^3 level, where colors hold 4096 patches and the mixed forms
  use the block form (A^{-1}, A^{-1} BT_c, B_c, S^{-1}): one FP32 sweep within 2e-5 and one
  BF16 sweep within 2e-2 of the oracle's dgetrf/dgetrs sweep.
"""
import numpy as np
import pytest

from conftest import params, rel, unpin_pressure

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def oseen32():
    from to_merge_master_parmoon_amd import problem
    return problem.oseen3d(32)


@pytest.fixture(scope="module")
def oracle32(oseen32):
    import pmg_oracle
    orc = pmg_oracle.Oracle(params(), oseen32.descs, refactor=0)
    yield orc
    orc.close()


@pytest.fixture(scope="module")
def oracle32_solution(oseen32, oracle32):
    f = oseen32.rhs(-1).copy()
    x = np.zeros_like(f)
    it, res, st = oracle32.fgmres(x, f)
    assert st == 0
    return x, it, oracle32.fgmres_history()


def _hier(prob, **kw):
    from to_merge_master_parmoon_amd.multigrid import build_hierarchy
    return build_hierarchy(prob, params(**kw))


def test_fp64_vcycle_32(oseen32, oracle32):
    mg = _hier(oseen32)
    v = np.random.default_rng(320).standard_normal(mg.n_dof[-1])
    z_g = mg.cycle(v)
    z_c, st = oracle32.vcycle(v)
    assert st == 0
    e = rel(z_g, z_c)
    print(f"32^3 FP64 V-cycle: {e:.2e}", flush=True)
    assert e <= 1e-12
    mg.close()


def test_fp64_fgmres_32(oseen32, oracle32_solution):
    x_c, it_c, h_c = oracle32_solution
    mg = _hier(oseen32)
    f = oseen32.rhs(-1).copy()
    x_g = np.zeros_like(f)
    it_g, res_g, st_g = mg.fgmres(x_g, f)
    h_g = mg.fgmres_history()
    e = rel(x_g, x_c)
    print(f"32^3 FP64 FGMRES: iterations {it_g} (oracle {it_c}), solution {e:.2e}", flush=True)
    assert st_g == 0 and abs(it_g - it_c) <= 1
    assert e <= 1e-10
    m = min(len(h_g), len(h_c))
    assert np.all(np.abs(h_g[:m] - h_c[:m]) <= 1e-9 * h_c[0])
    mg.close()


@pytest.mark.parametrize("precision", [1, 2])
def test_mixed_fgmres_32_vs_oracle(oseen32, oracle32_solution, precision):
    x_c, it_c, _ = oracle32_solution
    mg = _hier(oseen32, precision=precision)
    assert mg.shared_block(3)
    f = oseen32.rhs(-1).copy()
    x = np.zeros_like(f)
    it, res, st = mg.fgmres(x, f)
    nu = oseen32.descs[-1].n_udof
    xu, cu = unpin_pressure(x, nu), unpin_pressure(x_c, nu)
    ev = rel(xu[:3 * nu], cu[:3 * nu])
    ep = rel(xu[3 * nu:], cu[3 * nu:])
    print(f"32^3 precision {precision}: iterations {it} (oracle {it_c}), velocity {ev:.2e}, pressure {ep:.2e}",
          flush=True)
    assert st == 0 and abs(it - it_c) <= 2
    assert ev <= 1e-9 and ep <= 1e-9
    mg.close()


@pytest.mark.parametrize("precision,tol", [(1, 2e-5), (2, 2e-2)])
def test_block_form_sweep_32_vs_oracle(oseen32, oracle32, precision, tol):
    mg = _hier(oseen32, precision=precision)
    lvl = 3
    assert mg.level_info(lvl)["max_cells_per_color"] > 2048 and mg.shared_block(lvl)
    n = mg.n_dof[lvl]
    rng = np.random.default_rng(330 + precision)
    u0, f = rng.standard_normal(n), rng.standard_normal(n)
    u_g = mg.cell_vanka(lvl, u0.copy(), f, 1)
    u_c = u0.copy()
    assert oracle32.cell_vanka(lvl, u_c, f, 1) == 0
    e = rel(u_g - u0, u_c - u0)
    print(f"32^3 block-form sweep, precision {precision}: {e:.2e}", flush=True)
    assert e <= tol
    mg.close()
