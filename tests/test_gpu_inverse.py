"""Explicit patch inverses of the mixed-precision smoother, through the C ABI.

With patch_inverse = 1 (the mixed-precision default) every Vanka patch is inverted once per
operator and each local solve is a product with the stored data: with the shared velocity
block (the Oseen problems here) in the block form (A^{-1} of the 27 x 27 velocity block,
W_c = A^{-1} BT_c, B_c and the inverse of the 4 x 4 Schur complement), otherwise the full
85 x 85 inverse by FP32 Gauss-Jordan with the row pivots of dgetf2 (covered on the
nine-block operator by tests/test_gpu_blocks.py); patch_inverse = 0 keeps the LU factors
and triangular sweeps.  Pins:
  * one FP32 sweep with inverses against the FP64 oracle's sweep (dgetrf/dgetrs per patch)
    to FP32 accuracy, and with BF16-stored inverses to BF16 accuracy, on two levels; a
    wrong row of the inverse, a transposed inverse or a missing pivot permutation gives
    an O(1) difference;
  * the V-cycle and the FGMRES solve agree between the inverse and the LU forms (same
    iteration count, solutions within the condition number times the FGMRES tolerance,
    modulo the constant pressure).
"""
import numpy as np
import pytest

from conftest import params, rel, unpin_pressure

pytestmark = pytest.mark.gpu


def _hier(prob, **kw):
    from to_merge_master_parmoon_amd.multigrid import build_hierarchy
    return build_hierarchy(prob, params(**kw))


@pytest.mark.parametrize("precision,tol", [(1, 2e-5), (2, 2e-2)])
@pytest.mark.parametrize("lvl", [1, 2])
def test_inverse_sweep_matches_oracle(oseen16, precision, tol, lvl):
    import pmg_oracle
    mg = _hier(oseen16, precision=precision, patch_inverse=1)
    orc = pmg_oracle.Oracle(params(), oseen16.descs, refactor=0)
    n = mg.n_dof[lvl]
    rng = np.random.default_rng(60 + lvl)
    u0, f = rng.standard_normal(n), rng.standard_normal(n)
    u_g = mg.cell_vanka(lvl, u0.copy(), f, 1)
    u_c = u0.copy()
    assert orc.cell_vanka(lvl, u_c, f, 1) == 0
    e = rel(u_g - u0, u_c - u0)
    print("precision", precision, "level", lvl, "sweep rel", e, flush=True)
    assert e <= tol
    mg.close()


@pytest.mark.parametrize("precision", [1, 2])
def test_inverse_and_lu_forms_agree(oseen16, precision):
    mi = _hier(oseen16, precision=precision, patch_inverse=1)
    ml = _hier(oseen16, precision=precision, patch_inverse=0)
    v = np.random.default_rng(70).standard_normal(mi.n_dof[-1])
    e = rel(mi.cycle(v), ml.cycle(v))
    print("precision", precision, "V-cycle rel", e, flush=True)
    assert e <= (1e-4 if precision == 1 else 3e-2)
    f = oseen16.rhs(-1).copy()
    xi, xl = np.zeros_like(f), np.zeros_like(f)
    it_i, _, st_i = mi.fgmres(xi, f)
    it_l, _, st_l = ml.fgmres(xl, f)
    print("precision", precision, "FGMRES iterations", it_i, it_l, flush=True)
    assert st_i == 0 and st_l == 0 and abs(it_i - it_l) <= 1
    # both reach the relative residual 1e-10 with different preconditioners: their
    # solutions differ by up to the condition number times that, and (Dirichlet data on
    # the whole boundary) by a constant pressure, which each preconditioner leaves where
    # it puts it
    nu = oseen16.descs[-1].n_udof
    assert rel(unpin_pressure(xi, nu), unpin_pressure(xl, nu)) <= 1e-7
    mi.close()
    ml.close()


def test_block_form_on_large_colors():
    """Colors of more than 2048 patches (the 32^3 level) take the block form: one FP32 sweep
    there equals the sweep with LU factors (patch_inverse = 0, pinned against the oracle
    above) and the one with full 85 x 85 inverses on the nine-block kernels
    (shared_block = 0) to FP32 accuracy."""
    from to_merge_master_parmoon_amd import problem
    P = problem.oseen3d(32)
    lvl = 3
    res = {}
    for name, kw in (("block", dict(patch_inverse=1)), ("lu", dict(patch_inverse=0)),
                     ("full", dict(patch_inverse=1, shared_block=0))):
        mg = _hier(P, precision=1, **kw)
        assert mg.level_info(lvl)["max_cells_per_color"] > 2048
        n = mg.n_dof[lvl]
        rng = np.random.default_rng(80)
        u0, f = rng.standard_normal(n), rng.standard_normal(n)
        res[name] = mg.cell_vanka(lvl, u0.copy(), f, 1) - u0
        mg.close()
    e_lu, e_full = rel(res["block"], res["lu"]), rel(res["block"], res["full"])
    print("32^3 sweep: block vs LU", e_lu, "block vs full inverse", e_full, flush=True)
    assert e_lu <= 2e-5 and e_full <= 2e-5
