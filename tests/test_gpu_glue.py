"""The V-cycle glue entry points and the damping / smoothing-count branches of the GPU
path, against the FP64 oracle through the C ABI.

* pmg_correct_nodes (u_D := 0), pmg_set_dirichlet (u_D := f_D): bit-exact;
  pmg_correct_defect (rows >= N_Active := 0, IntoL20 of the pressure constants): the
  velocity part bit-exact, the pressure within 1e-15 relative (the mean is a reduction
  whose summation order differs); pmg_update (u += beta d, beta = 0.7): within the
  rounding of the product and of the sum per entry (the device forms beta d + u with one
  fused multiply-add, the oracle rounds the product first).
* sweeps and V-cycles with the three damping knobs away from 1 (per-cell damp 0.8,
  global relaxation alpha 0.9, correction beta 0.7) and with (nu1, nu2) in
  {(0,2), (2,0), (1,1), (2,2)}: FP64 within 1e-13 (sweep) / 1e-12 (V-cycle), FP32 within
  2e-5 / 1e-4 of the oracle.  These run the relaxation kernel, the post == 0 branch that
  keeps the CorrectNodes calls, and the closing SetDirichletNodes of a single sweep.
"""
import numpy as np
import pytest

from conftest import params, rel

pytestmark = pytest.mark.gpu

DAMP = dict(sc_smooth_damp_factor_coarse_saddle=0.8, sc_smooth_damp_factor_saddle=0.9,
            sc_gmg_damp_factor_saddle=0.7)


def _hier(prob, **kw):
    from to_merge_master_parmoon_amd.multigrid import build_hierarchy
    return build_hierarchy(prob, params(**kw))


def _oracle(prob, **kw):
    import pmg_oracle
    return pmg_oracle.Oracle(params(**kw), prob.descs, refactor=0)


@pytest.mark.parametrize("lvl", [0, 1, 2])
def test_glue_entry_points(oseen16, lvl):
    mg, orc = _hier(oseen16, **DAMP), _oracle(oseen16, **DAMP)
    n = mg.n_dof[lvl]
    d = oseen16.descs[lvl]
    rng = np.random.default_rng(400 + lvl)
    u, f, dv = rng.standard_normal(n), rng.standard_normal(n), rng.standard_normal(n)

    a, b = mg.correct_nodes(lvl, u.copy()), u.copy()
    orc.correct_nodes(lvl, b)
    assert np.array_equal(a, b)
    assert not np.array_equal(a, u)   # the call changed something

    a, b = mg.set_dirichlet_nodes(lvl, u.copy(), f), u.copy()
    orc.set_dirichlet(lvl, b, f)
    assert np.array_equal(a, b)

    a, b = mg.correct_defect(lvl, dv.copy()), dv.copy()
    orc.correct_defect(lvl, b)
    n3 = 3 * d.n_udof
    assert np.array_equal(a[:n3], b[:n3])
    assert rel(a[n3:], b[n3:]) <= 1e-15
    assert abs(a[n3::4].mean()) <= 1e-14 * np.abs(dv[n3::4]).max()

    a, b = mg.update(lvl, u.copy(), dv), u.copy()
    orc.update(lvl, b, dv)
    # one rounding of the product (unfused form) plus one of the sum
    ulp = np.spacing(np.abs(0.7 * dv)) + np.spacing(np.maximum(np.abs(a), np.abs(b)))
    assert np.all(np.abs(a - b) <= ulp)
    assert rel(a - u, 0.7 * dv) <= 1e-14
    mg.close()
    orc.close()


@pytest.mark.parametrize("precision,tol", [(0, 1e-13), (1, 2e-5)])
@pytest.mark.parametrize("lvl", [1, 2])
def test_damped_sweep(oseen16, precision, tol, lvl):
    mg, orc = _hier(oseen16, precision=precision, **DAMP), _oracle(oseen16, **DAMP)
    n = mg.n_dof[lvl]
    rng = np.random.default_rng(410 + lvl)
    u0, f = rng.standard_normal(n), rng.standard_normal(n)
    for sweeps in (1, 2):
        u_g = mg.cell_vanka(lvl, u0.copy(), f, sweeps)
        u_c = u0.copy()
        assert orc.cell_vanka(lvl, u_c, f, sweeps) == 0
        e = rel(u_g - u0, u_c - u0)
        print(f"precision {precision} level {lvl} {sweeps} damped sweeps: {e:.2e}", flush=True)
        assert e <= tol
    mg.close()
    orc.close()


@pytest.mark.parametrize("precision,tol", [(0, 1e-12), (1, 1e-4)])
@pytest.mark.parametrize("nus", [(0, 2), (2, 0), (1, 1), (2, 2)])
def test_damped_vcycle(oseen16, precision, tol, nus):
    kw = dict(DAMP, sc_pre_smooth_saddle=nus[0], sc_post_smooth_saddle=nus[1])
    mg, orc = _hier(oseen16, precision=precision, **kw), _oracle(oseen16, **kw)
    v = np.random.default_rng(420 + nus[0] + 3 * nus[1]).standard_normal(mg.n_dof[-1])
    z_g = mg.cycle(v)
    z_c, st = orc.vcycle(v)
    assert st == 0
    e = rel(z_g, z_c)
    print(f"precision {precision} V{nus} damped: {e:.2e}", flush=True)
    assert e <= tol
    mg.close()
    orc.close()
