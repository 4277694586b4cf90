"""GPU parity of every hot-path operation against the FP64 oracle, through the C ABI.

Tolerances (FP64 plain path): per-operation relative difference <= 1e-13 (defect,
matvec, transfers, one Vanka sweep), coarse solve <= 1e-12, one V-cycle <= 1e-12,
FGMRES iteration count within +-1 and final solution within 1e-10 relative.  The patch
LU and the coarse LU perform the oracle's fma sequence, so the local solves are
bit-identical; the remaining differences come from the summation order of the CSR rows.
"""
import numpy as np
import pytest
import sys

from conftest import params, rel

pytestmark = pytest.mark.gpu


def _hier(prob, **kw):
    from to_merge_master_parmoon_amd.multigrid import build_hierarchy
    return build_hierarchy(prob, params(**kw))


def _oracle(prob, **kw):
    import pmg_oracle
    return pmg_oracle.Oracle(params(**kw), prob.descs, refactor=0)


def _rand(n, seed):
    return np.random.default_rng(seed).standard_normal(n)


@pytest.mark.parametrize("lvl", [0, 1, 2])
def test_defect_matvec(oseen16, lvl):
    mg, orc = _hier(oseen16), _oracle(oseen16)
    n = mg.n_dof[lvl]
    u, f = _rand(n, 1), _rand(n, 2)
    d_g, r_g = mg.defect(lvl, u, f)
    d_c, r_c = orc.defect(lvl, u, f)
    assert rel(d_g, d_c) <= 1e-13
    assert abs(r_g - r_c) <= 1e-13 * r_c
    assert rel(mg.matvec(lvl, u), orc.matvec(lvl, u)) <= 1e-13
    assert rel(mg.defect_nse4(lvl, u, f), orc.defect_nse4(lvl, u, f)) <= 1e-13


def test_coloring_matches_oracle(oseen8):
    from to_merge_master_parmoon_amd.multigrid import color_cells
    import pmg_oracle
    for lvl in range(oseen8.n_levels):
        a = oseen8.arrays(lvl)
        d = oseen8.descs[lvl]
        nc, ptr, reo = color_cells(d.n_cells, a["u_global"], d.n_udof)
        onc, optr, oreo = pmg_oracle.color_cells(a["u_global"])
        assert nc == onc and np.array_equal(ptr, optr) and np.array_equal(reo, oreo)


@pytest.mark.parametrize("lvl", [1, 2])
def test_transfers(oseen16, lvl):
    mg, orc = _hier(oseen16), _oracle(oseen16)
    d = _rand(mg.n_dof[lvl], 3)
    fc_g = mg.restrict_defect(lvl, d)
    fc_c, d_after = orc.restrict_defect(lvl, d)
    assert np.array_equal(d_after, d)  # the reference restores the fine defect exactly
    assert rel(fc_g, fc_c) <= 1e-14
    uc = _rand(mg.n_dof[lvl - 1], 4)
    assert rel(mg.prolongate(lvl, uc), orc.prolongate(lvl, uc)) <= 1e-14


def test_coarse_solve(oseen8):
    mg, orc = _hier(oseen8), _oracle(oseen8)
    f = _rand(mg.n_dof[0], 5)
    u_g = mg.solve_exact(f)
    u_c, info = orc.solve_exact(f)
    assert info == 0
    assert rel(u_g, u_c) <= 1e-12


@pytest.mark.parametrize("lvl", [1, 2])
def test_vanka_sweep(oseen16, lvl):
    mg, orc = _hier(oseen16), _oracle(oseen16)
    n = mg.n_dof[lvl]
    f = _rand(n, 6)
    u0 = _rand(n, 7)
    u_g = mg.cell_vanka(lvl, u0.copy(), f, 1)
    u_c = u0.copy()
    assert orc.cell_vanka(lvl, u_c, f, 1) == 0
    assert rel(u_g - u0, u_c - u0) <= 1e-13
    # two sweeps
    u_g2 = mg.cell_vanka(lvl, u0.copy(), f, 2)
    u_c2 = u0.copy()
    orc.cell_vanka(lvl, u_c2, f, 2)
    assert rel(u_g2 - u0, u_c2 - u0) <= 1e-13


def test_vcycle(oseen16):
    mg, orc = _hier(oseen16), _oracle(oseen16)
    v = _rand(mg.n_dof[-1], 8)
    z_g = mg.cycle(v)
    z_c, st = orc.vcycle(v)
    assert st == 0
    assert rel(z_g, z_c) <= 1e-12
    # defect of the preconditioned vector, the quantity FGMRES builds on
    assert rel(mg.matvec(2, z_g), orc.matvec(2, z_c)) <= 1e-12


def test_fgmres_oseen16(oseen16):
    mg, orc = _hier(oseen16), _oracle(oseen16)
    f = oseen16.rhs(-1).copy()
    x_g = np.zeros_like(f)
    it_g, res_g, st_g = mg.fgmres(x_g, f)
    x_c = np.zeros_like(f)
    it_c, res_c, st_c = orc.fgmres(x_c, f)
    assert st_g == 0 and st_c == 0
    assert abs(it_g - it_c) <= 1
    assert rel(x_g, x_c) <= 1e-10
    h_g, h_c = mg.fgmres_history(), orc.fgmres_history()
    m = min(len(h_g), len(h_c))
    assert np.all(np.abs(h_g[:m] - h_c[:m]) <= 1e-9 * h_c[0])


def test_quadlin_exact_solution():
    """AnsatzQuadLin: the discrete solution is the interpolant (closed form)."""
    from to_merge_master_parmoon_amd import problem
    P = problem.quadlin(coarse=(2, 2, 2), n_levels=2, nu=1.0)
    mg = _hier(P, sc_lin_red_factor_saddle=1e-13)
    f = P.rhs(-1).copy()
    ex = P.exact(-1)
    x = np.zeros_like(f)
    mg.fgmres(x, f)
    nu = P.descs[-1].n_udof
    assert np.abs(x[:3 * nu] - ex[:3 * nu]).max() <= 1e-9 * np.abs(ex[:3 * nu]).max()
    pd = (x[3 * nu:] - ex[3 * nu:]).reshape(-1, 4)
    assert np.ptp(pd[:, 0]) <= 1e-8 and np.abs(pd[:, 1:]).max() <= 1e-8


def test_errors_are_status_codes(oseen8):
    from to_merge_master_parmoon_amd import _abi
    from to_merge_master_parmoon_amd.multigrid import NSEMultiGrid
    mg = NSEMultiGrid(params())
    mg.add_level(oseen8.descs[0])
    v = np.zeros(mg.n_dof[0])
    with pytest.raises(_abi.PmgError) as e:
        mg.cycle(v)   # before setup
    assert e.value.status == 4
    mg.setup()
    f = np.full(mg.n_dof[0], np.nan)
    with pytest.raises(_abi.PmgError) as e:
        mg.fgmres(np.zeros_like(f), f)
    assert e.value.status == 7


@pytest.mark.parametrize("precision", [0, 2])
def test_vcycle_graph_replays_match_eager(oseen16, precision):
    """Every replay of the captured V-cycle equals the kernel-by-kernel cycle bit for bit,
    with other work of the same context between the replays."""
    mg_e = _hier(oseen16, use_graph=0, precision=precision)
    mg_g = _hier(oseen16, use_graph=1, precision=precision)
    n = mg_g.n_dof[-1]
    for k in range(4):
        v = _rand(n, 300 + k)
        z_e = mg_e.cycle(v)
        mg_g.defect(2, v, v)
        z_g = mg_g.cycle(v)
        assert np.array_equal(z_e, z_g), f"replay {k}"
    f = oseen16.rhs(-1).copy()
    x_e, x_g = np.zeros_like(f), np.zeros_like(f)
    it_e = mg_e.fgmres(x_e, f)[0]
    it_g = mg_g.fgmres(x_g, f)[0]
    assert it_e == it_g and np.array_equal(x_e, x_g)
