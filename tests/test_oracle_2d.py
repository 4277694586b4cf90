"""The 2D twins of the oracle (oracle/pmg_oracle2d.cpp, Q2/P1disc quadrilaterals, 21-dof
Vanka patches): the steady Stokes configuration on the unit square (64 x 64 cells from a 4 x 4
coarse grid, 5 levels, V(2,2) cell Vanka with damping 0.8, FGMRES(20), rtol 1e-10).

Pins, each chosen so that a plausible slip (a transposed block, a wrong sign, a wrong
stride of IntoL20, a wrong child map, a missing damping) fails:
  * operator: defect of the CSR blocks against a dense long-double assembly of the same
    arrays with the Dirichlet-row and BT rules of CoupledDefect 2D; the discrete solution
    of a quadratic velocity / linear pressure pair is its interpolant (closed form);
    convergence orders against the manufactured solution u = curl(sin^2 sin^2),
    p = x^3 + y^3 - 1/2 (velocity nodes O(h^4), pressure means O(h^4), slopes O(h^3));
  * transfers: P maps the interpolant of any global Q2 / linear function on the coarse
    level to the one on the fine level, R = P^T, and the P1disc child map has the closed
    form c0' = c0 +- 3/2 c1 +- 3/2 c2, c1' = c1 / 2, c2' = c2 / 2;
  * Vanka: a colored sweep equals a dense restatement (per color: residual from the
    pre-color state, dense 21 x 21 submatrix solves, damped update); a solution of K u = f
    is a fixed point of a sweep;
  * coarse solve: a dense solve with pressure dof 0 pinned;
  * coloring: 4 colors (2 x 2 parity), no two same-color cells share a dof.
"""
import numpy as np
import pytest

from conftest import params


def _o2d(coarse=4, n_levels=2, nu=1.0, kind=0, **kw):
    import pmg_oracle
    base = dict(sc_gmres_restart=20, sc_smooth_damp_factor_coarse_saddle=0.8)
    base.update(kw)
    return pmg_oracle.Oracle2D(params(**base), coarse=coarse, n_levels=n_levels, nu=nu, rhs_kind=kind)


def _dense(o, lvl):
    """K from the CSR arrays; Dirichlet velocity rows only A11/A22, no BT (CoupledDefect 2D)."""
    c = o.csr(lvl)
    s = o.sizes(lvl)
    nu, na, npd = s["n_udof"], s["n_active"], s["n_pdof"]
    n = 2 * nu + npd
    K = np.zeros((n, n), dtype=np.longdouble)
    for r in range(nu):
        for k in range(c["a_rowptr"][r], c["a_rowptr"][r + 1]):
            col = c["a_kcol"][k]
            K[r, col] += c["a"][0][k]
            K[nu + r, nu + col] += c["a"][3][k]
            if r < na:
                K[r, nu + col] += c["a"][1][k]
                K[nu + r, col] += c["a"][2][k]
        if r < na:
            for k in range(c["bt_rowptr"][r], c["bt_rowptr"][r + 1]):
                col = 2 * nu + c["bt_kcol"][k]
                K[r, col] += c["bt"][0][k]
                K[nu + r, col] += c["bt"][1][k]
    for r in range(npd):
        for k in range(c["b_rowptr"][r], c["b_rowptr"][r + 1]):
            col = c["b_kcol"][k]
            K[2 * nu + r, col] += c["b"][0][k]
            K[2 * nu + r, nu + col] += c["b"][1][k]
    return K


def test_defect_matches_dense_product():
    o = _o2d(coarse=2, n_levels=2, internal_project_pressure=0)
    K = _dense(o, 1)
    rng = np.random.default_rng(5)
    x, f = rng.standard_normal(K.shape[0]), rng.standard_normal(K.shape[0])
    s = o.sizes(1)
    d, _ = o.defect(1, x, f)
    ref = (f.astype(np.longdouble) - K @ x.astype(np.longdouble)).astype(np.float64)
    ref[s["n_active"]:s["n_udof"]] = 0
    ref[s["n_udof"] + s["n_active"]:2 * s["n_udof"]] = 0
    assert np.abs(d - ref).max() <= 1e-13 * np.abs(ref).max()
    assert np.allclose(o.matvec(1, x), (K @ x.astype(np.longdouble)).astype(np.float64), rtol=0, atol=1e-12)


def test_defect_projection_stride_three():
    o = _o2d(coarse=2, n_levels=1)
    n = o.n_dof(0)
    nu = o.sizes(0)["n_udof"]
    rng = np.random.default_rng(6)
    d, _ = o.defect(0, rng.standard_normal(n), rng.standard_normal(n))
    assert abs(d[2 * nu::3].mean()) < 1e-14        # constants mean-zero
    assert abs(d[2 * nu + 1::3].mean()) > 1e-6     # slope coefficients untouched


def test_quadlin_is_reproduced():
    o = _o2d(coarse=2, n_levels=2, kind=1, sc_lin_red_factor_saddle=1e-13)
    f = o.rhs(1)
    x = np.zeros_like(f)
    it, res, st = o.fgmres(x, f)
    ex = o.exact(1)
    nu = o.sizes(1)["n_udof"]
    assert st == 0
    assert np.abs(x[:2 * nu] - ex[:2 * nu]).max() <= 1e-10
    pc = x[2 * nu::3] - x[2 * nu::3].mean()
    pe = ex[2 * nu::3] - ex[2 * nu::3].mean()
    assert np.abs(pc - pe).max() <= 1e-9
    assert np.abs(x[2 * nu + 1::3] - ex[2 * nu + 1::3]).max() <= 1e-9


def test_stokes_config_converges_with_manufactured_orders():
    errs = []
    for levels in (3, 4, 5):
        o = _o2d(coarse=4, n_levels=levels)
        f = o.rhs(levels - 1)
        x = np.zeros_like(f)
        it, res, st = o.fgmres(x, f)
        assert st == 0 and it <= 12
        ex = o.exact(levels - 1)
        nu = o.sizes(levels - 1)["n_udof"]
        eu = np.abs(x[:2 * nu] - ex[:2 * nu]).max()
        ep = np.abs((x[2 * nu::3] - x[2 * nu::3].mean()) - (ex[2 * nu::3] - ex[2 * nu::3].mean())).max()
        eg = np.abs(x[2 * nu + 1::3] - ex[2 * nu + 1::3]).max()
        errs.append((eu, ep, eg))
    for (a, b) in zip(errs, errs[1:]):
        assert 12 < a[0] / b[0] < 20     # velocity nodes: O(h^4)
        assert 10 < a[1] / b[1] < 20     # pressure cell means: O(h^4)
        assert 6 < a[2] / b[2] < 10      # pressure slopes: O(h^3)
    assert errs[-1][0] < 1e-6            # 64 x 64: the configuration's finest level


def test_transfers_reproduce_global_functions_and_are_adjoint():
    o = _o2d(coarse=2, n_levels=2)
    xy_c, xy_f = o.dof_coords(0), o.dof_coords(1)
    q = lambda p: 1 + 2 * p[:, 0] - p[:, 1] + 3 * p[:, 0] * p[:, 1] - p[:, 0] ** 2 + 0.5 * p[:, 1] ** 2
    sc, sf = o.sizes(0), o.sizes(1)
    vc = np.zeros(o.n_dof(0))
    vc[:sc["n_udof"]] = q(xy_c)
    vc[sc["n_udof"]:2 * sc["n_udof"]] = 2 * q(xy_c)
    # P1disc: the exact moments of a global linear function on the coarse cells
    okl = _o2d(coarse=2, n_levels=2, kind=1)
    ec, ef = okl.exact(0), okl.exact(1)
    vc[2 * sc["n_udof"]:] = ec[2 * sc["n_udof"]:]
    vf = o.prolongate(1, vc)
    assert np.abs(vf[:sf["n_udof"]] - q(xy_f)).max() < 1e-13
    assert np.abs(vf[sf["n_udof"]:2 * sf["n_udof"]] - 2 * q(xy_f)).max() < 1e-13
    assert np.abs(vf[2 * sf["n_udof"]:] - ef[2 * sf["n_udof"]:]).max() < 1e-13
    rng = np.random.default_rng(3)
    d, w = rng.standard_normal(o.n_dof(1)), rng.standard_normal(o.n_dof(0))
    lhs = o.restrict_defect(1, d) @ w
    rhs_ = d @ o.prolongate(1, w)
    assert abs(lhs - rhs_) <= 1e-13 * abs(lhs)


def test_p1disc_child_map_closed_form():
    o = _o2d(coarse=1, n_levels=2)
    s0, s1 = o.sizes(0), o.sizes(1)
    for k in range(3):
        vc = np.zeros(o.n_dof(0))
        vc[2 * s0["n_udof"] + k] = 1.0
        vf = o.prolongate(1, vc)[2 * s1["n_udof"]:].reshape(4, 3)
        for ch in range(4):
            sx = 1.5 if ch & 1 else -1.5
            sy = 1.5 if ch >> 1 else -1.5
            expect = {0: [1, 0, 0], 1: [sx, 0.5, 0], 2: [sy, 0, 0.5]}[k]
            assert np.allclose(vf[ch], expect, atol=1e-14), (k, ch, vf[ch])


def test_colored_vanka_equals_dense_restatement():
    o = _o2d(coarse=2, n_levels=2, internal_project_pressure=0)
    lvl = 1
    K = _dense(o, lvl).astype(np.float64)
    s = o.sizes(lvl)
    nu, na = s["n_udof"], s["n_active"]
    ug, pg = o.cell_dofs(lvl)
    rng = np.random.default_rng(9)
    n = K.shape[0]
    u0, f = rng.standard_normal(n), rng.standard_normal(n)
    u_o, info = o.cell_vanka(lvl, u0, f)
    assert info == 0
    import pmg_oracle
    nc, ptr, reo = pmg_oracle.color_cells(ug)
    assert nc == 4
    u = u0.copy()
    dirich = np.r_[np.arange(na, nu), nu + np.arange(na, nu)]
    u[dirich] = f[dirich]
    for c in range(nc):
        cells = reo[ptr[c]:ptr[c + 1]]
        r = f - K @ u
        du = np.zeros(n)
        for cell in cells:
            idx = np.r_[ug[cell], nu + ug[cell], 2 * nu + pg[cell]]
            Kl = K[np.ix_(idx, idx)].copy()
            # Dirichlet rows of the patch: no coupling to the other component, no BT
            for j, g in enumerate(ug[cell]):
                if g >= na:
                    Kl[j, 9:] = 0
                    Kl[9 + j, :9] = 0
                    Kl[9 + j, 18:] = 0
            du[idx] += 0.8 * np.linalg.solve(Kl, r[idx])
        u += du
    u[dirich] = f[dirich]
    assert np.abs(u_o - u).max() <= 1e-12 * np.abs(u).max()


def test_sweep_fixed_point():
    """A sweep leaves a solution of K u = f unchanged (every local residual is zero)."""
    o = _o2d(coarse=2, n_levels=2, internal_project_pressure=0)
    K = _dense(o, 1).astype(np.float64)
    x0 = np.random.default_rng(4).standard_normal(K.shape[0])
    f = K @ x0                      # Dirichlet rows are identity rows: f_D = x0_D
    u, info = o.cell_vanka(1, x0.copy(), f)
    assert info == 0
    assert np.abs(u - x0).max() <= 1e-11 * np.abs(x0).max()


def test_coarse_solve_is_pinned_dense_solve():
    o = _o2d(coarse=2, n_levels=1)
    K = _dense(o, 0).astype(np.float64)
    nu = o.sizes(0)["n_udof"]
    f = np.random.default_rng(2).standard_normal(K.shape[0])
    u, info = o.solve_exact(f)
    Kp, fp = K.copy(), f.copy()
    Kp[2 * nu, :] = 0
    Kp[2 * nu, 2 * nu] = 1
    fp[2 * nu] = 0
    assert info == 0
    assert np.allclose(u, np.linalg.solve(Kp, fp), rtol=0, atol=1e-11 * np.abs(u).max())
