"""The two forms of the velocity rows on the GPU, through the C ABI.

When A11 = A22 = A33 and the coupling blocks A12..A32 vanish on the active rows (the
Galerkin Oseen form of NSType4Galerkin3D, src/FE/NSE3D_FixPo.C:1919-2071) the library
streams one value array for the three velocity components instead of nine.  Pins:
  * the form is detected on the generated Oseen hierarchy, and not on a copy whose active
    rows carry coupling blocks and a scaled A22 (the general type-4 operator);
  * with coupling blocks the nine-block kernels match the FP64 oracle as tightly as the
    shared form does on the Oseen problem (defect, matvec, sweep, V-cycle, FGMRES);
  * the two forms give the same V-cycle on the Oseen problem (shared_block = 0 forces the
    nine blocks);
  * re-assembly on the device re-checks the form: a coupled level becomes shared.
"""
import numpy as np
import pytest

from conftest import params, rel, unpin_pressure

pytestmark = pytest.mark.gpu


def _coupled(n=16):
    """The Oseen hierarchy with A12 = 0.05 A11, A21 = -0.05 A11, A13 = 0.03 A11,
    A31 = -0.03 A11, A23 = 0.02 A11, A32 = -0.02 A11 and A22 = 1.1 A11 on the active rows
    (Dirichlet rows keep their identity diagonal blocks)."""
    from to_merge_master_parmoon_amd import problem
    P = problem.oseen3d(n)
    for lvl in range(P.n_levels):
        arr = P.arrays(lvl)
        d = P.descs[lvl]
        ne = int(arr["a_rowptr"][d.n_active])
        a = arr["a"]
        a0 = a[0][:ne].copy()
        for b, f in ((1, 0.05), (3, -0.05), (2, 0.03), (6, -0.03), (5, 0.02), (7, -0.02)):
            a[b][:ne] = f * a0
        a[4][:ne] = 1.1 * a0
    return P


@pytest.fixture(scope="module")
def coupled16():
    return _coupled(16)


def _hier(prob, **kw):
    from to_merge_master_parmoon_amd.multigrid import build_hierarchy
    return build_hierarchy(prob, params(**kw))


def _oracle(prob, **kw):
    import pmg_oracle
    return pmg_oracle.Oracle(params(**kw), prob.descs, refactor=0)


def _rand(n, seed):
    return np.random.default_rng(seed).standard_normal(n)


def test_form_detection(oseen16, coupled16):
    mg = _hier(oseen16)
    assert all(mg.shared_block(l) for l in range(mg.n_levels))
    mg.close()
    mg = _hier(oseen16, shared_block=0)
    assert not any(mg.shared_block(l) for l in range(mg.n_levels))
    mg.close()
    mg = _hier(coupled16)
    assert not any(mg.shared_block(l) for l in range(mg.n_levels))
    mg.close()


@pytest.mark.parametrize("lvl", [1, 2])
def test_nine_blocks_match_oracle(coupled16, lvl):
    mg, orc = _hier(coupled16), _oracle(coupled16)
    n = mg.n_dof[lvl]
    u, f = _rand(n, 21), _rand(n, 22)
    d_g, r_g = mg.defect(lvl, u, f)
    d_c, r_c = orc.defect(lvl, u, f)
    assert rel(d_g, d_c) <= 1e-13
    assert abs(r_g - r_c) <= 1e-13 * r_c
    assert rel(mg.matvec(lvl, u), orc.matvec(lvl, u)) <= 1e-13
    u_g = mg.cell_vanka(lvl, u.copy(), f, 2)
    u_c = u.copy()
    assert orc.cell_vanka(lvl, u_c, f, 2) == 0
    assert rel(u_g - u, u_c - u) <= 1e-13
    mg.close()


def test_nine_blocks_vcycle_and_fgmres(coupled16):
    mg, orc = _hier(coupled16), _oracle(coupled16)
    v = _rand(mg.n_dof[-1], 23)
    z_g = mg.cycle(v)
    z_c, st = orc.vcycle(v)
    assert st == 0
    assert rel(z_g, z_c) <= 1e-12
    f = coupled16.rhs(-1).copy()
    x_g, x_c = np.zeros_like(f), np.zeros_like(f)
    it_g, _, st_g = mg.fgmres(x_g, f)
    it_c, _, st_c = orc.fgmres(x_c, f)
    print("coupled 16^3 FGMRES iterations", it_g, it_c, flush=True)
    assert st_g == 0 and st_c == 0 and abs(it_g - it_c) <= 1
    assert rel(x_g, x_c) <= 1e-10
    mg.close()


@pytest.mark.parametrize("precision", [0, 2])
def test_forms_agree_on_oseen(oseen16, precision):
    mg1 = _hier(oseen16, precision=precision)
    mg0 = _hier(oseen16, precision=precision, shared_block=0)
    # FP64: same arithmetic up to the summation order; BF16: the two forms store different
    # local data (block form vs full patch inverse) in BF16, so they agree to its accuracy
    tol = 1e-12 if precision == 0 else 2e-3
    for k in range(2):
        v = _rand(mg1.n_dof[-1], 40 + k)
        assert rel(mg1.cycle(v), mg0.cycle(v)) <= tol
    f = oseen16.rhs(-1).copy()
    x1, x0 = np.zeros_like(f), np.zeros_like(f)
    it1 = mg1.fgmres(x1, f)[0]
    it0 = mg0.fgmres(x0, f)[0]
    # both meet the relative residual 1e-10; the constant pressure is left where each
    # preconditioner puts it
    nu = oseen16.descs[-1].n_udof
    assert abs(it1 - it0) <= 1 and rel(unpin_pressure(x1, nu), unpin_pressure(x0, nu)) <= 1e-7
    mg1.close()
    mg0.close()


def test_reassembly_rechecks_the_form():
    from to_merge_master_parmoon_amd.multigrid import build_hierarchy
    P = _coupled(8)
    mg = build_hierarchy(P, params())
    for l in range(P.n_levels):
        mg.level_geometry(l, P.cell_vertices(l))
    assert not mg.shared_block(1)
    x = _rand(mg.n_dof[1], 50)
    mg.assemble_oseen(1, P.conv_quad(1), 1e-2, 0.0)
    assert mg.shared_block(1) and not mg.shared_block(0)
    # the re-assembled operator is the generator's uncoupled one
    from to_merge_master_parmoon_amd import problem
    Q = problem.oseen3d(8)
    mq = build_hierarchy(Q, params())
    assert rel(mg.matvec(1, x), mq.matvec(1, x)) <= 1e-13
    mg.close()
    mq.close()
