"""The pressure-dof order of the nodal Vanka as level sets (pmg_nodal_schedule, host code of
libpmg).

TNSE_MGLevel4::NodalVanka (src/FE/NSE_MGLevel4.C:4098-4549, smoothers 3 and 4) visits the
pressure dofs in order and updates u after every patch (patch i = pressure dof i and the
velocity dofs of B row i).  The GPU runs the patches as level sets.  Pins, on small meshes
against the dense operator pattern:
  * every pair of patches whose rows read each other's dofs keeps its order (set of the
    earlier dof < set of the later one), and no such pair shares a set;
  * sets are as early as possible, so a wrong closure or an off-by-one changes the count;
  * a sweep over the sets (one residual per set, all its patches solved from it) equals the
    dense sequential nodal sweep (1e-12) for both local solves, while one set holding every
    patch (a Jacobi sweep) differs from it.
"""
import ctypes as C

import numpy as np
import pytest

from conftest import rel
from dense_ref import dense_operator
from test_oracle_nodal import _nodal_dense


def _schedule(prob, lvl):
    from to_merge_master_parmoon_amd import _abi
    lib = _abi.load_pmg()
    d = prob.descs[lvl]
    a = prob.arrays(lvl)
    keep = [np.ascontiguousarray(a[k], dtype=np.int32) for k in
            ("a_rowptr", "a_kcol", "bt_rowptr", "bt_kcol", "b_rowptr", "b_kcol")]
    ptr = np.empty(d.n_pdof + 1, dtype=np.int32)
    order = np.empty(d.n_pdof, dtype=np.int32)
    ns = C.c_int()
    st = lib.pmg_nodal_schedule(d.n_udof, d.n_pdof, *[k.ctypes.data_as(_abi.c_int_p) for k in keep], C.byref(ns),
                                ptr.ctypes.data_as(_abi.c_int_p), order.ctypes.data_as(_abi.c_int_p))
    assert st == 0
    return [order[ptr[k]:ptr[k + 1]] for k in range(ns.value)]


def _patch(prob, lvl, i):
    d = prob.descs[lvl]
    a = prob.arrays(lvl)
    N = d.n_udof
    ud = a["b_kcol"][a["b_rowptr"][i]:a["b_rowptr"][i + 1]].astype(np.int64)
    return np.concatenate([ud, ud + N, ud + 2 * N, [3 * N + i]])


def _coupling(prob, lvl):
    """patches i, j coupled iff a row of one has a stored entry in a column of the other"""
    K = dense_operator(prob, lvl)
    pat = K != 0.0
    d = prob.descs[lvl]
    a = prob.arrays(lvl)
    # stored zeros count as entries: add the CSR pattern explicitly
    N = d.n_udof
    rows = np.repeat(np.arange(N), np.diff(a["a_rowptr"]))
    for r in range(3):
        for c in range(3):
            pat[r * N + rows, c * N + a["a_kcol"]] = True
    rb = np.repeat(np.arange(N), np.diff(a["bt_rowptr"]))
    for r in range(3):
        pat[r * N + rb, 3 * N + a["bt_kcol"]] = True
    rp = np.repeat(np.arange(d.n_pdof), np.diff(a["b_rowptr"]))
    for c in range(3):
        pat[3 * N + rp, c * N + a["b_kcol"]] = True
    idx = [_patch(prob, lvl, i) for i in range(d.n_pdof)]
    n = d.n_pdof
    cpl = np.zeros((n, n), dtype=bool)
    for j in range(n):
        cols = np.any(pat[idx[j]], axis=0)
        for i in range(n):
            cpl[i, j] = cols[idx[i]].any()
    return cpl | cpl.T


@pytest.mark.parametrize("n,faces", [(2, 63), (3, 21)])
def test_nodal_schedule_orders_coupled_patches(n, faces):
    from to_merge_master_parmoon_amd import problem
    prob = problem.oseen3d(n, coarsest=n, dirichlet_faces=faces)
    sets = _schedule(prob, 0)
    npd = prob.descs[0].n_pdof
    where = np.empty(npd, dtype=int)
    for k, s in enumerate(sets):
        assert np.all(np.diff(s) > 0)
        where[s] = k
    assert sorted(np.concatenate(sets).tolist()) == list(range(npd))
    cpl = _coupling(prob, 0)
    for j in range(npd):
        earlier = [i for i in range(j) if cpl[i, j]]
        assert all(where[i] < where[j] for i in earlier)
        assert where[j] == (max(where[i] for i in earlier) + 1 if earlier else 0)
    print(f"{n}^3 cells, {npd} pressure dofs: {len(sets)} sets", flush=True)


def _sweep_sets(prob, K, sets, u0, f, damp, alpha, smoother):
    d = prob.descs[0]
    N, NA = d.n_udof, d.n_active
    u = u0.copy()
    for c in range(3):
        u[c * N + NA:(c + 1) * N] = f[c * N + NA:(c + 1) * N]
    uold = u.copy()
    for s in sets:
        r = f - K @ u
        upd = []
        for i in s:
            idx = _patch(prob, 0, i)
            Kl = K[np.ix_(idx, idx)]
            if smoother == 4:
                dl = np.linalg.solve(Kl, r[idx])
            else:
                nu_ = (idx.size - 1) // 3
                Dg = np.tile(np.diag(Kl)[:nu_], 3)
                B, BT, rl = Kl[-1, :-1], Kl[:-1, -1], r[idx]
                dp = (rl[-1] - np.sum(B * rl[:-1] / Dg)) / (-np.sum(B * BT / Dg))
                dl = np.concatenate([(rl[:-1] - BT * dp) / Dg, [dp]])
            upd.append((idx, dl))
        for idx, dl in upd:
            u[idx] += damp * dl
    u = uold + alpha * (u - uold)
    for c in range(3):
        u[c * N + NA:(c + 1) * N] = f[c * N + NA:(c + 1) * N]
    return u


@pytest.mark.parametrize("smoother", [3, 4])
def test_set_sweep_is_sequential_nodal_sweep(smoother):
    from to_merge_master_parmoon_amd import problem
    prob = problem.oseen3d(3, coarsest=3)
    K = dense_operator(prob, 0)
    rng = np.random.default_rng(40 + smoother)
    u0, f = rng.standard_normal(K.shape[0]), rng.standard_normal(K.shape[0])
    want = _nodal_dense(prob, 0, K, u0, f, 0.8, 0.9, smoother)
    got = _sweep_sets(prob, K, _schedule(prob, 0), u0, f, 0.8, 0.9, smoother)
    assert rel(got, want) <= 1e-12
    jacobi = _sweep_sets(prob, K, [np.arange(prob.descs[0].n_pdof)], u0, f, 0.8, 0.9, smoother)
    assert rel(jacobi, want) > 1e-3
