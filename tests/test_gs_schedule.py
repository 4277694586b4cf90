"""The Gauss-Seidel cell order as level sets (pmg_sequential_schedule, host code of libpmg).

TNSE_MGLevel4::CellVanka's Gauss-Seidel variant (src/FE/NSE_MGLevel4.C:1653-2085) updates
u after every cell in collection order.  The GPU runs it as a colored sweep over level
sets.  Pins, on small meshes against the dense operator pattern:
  * every pair of cells whose patch rows read each other's patch dofs keeps its collection
    order (set of the earlier cell < set of the later one), and no such pair shares a set;
  * sets are as early as possible (each cell sits one set after its latest earlier coupled
    cell), so a wrong closure or an off-by-one changes the set count;
  * a colored sweep over the sets equals the dense Gauss-Seidel sweep and the oracle's
    Gauss-Seidel variant (1e-12), while the ordinary coloring's sweep differs from both.
"""
import ctypes as C

import numpy as np
import pytest

import pmg_oracle
from conftest import params, rel
from dense_ref import dense_operator, patch_indices
from test_oracle_vanka import _cube, _sweep_gs


def _schedule(prob, lvl):
    from to_merge_master_parmoon_amd import _abi
    lib = _abi.load_pmg()
    d = prob.descs[lvl]
    a = prob.arrays(lvl)
    ip = lambda x: np.ascontiguousarray(x, dtype=np.int32).ctypes.data_as(_abi.c_int_p)
    keep = [np.ascontiguousarray(a[k], dtype=np.int32) for k in
            ("u_global", "p_global", "a_rowptr", "a_kcol", "bt_rowptr", "bt_kcol", "b_rowptr", "b_kcol")]
    ptr = np.empty(d.n_cells + 1, dtype=np.int32)
    reo = np.empty(d.n_cells, dtype=np.int32)
    ns = C.c_int()
    st = lib.pmg_sequential_schedule(d.n_cells, *[k.ctypes.data_as(_abi.c_int_p) for k in keep[:2]], d.n_udof,
                                     d.n_pdof, *[k.ctypes.data_as(_abi.c_int_p) for k in keep[2:]], C.byref(ns),
                                     ptr.ctypes.data_as(_abi.c_int_p), reo.ctypes.data_as(_abi.c_int_p))
    assert st == 0
    return [reo[ptr[k]:ptr[k + 1]] for k in range(ns.value)]


def _coupling(prob, lvl):
    """cells i, j coupled iff a row of one patch has a stored entry in a column of the other"""
    d = prob.descs[lvl]
    a = prob.arrays(lvl)
    n = 3 * d.n_udof + d.n_pdof
    pat = np.zeros((n, n), dtype=bool)
    N = d.n_udof
    rows = np.repeat(np.arange(N), np.diff(a["a_rowptr"]))
    for r in range(3):
        pat[r * N + rows, r * N + a["a_kcol"]] = True   # (the coupling blocks share the pattern)
    rb = np.repeat(np.arange(N), np.diff(a["bt_rowptr"]))
    for r in range(3):
        pat[r * N + rb, 3 * N + a["bt_kcol"]] = True
    rp = np.repeat(np.arange(d.n_pdof), np.diff(a["b_rowptr"]))
    for c in range(3):
        pat[3 * N + rp, c * N + a["b_kcol"]] = True
    idx = [patch_indices(prob, lvl, c) for c in range(d.n_cells)]
    cpl = np.zeros((d.n_cells, d.n_cells), dtype=bool)
    for j in range(d.n_cells):
        cols = np.any(pat[idx[j]], axis=0)
        for i in range(d.n_cells):
            cpl[i, j] = cols[idx[i]].any()
    return cpl | cpl.T


@pytest.mark.parametrize("n,faces", [(3, 63), (4, 21)])
def test_schedule_orders_coupled_cells(n, faces):
    prob = _cube(n, faces)
    sets = _schedule(prob, 0)
    nc = prob.descs[0].n_cells
    where = np.empty(nc, dtype=int)
    for k, s in enumerate(sets):
        assert np.all(np.diff(s) > 0)
        where[s] = k
    assert sorted(np.concatenate(sets).tolist()) == list(range(nc))
    cpl = _coupling(prob, 0)
    for j in range(nc):
        earlier = [i for i in range(j) if cpl[i, j]]
        assert all(where[i] < where[j] for i in earlier)
        # as early as possible
        assert where[j] == (max(where[i] for i in earlier) + 1 if earlier else 0)
    print(f"{n}^3 cells: {len(sets)} sets", flush=True)


def test_set_sweep_is_gauss_seidel():
    prob = _cube(3, 63)
    K = dense_operator(prob, 0)
    rng = np.random.default_rng(12)
    u0, f = rng.standard_normal(K.shape[0]), rng.standard_normal(K.shape[0])
    d = prob.descs[0]
    N, NA = d.n_udof, d.n_active

    def sweep_sets(sets, damp):
        u = u0.copy()
        for c in range(3):
            u[c * N + NA:(c + 1) * N] = f[c * N + NA:(c + 1) * N]
        for cells in sets:
            r = f - K @ u
            upd = [(patch_indices(prob, 0, c), None) for c in cells]
            upd = [(ix, np.linalg.solve(K[np.ix_(ix, ix)], r[ix])) for ix, _ in upd]
            for ix, dl in upd:
                u[ix] += damp * dl
        for c in range(3):
            u[c * N + NA:(c + 1) * N] = f[c * N + NA:(c + 1) * N]
        return u

    gs = _sweep_gs(prob, 0, K, u0, f, 0.8)
    via_sets = sweep_sets(_schedule(prob, 0), 0.8)
    assert rel(via_sets, gs) <= 1e-12
    orc = pmg_oracle.Oracle(params(sc_smooth_damp_factor_coarse_saddle=0.8), prob.descs, variant=1)
    uo = u0.copy()
    orc.cell_vanka(0, uo, f)
    orc.close()
    assert rel(uo, gs) <= 1e-12
    nc, ptr, reo = pmg_oracle.color_cells(prob.arrays(0)["u_global"])
    colored = sweep_sets([reo[ptr[k]:ptr[k + 1]] for k in range(nc)], 0.8)
    assert rel(colored, gs) > 1e-3
