"""Dense-algebra statements of the type-4 objects, built directly from a level's CSR
arrays, used to pin the oracle on small meshes.

Conventions (MGComponents3D.C:2872-3035, CoupledMatVect): the vector is [u1 | u2 | u3 | p];
A block b = 3*r + c couples velocity component r (row) with component c (column); active
velocity rows (i < N_Active) carry all nine A blocks and the three BT blocks; rows
N_Active..N_UDOF carry only the diagonal blocks A11, A22, A33; pressure row i carries
B1, B2, B3.
"""
from __future__ import annotations

import numpy as np


def dense_operator(prob, lvl: int, dtype=np.float64) -> np.ndarray:
    d = prob.descs[lvl]
    a = prob.arrays(lvl)
    N, NA, NP = d.n_udof, d.n_active, d.n_pdof
    n = 3 * N + NP
    K = np.zeros((n, n), dtype=dtype)
    rows_a = np.repeat(np.arange(N), np.diff(a["a_rowptr"]))
    cols_a = a["a_kcol"]
    act = rows_a < NA
    for r in range(3):
        for c in range(3):
            v = a["a"][3 * r + c]
            if r == c:
                K[r * N + rows_a, c * N + cols_a] = v
            else:
                K[r * N + rows_a[act], c * N + cols_a[act]] = v[act]
    rows_bt = np.repeat(np.arange(N), np.diff(a["bt_rowptr"]))
    act_bt = rows_bt < NA
    for r in range(3):
        K[r * N + rows_bt[act_bt], 3 * N + a["bt_kcol"][act_bt]] = a["bt"][r][act_bt]
    rows_b = np.repeat(np.arange(NP), np.diff(a["b_rowptr"]))
    for c in range(3):
        K[3 * N + rows_b, c * N + a["b_kcol"]] = a["b"][c]
    return K


def into_l20(p: np.ndarray) -> np.ndarray:
    """IntoL20Vector3D case -11 (MGComponents3D.C:1756-1780): subtract the mean of the
    P1disc constants (every 4th entry) from those entries."""
    q = p.copy()
    q[0::4] -= q[0::4].mean()
    return q


def patch_indices(prob, lvl: int, cell: int) -> np.ndarray:
    """Global indices of the 85 local unknowns of a cell, in CellVanka's local order:
    u1 of the 27 dofs, u2, u3, then the 4 pressure dofs (NSE_MGLevel4.C:902-1061)."""
    d = prob.descs[lvl]
    a = prob.arrays(lvl)
    ug = a["u_global"][cell]
    pg = a["p_global"][cell]
    N = d.n_udof
    return np.concatenate([ug, ug + N, ug + 2 * N, 3 * N + pg]).astype(np.int64)


def randomize_values(prob, lvl: int, seed: int):
    """Overwrite every stored operator value of a level with N(0,1) samples (same pattern)."""
    rng = np.random.default_rng(seed)
    a = prob.arrays(lvl)
    for arr in a["a"] + a["b"] + a["bt"]:
        arr[:] = rng.standard_normal(arr.size)
