"""Level dumps (pmg_level_read / pmg_level_write; SaveData / ReadData of
src/General/MainUtilities.C:3360-3440 for vectors), host code of libpmg.

* a dump written by hand from the documented layout (numpy raw files and a manifest,
  independent of the library's writer) is read back array for array;
* the library's writer and reader round-trip a whole Oseen hierarchy, and the oracle
  solves the read-back hierarchy bit for bit like the original;
* vectors are raw doubles exactly as SaveData writes them;
* truncated, missing or inconsistent dumps are refused with PMG_ERR_INVALID_ARGUMENT.
"""
import os

import numpy as np
import pytest

import pmg_oracle
from conftest import params


def _arrays(desc, prob, lvl):
    a = prob.arrays(lvl)
    return a


def _hand_written(path, prob, lvl):
    """the layout of include/pmg.h written directly with numpy"""
    os.makedirs(path, exist_ok=True)
    d = prob.descs[lvl]
    a = prob.arrays(lvl)
    w = lambda name, x, dt: np.ascontiguousarray(x, dtype=dt).tofile(os.path.join(path, name))
    for k in ("a_rowptr", "a_kcol", "b_rowptr", "b_kcol", "bt_rowptr", "bt_kcol"):
        w(k + ".i32", a[k], np.int32)
    for k, nm in enumerate(["A11", "A12", "A13", "A21", "A22", "A23", "A31", "A32", "A33"]):
        w(nm + ".f64", a["a"][k], np.float64)
    for k in range(3):
        w(f"B{k + 1}.f64", a["b"][k], np.float64)
        w(f"B{k + 1}T.f64", a["bt"][k], np.float64)
    nc = d.n_cells
    w("u_begin.i32", np.arange(nc + 1) * 27, np.int32)
    w("u_global.i32", a["u_global"].ravel(), np.int32)
    w("p_begin.i32", np.arange(nc + 1) * 4, np.int32)
    w("p_global.i32", a["p_global"].ravel(), np.int32)
    nch = 0
    if lvl > 0:
        nch = 8
        w("parent.i32", a["parent"], np.int32)
        w("child.i32", a["child"], np.int32)
        w("qq_u.f64", a["qq_u"].ravel(), np.float64)
        w("qq_p.f64", a["qq_p"].ravel(), np.float64)
    with open(os.path.join(path, "manifest.txt"), "w") as f:
        for k in ("n_udof", "n_active", "hanging_bound", "n_dirichlet", "n_pdof", "a_nnz", "b_nnz", "bt_nnz",
                  "n_cells", "n_ulocal", "n_plocal"):
            f.write(f"{k} {getattr(d, k)}\n")
        f.write(f"n_children {nch}\nn_colors 0\n")


def _same_level(d1, d2):
    for k in ("n_udof", "n_active", "hanging_bound", "n_dirichlet", "n_pdof", "a_nnz", "b_nnz", "bt_nnz", "n_cells",
              "n_ulocal", "n_plocal"):
        assert getattr(d1, k) == getattr(d2, k), k
    v = lambda p, n: np.ctypeslib.as_array(p, shape=(n,))
    assert np.array_equal(v(d1.a_rowptr, d1.n_udof + 1), v(d2.a_rowptr, d2.n_udof + 1))
    assert np.array_equal(v(d1.a_kcol, d1.a_nnz), v(d2.a_kcol, d2.a_nnz))
    for k in range(9):
        assert np.array_equal(v(d1.a[k], d1.a_nnz), v(d2.a[k], d2.a_nnz))
    for k in range(3):
        assert np.array_equal(v(d1.b[k], d1.b_nnz), v(d2.b[k], d2.b_nnz))
        assert np.array_equal(v(d1.bt[k], d1.bt_nnz), v(d2.bt[k], d2.bt_nnz))
    for cell in (0, d1.n_cells - 1):
        u1 = [d1.u_global[d1.u_begin[cell] + j] for j in range(27)]
        u2 = [d2.u_global[d2.u_begin[cell] + j] for j in range(27)]
        assert u1 == u2


def test_hand_written_dump_is_read(tmp_path):
    from to_merge_master_parmoon_amd import problem
    from to_merge_master_parmoon_amd.dump import LevelDump
    prob = problem.oseen3d(4, coarsest=2)
    for lvl in range(prob.n_levels):
        path = str(tmp_path / f"level{lvl}")
        _hand_written(path, prob, lvl)
        lv = LevelDump(path)
        _same_level(lv.desc, prob.descs[lvl])
        if lvl > 0:
            q = np.ctypeslib.as_array(lv.desc.qq_u, shape=(8 * 729,))
            assert np.array_equal(q, prob.arrays(lvl)["qq_u"].ravel())
        lv.close()


def test_round_trip_and_oracle_solve(tmp_path):
    from to_merge_master_parmoon_amd import problem
    from to_merge_master_parmoon_amd.dump import DumpedHierarchy, read_vector, write_hierarchy, write_vector
    prob = problem.oseen3d(8)
    write_hierarchy(str(tmp_path), prob.descs)
    f = prob.rhs(-1).copy()
    write_vector(str(tmp_path / "rhs"), f)
    assert np.array_equal(np.fromfile(str(tmp_path / "rhs"), dtype=np.float64), f)   # SaveData's bytes
    H = DumpedHierarchy(str(tmp_path))
    assert H.n_levels == prob.n_levels
    for d1, d2 in zip(H.descs, prob.descs):
        _same_level(d1, d2)
    f2 = read_vector(str(tmp_path / "rhs"), f.size)
    prm = params()
    x1, x2 = np.zeros_like(f), np.zeros_like(f)
    o1 = pmg_oracle.Oracle(prm, prob.descs)
    o2 = pmg_oracle.Oracle(prm, H.descs)
    r1, r2 = o1.fgmres(x1, f), o2.fgmres(x2, f2)
    assert r1[0] == r2[0] and np.array_equal(x1, x2)
    o1.close()
    o2.close()
    H.close()


def test_malformed_dumps_are_refused(tmp_path):
    from to_merge_master_parmoon_amd import _abi, problem
    from to_merge_master_parmoon_amd.dump import LevelDump, read_vector
    prob = problem.oseen3d(4, coarsest=2)
    path = str(tmp_path / "level1")
    _hand_written(path, prob, 1)
    kc = os.path.join(path, "a_kcol.i32")
    good = open(kc, "rb").read()
    open(kc, "wb").write(good[:-4])               # truncated
    with pytest.raises(_abi.PmgError) as e:
        LevelDump(path)
    assert e.value.status == 1
    bad = np.frombuffer(good, dtype=np.int32).copy()
    bad[3] = prob.descs[1].n_udof                 # column out of range
    bad.tofile(kc)
    with pytest.raises(_abi.PmgError):
        LevelDump(path)
    open(kc, "wb").write(good)
    os.remove(os.path.join(path, "qq_p.f64"))     # missing transfer table
    with pytest.raises(_abi.PmgError):
        LevelDump(path)
    with pytest.raises(_abi.PmgError):
        read_vector(str(tmp_path / "absent"), 4)
