"""Synthetic Q2/P1disc Oseen problem hierarchies (host C++ generator, libpmg_problem.so).

These stand in for what ParMooN's set-up hands to its multigrid (mesh refinement,
FE spaces and NSType4Galerkin3D assembly; see csrc/problem/pmg_problem.h).  The named
configurations restate the benchmark workloads:

* ``oseen3d(n)``  - unit cube, n^3 hexahedra (n = 4 * 2^k), Q2/P1disc, nu = 1e-2,
  b = curl(psi) with 8 seeded Fourier modes scaled to max|b| = 1 (seed 1001),
  f = 8 seeded Fourier modes (seed 1002), u = 0 on the boundary; levels 4^3 ... n^3.
* ``quadlin(...)`` - the closed-form AnsatzQuadLin case (u quadratic, p linear).
* ``batch_instance`` / ``batched_oseen`` - the batched workload: independent unit-cube
  Oseen problems of n^3 cells, n drawn from {8, 12, 16, 20, 24} (seed 2001), nu
  log-uniform in [1e-3, 1] (seed 2002), Dirichlet on all six faces or on x=0, y=0, z=0
  only (probability 1/2 each, seed 2003), coarsest-grid interior vertices moved by
  U[-0.1h, 0.1h]^3 (seed 4000+i; the finer levels refine the perturbed trilinear cells),
  Fourier convection (seed 3000+i) and right-hand side (seed 6000+i); levels n, n/2, ...
  down to the first size <= 4 or odd (12 -> 6 -> 3, 20 -> 10 -> 5).
* ``channel`` - one Crank-Nicolson step system of the time-dependent channel workload:
  box [0,2.5] x [0,0.41]^2, hex cells 8k x k x k on the finest level (96 x 16 x 16 for
  k = 12 ... 16), parabolic inflow with mean velocity 1 (max 2.25) at x = 0, no-slip walls,
  natural outflow at x = 2.5, Re = 100 (nu = 0.0041), Oseen convection by the inflow profile,
  sigma = 1 / (theta dt) = 200 mass term (dt = 0.01, theta = 1/2), no pressure projection.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _abi


class Problem:
    def __init__(self, **cfg):
        lib = _abi.load_problem()
        c = _abi.PmgpConfig()
        lib.pmgp_config_default(C.byref(c))
        for k, v in cfg.items():
            if k == "conv_const":
                c.conv_const[:] = list(v)
            elif not hasattr(c, k):
                raise AttributeError(k)
            else:
                setattr(c, k, v)
        self.config = c
        self._lib = lib
        h = C.c_void_p()
        st = lib.pmgp_create(C.byref(c), C.byref(h))
        if st != 0:
            raise _abi.PmgError(st, "pmgp_create")
        self._h = h
        self.n_levels = lib.pmgp_n_levels(h)
        self.descs = []
        for lvl in range(self.n_levels):
            d = _abi.PmgLevelDesc()
            lib.pmgp_level_desc(h, lvl, C.byref(d))
            self.descs.append(d)

    def close(self):
        if getattr(self, "_h", None):
            self._lib.pmgp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def n_dof(self, lvl: int) -> int:
        return self._lib.pmgp_n_dof(self._h, lvl)

    def _vec(self, fn, lvl, n):
        p = fn(self._h, lvl)
        if not p:
            return None
        return np.ctypeslib.as_array(p, shape=(n,))

    def rhs(self, lvl: int = -1) -> np.ndarray:
        lvl = lvl % self.n_levels
        return self._vec(self._lib.pmgp_rhs, lvl, self.n_dof(lvl))

    def exact(self, lvl: int = -1):
        lvl = lvl % self.n_levels
        return self._vec(self._lib.pmgp_exact, lvl, self.n_dof(lvl))

    def dof_coords(self, lvl: int = -1) -> np.ndarray:
        lvl = lvl % self.n_levels
        d = self.descs[lvl]
        return self._vec(self._lib.pmgp_dof_coords, lvl, 3 * d.n_udof).reshape(-1, 3)

    def cell_vertices(self, lvl: int = -1) -> np.ndarray:
        lvl = lvl % self.n_levels
        out = np.empty(self.descs[lvl].n_cells * 24)
        self._lib.pmgp_cell_vertices(self._h, lvl, out.ctypes.data_as(_abi.c_double_p))
        return out

    def conv_quad(self, lvl: int = -1) -> np.ndarray:
        """The configuration's convection field at the 27 Gauss points of every cell."""
        lvl = lvl % self.n_levels
        out = np.empty(self.descs[lvl].n_cells * 81)
        self._lib.pmgp_conv_quad(self._h, lvl, out.ctypes.data_as(_abi.c_double_p))
        return out

    def sizes(self, lvl: int = -1) -> dict:
        d = self.descs[lvl % self.n_levels]
        return dict(n_udof=d.n_udof, n_active=d.n_active, n_dirichlet=d.n_dirichlet, n_pdof=d.n_pdof,
                    n_cells=d.n_cells, a_nnz=d.a_nnz, b_nnz=d.b_nnz, bt_nnz=d.bt_nnz,
                    n_dof=3 * d.n_udof + d.n_pdof)

    # numpy views of a level's arrays (borrowed from the generator)
    def arrays(self, lvl: int) -> dict:
        d = self.descs[lvl]
        nu, npd, na, nb, nbt, nc = d.n_udof, d.n_pdof, d.a_nnz, d.b_nnz, d.bt_nnz, d.n_cells
        v = lambda p, n: np.ctypeslib.as_array(p, shape=(n,))
        out = dict(
            a_rowptr=v(d.a_rowptr, nu + 1), a_kcol=v(d.a_kcol, na),
            a=[v(d.a[i], na) for i in range(9)],
            b_rowptr=v(d.b_rowptr, npd + 1), b_kcol=v(d.b_kcol, nb), b=[v(d.b[i], nb) for i in range(3)],
            bt_rowptr=v(d.bt_rowptr, nu + 1), bt_kcol=v(d.bt_kcol, nbt), bt=[v(d.bt[i], nbt) for i in range(3)],
            u_global=v(d.u_global, 27 * nc).reshape(nc, 27), p_global=v(d.p_global, 4 * nc).reshape(nc, 4),
        )
        if lvl > 0:
            out["parent"] = v(d.parent, nc)
            out["child"] = v(d.child, nc)
            out["qq_u"] = v(d.qq_u, 8 * 27 * 27).reshape(8, 27, 27)
            out["qq_p"] = v(d.qq_p, 8 * 16).reshape(8, 4, 4)
        return out


def oseen3d(n: int = 32, coarsest: int = 4, nu: float = 1e-2, seed_conv: int = 1001, seed_rhs: int = 1002,
            dirichlet_faces: int = 63, perturb: float = 0.0, seed_perturb: int = 4000) -> Problem:
    """Oseen on the unit cube, levels coarsest^3 ... n^3 (n = coarsest * 2^k)."""
    levels = 1
    m = coarsest
    while m < n:
        m *= 2
        levels += 1
    if m != n:
        raise ValueError("n must be coarsest * 2^k")
    return Problem(cx=coarsest, cy=coarsest, cz=coarsest, n_levels=levels, nu=nu,
                   conv_kind=_abi.CONV_FOURIER, seed_conv=seed_conv, conv_max=1.0,
                   rhs_kind=_abi.RHS_FOURIER, seed_rhs=seed_rhs, dirichlet_faces=dirichlet_faces,
                   perturb=perturb, seed_perturb=seed_perturb)


def quadlin(coarse=(1, 1, 1), n_levels: int = 3, nu: float = 1.0, conv=(0.0, 0.0, 0.0),
            dirichlet_faces: int = 63) -> Problem:
    """Examples/NSE_3D/AnsatzQuadLin.h: u in Q2, p in P1 reproduced exactly."""
    kind = _abi.CONV_CONST if any(conv) else _abi.CONV_NONE
    return Problem(cx=coarse[0], cy=coarse[1], cz=coarse[2], n_levels=n_levels, nu=nu, conv_kind=kind,
                   conv_const=conv, rhs_kind=_abi.RHS_EXACT, exact_kind=_abi.EXACT_QUADLIN,
                   dirichlet_faces=dirichlet_faces)


DIRICHLET_ALL = 63
DIRICHLET_X0_Y0_Z0 = 1 | 4 | 16


def coarsest_of(n: int) -> tuple[int, int]:
    """(coarsest size, number of levels) of the halving hierarchy of an n^3 cube."""
    m, levels = n, 1
    while m > 4 and m % 2 == 0:
        m //= 2
        levels += 1
    return m, levels


def batch_instance(n: int, nu: float, dirichlet_faces: int, i: int, perturb: float = 0.1) -> Problem:
    """Problem i of the batched workload (see the module docstring)."""
    m, levels = coarsest_of(n)
    return Problem(cx=m, cy=m, cz=m, n_levels=levels, nu=nu, conv_kind=_abi.CONV_FOURIER, seed_conv=3000 + i,
                   conv_max=1.0, rhs_kind=_abi.RHS_FOURIER, seed_rhs=6000 + i, dirichlet_faces=dirichlet_faces,
                   perturb=perturb, seed_perturb=4000 + i)


def batched_specs(count: int = 64, sizes=(8, 12, 16, 20, 24)):
    """(n, nu, dirichlet_faces) of the batched workload's problems."""
    ns = np.random.default_rng(2001).choice(np.asarray(sizes), size=count)
    nus = 10.0 ** np.random.default_rng(2002).uniform(-3.0, 0.0, size=count)
    six = np.random.default_rng(2003).random(count) < 0.5
    return [(int(ns[i]), float(nus[i]), DIRICHLET_ALL if six[i] else DIRICHLET_X0_Y0_Z0) for i in range(count)]


def batched_oseen(count: int = 64, sizes=(8, 12, 16, 20, 24)) -> list[Problem]:
    return [batch_instance(n, nu, f, i) for i, (n, nu, f) in enumerate(batched_specs(count, sizes))]


CHANNEL_DIRICHLET = 1 | 4 | 8 | 16 | 32     # all faces but the outflow x = 2.5


def channel(n_levels: int = 4, coarse=(12, 2, 2), nu: float = 0.0041, u_max: float = 2.25,
            dt: float = 0.01, theta: float = 0.5) -> Problem:
    """The Crank-Nicolson step system of the channel workload (see the module docstring)."""
    return Problem(cx=coarse[0], cy=coarse[1], cz=coarse[2], lx=2.5, ly=0.41, lz=0.41, n_levels=n_levels, nu=nu,
                   conv_kind=_abi.CONV_CHANNEL, conv_max=u_max, rhs_kind=_abi.RHS_ZERO,
                   dirichlet_faces=CHANNEL_DIRICHLET, mass=1.0 / (theta * dt))
