"""Crank-Nicolson time stepping with a fixed-point Oseen linearisation, on the device.

The time loop of the reference's TNSE3D program (3DPrograms/TNSE3D_ParMooN.C:649-882) for
one sub-step per time step with the theta scheme, theta = 1/2:

    per time step n -> n+1 (tau = dt, sigma = 1 / (theta tau)):
      right-hand side   B = sigma M u^n - A(u^n) u^n       (TSystemTNSE3D::AssembleSystMat,
                                                            src/System/SystemTNSE3D.C:1118-1190,
                                                            scaled by 1 / (theta tau), no sources)
      u^{n+1,0} = u^n
      for k = 1 .. n_picard:                               (the j-loop of TNSE3D_ParMooN.C)
        restrict u^{n+1,k-1} to every level                (RestrictToAllGrids,
                                                            SystemTNSE3D.C:1400-1415)
        assemble sigma M + A(u_level) on every level       (AssembleNonLinear / NSType4Galerkin3D)
        solve with FGMRES + GMG cell Vanka from u^{n+1,k-1} (TSystemTNSE3D::Solve)

with the pressure fully implicit (the pressure unknown is p / theta).  Every vector stays
in device memory from one step to the next: the velocity at the Gauss points
(pmg_eval_velocity), the re-assembly (pmg_assemble_oseen), the right-hand side
(pmg_theta_rhs), the restriction (pmg_restrict_to_all_grids), the value-dependent set-up
(pmg_refactor) and the solve (pmg_fgmres_solve) all take device pointers.
"""
from __future__ import annotations

import ctypes as C
import time

import numpy as np

from .multigrid import DeviceVector, build_hierarchy

MEM_DEVICE = 1


class ThetaTimeLoop:
    """Device-resident time loop for a problem hierarchy (problem.Problem: level
    descriptors, cell geometry and Dirichlet values in its right-hand side)."""

    def __init__(self, prob, params, nu: float, dt: float, theta: float = 0.5, dirichlet: np.ndarray | None = None,
                 u0: np.ndarray | None = None):
        if abs(theta - 0.5) > 1e-15:
            raise ValueError("the right-hand side assumes Crank-Nicolson (theta = 1/2)")
        self.prob, self.nu, self.dt, self.theta = prob, nu, dt, theta
        self.sigma = 1.0 / (theta * dt)
        self.mg = mg = build_hierarchy(prob, params)
        self.L = L = prob.n_levels
        for l in range(L):
            mg.level_geometry(l, prob.cell_vertices(l))
        n = mg.n_dof[-1]
        d = prob.descs[-1]
        g = np.array(prob.rhs(-1) if dirichlet is None else dirichlet, dtype=np.float64)
        u = np.zeros(n) if u0 is None else np.array(u0, dtype=np.float64)
        for c in range(3):   # the initial state carries the Dirichlet values
            u[c * d.n_udof + d.n_active:(c + 1) * d.n_udof] = g[c * d.n_udof + d.n_active:(c + 1) * d.n_udof]
        self.g = DeviceVector(mg, n, g)
        self.u = DeviceVector(mg, n, u)
        self.rhs = DeviceVector(mg, n, np.zeros(n))
        self.ul = [DeviceVector(mg, mg.n_dof[l], np.zeros(mg.n_dof[l])) for l in range(L - 1)]
        self.bq = [DeviceVector(mg, prob.descs[l].n_cells * 81) for l in range(L)]
        self._ul_ptrs = (C.c_void_p * max(1, L - 1))(*[v.ptr for v in self.ul])
        self.history = []

    def _check(self, st, what):
        self.mg._check(st, what)

    def _assemble_all(self):
        lib, ctx = self.mg._lib, self.mg._ctx
        self._check(lib.pmg_restrict_to_all_grids(ctx, self.u.ptr, self._ul_ptrs, MEM_DEVICE), "restrict_to_all_grids")
        for l in range(self.L):
            src = self.u if l == self.L - 1 else self.ul[l]
            self._check(lib.pmg_eval_velocity(ctx, l, src.ptr, self.bq[l].ptr, MEM_DEVICE), "eval_velocity")
            self._check(lib.pmg_assemble_oseen(ctx, l, self.bq[l].ptr, self.nu, self.sigma, MEM_DEVICE),
                        "assemble_oseen")
        self.mg.refactor()

    def step(self, n_picard: int = 3):
        """One time step; returns (wall seconds, FGMRES iterations per fixed-point iteration)."""
        lib, ctx, top = self.mg._lib, self.mg._ctx, self.L - 1
        self.mg.synchronize()
        t0 = time.perf_counter()
        self._check(lib.pmg_eval_velocity(ctx, top, self.u.ptr, self.bq[top].ptr, MEM_DEVICE), "eval_velocity")
        self._check(lib.pmg_theta_rhs(ctx, top, self.u.ptr, self.bq[top].ptr, self.nu, self.sigma, self.g.ptr,
                                      self.rhs.ptr, MEM_DEVICE), "theta_rhs")
        its = []
        for _ in range(n_picard):
            self._assemble_all()
            it, _ = self.mg.fgmres_device(self.u, self.rhs)
            its.append(it)
        self.mg.synchronize()
        dt = time.perf_counter() - t0
        self.history.append((dt, its))
        return dt, its

    def solution(self) -> np.ndarray:
        return self.u.download()

    def close(self):
        self.mg.close()


def perturbed_inflow(prob, amplitude: float = 0.01, seed: int = 5001) -> np.ndarray:
    """The problem's Dirichlet values with the normal (x) velocity on the Dirichlet dofs
    scaled by 1 + amplitude * U[-1, 1] (seeded): the parabolic inflow plus a 1 % random
    perturbation; walls keep zero."""
    g = np.array(prob.rhs(-1), dtype=np.float64)
    d = prob.descs[-1]
    dn = np.arange(d.n_active, d.n_udof)
    xi = np.random.default_rng(seed).uniform(-1.0, 1.0, dn.size)
    g[dn] *= 1.0 + amplitude * xi
    return g
