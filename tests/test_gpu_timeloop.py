"""The channel workload's time loop on the device (to_merge_master_parmoon_amd.timeloop):
Crank-Nicolson with fixed-point Oseen iterations, state kept in device memory.

* pmg_theta_rhs (matrix-free sigma M u - A(u) u, TSystemTNSE3D::AssembleSystMat) equals
  the same right-hand side formed through the assembled operator: pmg_assemble_oseen with
  the mass coefficient -sigma, then -K [u; 0] on the active velocity rows (1e-13);
* a few time steps of the loop: the fixed-point iteration contracts (the nonlinear
  residual of the step falls from one iteration to the next), and the mixed-precision
  loop (BF16 patch data, FP32 smoother) stays within 1e-8 of the FP64 loop at every step.
"""
import numpy as np
import pytest

from conftest import params, rel

pytestmark = pytest.mark.gpu

NU, DT = 0.0041, 0.01


def _small():
    from to_merge_master_parmoon_amd import problem
    return problem.channel(n_levels=3, coarse=(6, 1, 1))


def test_theta_rhs_matches_assembled_operator():
    from to_merge_master_parmoon_amd.multigrid import DeviceVector
    from to_merge_master_parmoon_amd.timeloop import ThetaTimeLoop, perturbed_inflow
    p = _small()
    tl = ThetaTimeLoop(p, params(internal_project_pressure=0), NU, DT, dirichlet=perturbed_inflow(p))
    mg, lib, ctx, top = tl.mg, tl.mg._lib, tl.mg._ctx, p.n_levels - 1
    d = p.descs[-1]
    n = mg.n_dof[-1]
    u = np.random.default_rng(900).standard_normal(n)
    uv = DeviceVector(mg, n, u)
    bq = tl.bq[top]
    mg._check(lib.pmg_eval_velocity(ctx, top, uv.ptr, bq.ptr, 1))
    mg._check(lib.pmg_theta_rhs(ctx, top, uv.ptr, bq.ptr, NU, tl.sigma, tl.g.ptr, tl.rhs.ptr, 1))
    r1 = tl.rhs.download()
    mg._check(lib.pmg_assemble_oseen(ctx, top, bq.ptr, NU, -tl.sigma, 1))
    u0p = u.copy()
    u0p[3 * d.n_udof:] = 0.0
    y = mg.matvec(top, u0p)
    g = tl.g.download()
    N, NA = d.n_udof, d.n_active
    want = np.zeros(n)
    for c in range(3):
        want[c * N:c * N + NA] = -y[c * N:c * N + NA]
        want[c * N + NA:(c + 1) * N] = g[c * N + NA:(c + 1) * N]
    e = rel(r1, want)
    print(f"theta right-hand side vs assembled operator: {e:.2e}", flush=True)
    assert e <= 1e-13
    tl.close()


def _nonlinear_residual(tl):
    """||rhs - K(u) u|| / ||rhs|| of the current iterate with the operator re-assembled at it"""
    tl._assemble_all()
    top = tl.L - 1
    u, rhs = tl.u.download(), tl.rhs.download()
    return np.linalg.norm(tl.mg.defect_nse4(top, u, rhs)) / np.linalg.norm(rhs)


def test_time_loop_fixed_point_and_mixed_precision():
    from to_merge_master_parmoon_amd.timeloop import ThetaTimeLoop, perturbed_inflow
    p = _small()
    g = perturbed_inflow(p)
    loops = {prec: ThetaTimeLoop(p, params(internal_project_pressure=0, precision=prec), NU, DT, dirichlet=g)
             for prec in (0, 2)}
    for step in range(3):
        sols = {}
        for prec, tl in loops.items():
            lib, ctx, top = tl.mg._lib, tl.mg._ctx, tl.L - 1
            tl.mg._check(lib.pmg_eval_velocity(ctx, top, tl.u.ptr, tl.bq[top].ptr, 1))
            tl.mg._check(lib.pmg_theta_rhs(ctx, top, tl.u.ptr, tl.bq[top].ptr, NU, tl.sigma, tl.g.ptr, tl.rhs.ptr, 1))
            res = []
            for k in range(3):
                tl._assemble_all()
                it, _ = tl.mg.fgmres_device(tl.u, tl.rhs)
                res.append(_nonlinear_residual(tl))
            print(f"step {step} precision {prec}: nonlinear residuals {['%.2e' % r for r in res]}", flush=True)
            assert res[1] < res[0] and res[2] < res[1]
            sols[prec] = tl.solution()
        e = rel(sols[2], sols[0])
        print(f"step {step}: mixed vs FP64 {e:.2e}", flush=True)
        assert e <= 1e-8
    for tl in loops.values():
        tl.close()
