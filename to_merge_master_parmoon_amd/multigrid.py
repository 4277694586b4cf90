"""Python face of libpmg with the reference's operator names.

``NSEMultiGrid`` mirrors ParMooN's TNSE_MultiGrid + TNSE_MGLevel4 + TMultiGridIte +
TFgmresIte as used by TSystemNSE3D (src/System/SystemNSE3D.C:400-450, 806-878,
1199-1235): levels are added coarsest first, ``setup()`` does the value-dependent
factorizations, and the level methods keep the reference's names and argument meaning
(``defect`` = TNSE_MGLevel4::Defect, ``cell_vanka`` = CellVanka, ``restrict_defect`` =
DefectRestriction, ``prolongate`` = Prolongate, ``solve_exact`` = SolveExact, ``cycle`` =
TMultiGridIte::Iterate, ``fgmres`` = TFgmresIte::Iterate).  Errors the reference turns into
exit(4711) raise PmgError with the status code instead.  Every call runs on the GPU through
the C ABI; there is no host fallback.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _abi
from ._abi import MEM_DEVICE, MEM_HOST, PmgError


def _ptr(a: np.ndarray):
    if a.dtype != np.float64 or not a.flags["C_CONTIGUOUS"]:
        raise TypeError("float64 C-contiguous array required")
    return a.ctypes.data_as(C.c_void_p)


def _autotune(obj, n_levels: int, reps: int) -> list:
    obj._check(obj._lib.pmg_autotune(obj._ctx, reps), "pmg_autotune")
    out = []
    for l in range(n_levels):
        v = (C.c_int * 6)()
        obj._check(obj._lib.pmg_level_tuning(obj._ctx, l, v), "pmg_level_tuning")
        out.append(dict(res_items_per_wg=[v[0], v[1]], res_xcd_groups=[v[2], v[3]], op_rows_per_wg=[v[4], v[5]]))
    return out


def save_tuning(shapes: list, path: str):
    """Write the launch shapes of autotune() as JSON."""
    import json
    with open(path, "w") as fh:
        json.dump({"levels": shapes}, fh, indent=1)


def load_tuning(obj, path: str) -> list:
    """Apply launch shapes saved by save_tuning to the levels of obj (pmg_set_level_tuning)."""
    import json
    with open(path) as fh:
        shapes = json.load(fh)["levels"]
    for l, s in enumerate(shapes):
        v = (C.c_int * 6)(*s["res_items_per_wg"], *s["res_xcd_groups"], *s["op_rows_per_wg"])
        obj._check(obj._lib.pmg_set_level_tuning(obj._ctx, l, v), "pmg_set_level_tuning")
    return shapes


class DeviceVector:
    """A float64 vector in device memory owned by a context."""

    def __init__(self, mg: "NSEMultiGrid", n: int, init: np.ndarray | None = None):
        self.mg, self.n = mg, n
        p = C.c_void_p()
        mg._check(mg._lib.pmg_device_alloc(mg._ctx, 8 * n, C.byref(p)))
        self.ptr = p
        if init is not None:
            self.upload(init)

    def upload(self, a: np.ndarray):
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert a.size == self.n
        self.mg._check(self.mg._lib.pmg_memcpy(self.mg._ctx, self.ptr, _ptr(a), 8 * self.n, 0))

    def download(self) -> np.ndarray:
        out = np.empty(self.n)
        self.mg._check(self.mg._lib.pmg_memcpy(self.mg._ctx, _ptr(out), self.ptr, 8 * self.n, 1))
        return out

    def free(self):
        if self.ptr:
            self.mg._lib.pmg_device_free(self.mg._ctx, self.ptr)
            self.ptr = None


class NSEMultiGrid:
    def __init__(self, params: _abi.PmgParams | None = None, device: int = 0, **kw):
        self._lib = _abi.load_pmg()
        self.params = params if params is not None else _abi.default_params(**kw)
        ctx = C.c_void_p()
        st = self._lib.pmg_ctx_create(device, C.byref(self.params), C.byref(ctx))
        if st != 0:
            raise PmgError(st, "pmg_ctx_create")
        self._ctx = ctx
        self.n_dof: list[int] = []

    # ------------------------------------------------------------------ plumbing
    def _check(self, st: int, what: str = ""):
        if st != 0:
            msg = self._lib.pmg_last_error_string(self._ctx)
            raise PmgError(st, f"{what} {msg.decode() if msg else ''}".strip())
        return st

    def close(self):
        if getattr(self, "_ctx", None):
            self._lib.pmg_ctx_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, **kw):
        for k, v in kw.items():
            setattr(self.params, k, v)
        self._check(self._lib.pmg_set_params(self._ctx, C.byref(self.params)), "pmg_set_params")

    # ------------------------------------------------------------------ levels
    def add_level(self, desc: _abi.PmgLevelDesc) -> int:
        lvl = C.c_int()
        self._check(self._lib.pmg_level_add(self._ctx, C.byref(desc), C.byref(lvl)), "pmg_level_add")
        n = C.c_int()
        self._lib.pmg_level_info(self._ctx, lvl.value, C.byref(n), None, None)
        self.n_dof.append(n.value)
        return lvl.value

    def update_values(self, level: int, desc: _abi.PmgLevelDesc):
        self._check(self._lib.pmg_level_update_values(self._ctx, level, C.byref(desc)), "update_values")

    def setup(self):
        self._check(self._lib.pmg_setup(self._ctx), "pmg_setup")

    def refactor(self):
        """Recompute all value-dependent factors from the resident operator."""
        self._check(self._lib.pmg_refactor(self._ctx), "pmg_refactor")

    def level_info(self, level: int) -> dict:
        n, nc, mc = C.c_int(), C.c_int(), C.c_int()
        self._check(self._lib.pmg_level_info(self._ctx, level, C.byref(n), C.byref(nc), C.byref(mc)))
        return dict(n_dof=n.value, n_colors=nc.value, max_cells_per_color=mc.value)

    def shared_block(self, level: int) -> bool:
        """True when the level's velocity rows stream one shared A block (A11 = A22 = A33,
        no coupling blocks on the active rows), False when they stream all nine."""
        sb = C.c_int()
        self._check(self._lib.pmg_level_form(self._ctx, level, C.byref(sb)), "pmg_level_form")
        return bool(sb.value)

    @property
    def n_levels(self) -> int:
        return len(self.n_dof)

    # ------------------------------------------------------------------ level operations (host arrays)
    def defect(self, level: int, u: np.ndarray, f: np.ndarray):
        d = np.empty_like(u)
        res = C.c_double()
        self._check(self._lib.pmg_defect(self._ctx, level, _ptr(u), _ptr(f), _ptr(d), C.byref(res), MEM_HOST),
                    "pmg_defect")
        return d, res.value

    def matvec(self, level: int, x: np.ndarray) -> np.ndarray:
        y = np.empty_like(x)
        self._check(self._lib.pmg_matvec(self._ctx, level, _ptr(x), _ptr(y), MEM_HOST), "pmg_matvec")
        return y

    def defect_nse4(self, level: int, x: np.ndarray, b: np.ndarray) -> np.ndarray:
        r = np.empty_like(x)
        self._check(self._lib.pmg_defect_nse4(self._ctx, level, _ptr(x), _ptr(b), _ptr(r), MEM_HOST), "defect_nse4")
        return r

    def cell_vanka(self, level: int, u: np.ndarray, f: np.ndarray, nsweeps: int = 1) -> np.ndarray:
        self._check(self._lib.pmg_vanka_smooth(self._ctx, level, _ptr(u), _ptr(f), nsweeps, MEM_HOST), "vanka")
        return u

    def restrict_defect(self, level: int, d_fine: np.ndarray) -> np.ndarray:
        out = np.empty(self.n_dof[level - 1])
        self._check(self._lib.pmg_restrict_defect(self._ctx, level, _ptr(d_fine), _ptr(out), MEM_HOST), "restrict")
        return out

    def prolongate(self, level: int, u_coarse: np.ndarray) -> np.ndarray:
        out = np.empty(self.n_dof[level])
        self._check(self._lib.pmg_prolongate(self._ctx, level, _ptr(u_coarse), _ptr(out), MEM_HOST), "prolongate")
        return out

    def correct_nodes(self, level: int, u: np.ndarray) -> np.ndarray:
        self._check(self._lib.pmg_correct_nodes(self._ctx, level, _ptr(u), MEM_HOST))
        return u

    def set_dirichlet_nodes(self, level: int, u: np.ndarray, f: np.ndarray) -> np.ndarray:
        self._check(self._lib.pmg_set_dirichlet(self._ctx, level, _ptr(u), _ptr(f), MEM_HOST))
        return u

    def correct_defect(self, level: int, d: np.ndarray) -> np.ndarray:
        self._check(self._lib.pmg_correct_defect(self._ctx, level, _ptr(d), MEM_HOST))
        return d

    def update(self, level: int, u: np.ndarray, d: np.ndarray) -> np.ndarray:
        self._check(self._lib.pmg_update(self._ctx, level, _ptr(u), _ptr(d), MEM_HOST))
        return u

    def solve_exact(self, f: np.ndarray) -> np.ndarray:
        u = np.empty_like(f)
        self._check(self._lib.pmg_coarse_solve(self._ctx, _ptr(u), _ptr(f), MEM_HOST), "coarse solve")
        return u

    # ------------------------------------------------------------------ re-assembly
    def level_geometry(self, level: int, cell_vertices: np.ndarray):
        cv = np.ascontiguousarray(cell_vertices, dtype=np.float64)
        self._check(self._lib.pmg_level_geometry(self._ctx, level, _ptr(cv)), "level_geometry")

    def eval_velocity(self, level: int, u: np.ndarray, n_cells: int) -> np.ndarray:
        out = np.empty(n_cells * 81)
        self._check(self._lib.pmg_eval_velocity(self._ctx, level, _ptr(u), _ptr(out), MEM_HOST), "eval_velocity")
        return out

    def assemble_oseen(self, level: int, b_quad: np.ndarray, nu: float, sigma: float = 0.0):
        """NSType4Galerkin3D on the device; call refactor() (or setup()) before solving."""
        bq = np.ascontiguousarray(b_quad, dtype=np.float64)
        self._check(self._lib.pmg_assemble_oseen(self._ctx, level, _ptr(bq), nu, sigma, MEM_HOST), "assemble_oseen")

    # ------------------------------------------------------------------ drivers
    def cycle(self, v: np.ndarray) -> np.ndarray:
        z = np.empty_like(v)
        self._check(self._lib.pmg_vcycle(self._ctx, _ptr(v), _ptr(z), MEM_HOST), "vcycle")
        return z

    def fgmres(self, sol: np.ndarray, rhs: np.ndarray, raise_on_error: bool = True):
        it, res = C.c_int(), C.c_double()
        st = self._lib.pmg_fgmres_solve(self._ctx, _ptr(sol), _ptr(rhs), C.byref(it), C.byref(res), MEM_HOST)
        if raise_on_error:
            self._check(st, "fgmres")
        return it.value, res.value, st

    def fgmres_device(self, sol: DeviceVector, rhs: DeviceVector):
        it, res = C.c_int(), C.c_double()
        st = self._lib.pmg_fgmres_solve(self._ctx, sol.ptr, rhs.ptr, C.byref(it), C.byref(res), MEM_DEVICE)
        self._check(st, "fgmres")
        return it.value, res.value

    def fgmres_history(self, cap: int = 100000) -> np.ndarray:
        out = np.empty(cap)
        n = self._lib.pmg_fgmres_history(self._ctx, out.ctypes.data_as(_abi.c_double_p), cap)
        return out[:n].copy()

    def synchronize(self):
        self._check(self._lib.pmg_synchronize(self._ctx))

    # ------------------------------------------------------------------ timing
    def autotune(self, reps: int = 20) -> list:
        """Time the candidate launch shapes of every smoothed level on the device and keep
        the fastest (pmg_autotune); returns the chosen shapes per level."""
        return _autotune(self, self.n_levels, reps)

    def set_profiling(self, on: bool):
        self._lib.pmg_set_profiling(self._ctx, 1 if on else 0)

    def timers(self) -> dict:
        t = _abi.PmgTimers()
        self._lib.pmg_timers_get(self._ctx, C.byref(t))
        out = {}
        for k in _abi.KSTATS:
            ks = getattr(t, k)
            out[k] = dict(ms=ks.ms, calls=ks.calls, bytes=ks.bytes, flops=ks.flops)
        out["sweep_ms"], out["sweep_calls"] = t.sweep_ms, t.sweep_calls
        ks = t.vanka_res_finest
        out["vanka_res_finest"] = dict(ms=ks.ms, calls=ks.calls, bytes=ks.bytes, flops=ks.flops)
        return out

    def reset_timers(self):
        self._lib.pmg_timers_reset(self._ctx)


def color_cells(n_cells: int, u_global: np.ndarray, n_udof: int):
    """TParFEMapper3D::colorCell through the C ABI (host code, no GPU needed)."""
    lib = _abi.load_pmg()
    ug = np.ascontiguousarray(u_global, dtype=np.int32).reshape(-1)
    nl = ug.size // n_cells
    ub = (np.arange(n_cells, dtype=np.int32) * nl).astype(np.int32)
    ptr = np.empty(n_cells + 1, dtype=np.int32)
    reo = np.empty(n_cells, dtype=np.int32)
    nc = C.c_int()
    ip = lambda a: a.ctypes.data_as(_abi.c_int_p)
    st = lib.pmg_color_cells(n_cells, nl, ip(ub), ip(ug), n_udof, C.byref(nc), ip(ptr), ip(reo))
    if st:
        raise PmgError(st, "pmg_color_cells")
    return nc.value, ptr[: nc.value + 1].copy(), reo


def build_hierarchy(problem, params=None, device: int = 0, **kw) -> NSEMultiGrid:
    """AddLevel for every level of a generated problem (coarsest first), then setup()."""
    mg = NSEMultiGrid(params, device=device, **kw)
    for d in problem.descs:
        mg.add_level(d)
    mg.setup()
    return mg


class NSEBatch:
    """Many independent problems solved together (pmg_batch_create / pmg_batch_fgmres_solve).

    ``problems`` are generated hierarchies (``problem.Problem``); ``project`` gives each its
    INTERNAL_PROJECT_PRESSURE (default: 1 when all six faces are Dirichlet).  The levels are
    merged block-diagonally, aligned at the finest level, so every sweep, defect and transfer
    is one launch for all problems; ``fgmres`` returns per problem what ``NSEMultiGrid.fgmres``
    returns for that problem alone.
    """

    def __init__(self, problems, params: _abi.PmgParams | None = None, device: int = 0, project=None, **kw):
        self._lib = _abi.load_pmg()
        self.params = params if params is not None else _abi.default_params(**kw)
        self.problems = list(problems)
        if project is None:
            project = [1 if p.config.dirichlet_faces == 63 else 0 for p in self.problems]
        self._levels = [(_abi.PmgLevelDesc * p.n_levels)(*p.descs) for p in self.problems]
        insts = (_abi.PmgInstanceDesc * len(self.problems))()
        for i, p in enumerate(self.problems):
            insts[i].n_levels = p.n_levels
            insts[i].levels = C.cast(self._levels[i], C.POINTER(_abi.PmgLevelDesc))
            insts[i].project_pressure = int(project[i])
        ctx = C.c_void_p()
        st = self._lib.pmg_batch_create(device, C.byref(self.params), len(self.problems), insts, C.byref(ctx))
        if st != 0:
            raise PmgError(st, "pmg_batch_create")
        self._ctx = ctx
        self.n_dof = [p.n_dof(p.n_levels - 1) for p in self.problems]

    def _check(self, st: int, what: str = ""):
        if st != 0:
            msg = self._lib.pmg_last_error_string(self._ctx)
            raise PmgError(st, f"{what} {msg.decode() if msg else ''}".strip())
        return st

    def autotune(self, reps: int = 20) -> list:
        """pmg_autotune on the merged levels (after the set-up)."""
        return _autotune(self, max(p.n_levels for p in self.problems), reps)

    def close(self):
        if getattr(self, "_ctx", None):
            self._lib.pmg_ctx_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def setup(self):
        self._check(self._lib.pmg_setup(self._ctx), "pmg_setup")

    def refactor(self):
        self._check(self._lib.pmg_refactor(self._ctx), "pmg_refactor")

    def __len__(self) -> int:
        return self._lib.pmg_batch_size(self._ctx)

    def fgmres(self, sols, rhss, raise_on_error: bool = True):
        """sols / rhss: one float64 vector per problem (sols updated in place)."""
        nb = len(self.problems)
        sp = (C.c_void_p * nb)(*[_ptr(s) for s in sols])
        rp = (C.c_void_p * nb)(*[_ptr(r) for r in rhss])
        it = (C.c_int * nb)()
        res = (C.c_double * nb)()
        stv = (C.c_int * nb)()
        st = self._lib.pmg_batch_fgmres_solve(self._ctx, sp, rp, it, res, stv, MEM_HOST)
        if raise_on_error:
            self._check(st, "batch fgmres")
        return list(it), list(res), list(stv)

    def synchronize(self):
        self._check(self._lib.pmg_synchronize(self._ctx))

    def set_profiling(self, on: bool):
        self._lib.pmg_set_profiling(self._ctx, 1 if on else 0)

    def timers(self) -> dict:
        return NSEMultiGrid.timers(self)

    def reset_timers(self):
        self._lib.pmg_timers_reset(self._ctx)
