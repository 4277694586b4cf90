"""The C ABI of libpmg without a GPU: every function declared in include/pmg.h is
exported, the ctypes mirrors have the C layout, defaults follow TDatabase::ParamDB,
loading fails loudly, and the host-only entry points (status strings, coloring) work."""
import ctypes as C
import re
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

from conftest import ROOT, _have_gpu

HEADER = ROOT / "include" / "pmg.h"


def _declared():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pmg_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    from to_merge_master_parmoon_amd import _abi
    lib = _abi.load_pmg()
    names = _declared()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing
    assert set(names) == set(_abi.PMG_FUNCS)


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_struct_layout_matches_header(tmp_path):
    from to_merge_master_parmoon_amd import _abi
    structs = {"pmg_params": _abi.PmgParams, "pmg_level_desc": _abi.PmgLevelDesc, "pmg_timers": _abi.PmgTimers}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "pmg.h"', "int main(void) {"]
    for cname, cls in structs.items():
        lines.append(f'printf("{cname} sizeof %zu\\n", sizeof({cname}));')
        for f, _ in cls._fields_:
            lines.append(f'printf("{cname} {f} %zu\\n", offsetof({cname}, {f}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", f"-I{ROOT / 'include'}", str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {}
    for ln in out:
        if ln:
            s, f, v = ln.split()
            got[(s, f)] = int(v)
    for cname, cls in structs.items():
        assert got[(cname, "sizeof")] == C.sizeof(cls)
        for f, _ in cls._fields_:
            assert got[(cname, f)] == getattr(cls, f).offset, (cname, f)


def test_param_defaults_follow_paramdb():
    from to_merge_master_parmoon_amd import _abi
    p = _abi.default_params()
    # src/General/Database.C:674-749, 871
    assert (p.sc_smoother_saddle, p.sc_pre_smooth_saddle, p.sc_post_smooth_saddle) == (2, 2, 2)
    assert p.sc_smooth_damp_factor_saddle == 1.0 and p.sc_smooth_damp_factor_coarse_saddle == 1.0
    assert p.sc_gmg_damp_factor_saddle == 1.0 and p.internal_project_pressure == 1
    assert (p.sc_gmres_restart, p.sc_lin_maxit_saddle, p.sc_minit, p.sc_amg_prec_it) == (10, 10000, 0, 1)
    assert p.sc_lin_red_factor_saddle == 0.0 and p.sc_lin_res_norm_min_saddle == 1e-10
    assert p.sc_div_factor == 1e10
    # the only coarse solver of this path: the dense direct solve (17); ParamDB's default
    # (2) selects an iterative smoother that this library does not provide
    assert p.sc_coarse_smoother_saddle == 17
    assert p.precision == _abi.PREC_FP64


def test_status_strings():
    from to_merge_master_parmoon_amd import _abi
    lib = _abi.load_pmg()
    seen = set()
    for code in _abi.STATUS:
        s = lib.pmg_status_string(code)
        assert s and s not in seen
        seen.add(s)
    assert lib.pmg_status_string(12345)


def test_missing_library_fails_loudly():
    from to_merge_master_parmoon_amd import _abi
    with pytest.raises(ImportError, match="no fallback"):
        _abi._load("libpmg_missing.so", {})


@pytest.mark.skipif(_have_gpu(), reason="checks the error path without a device")
def test_ctx_create_without_device_reports_error():
    from to_merge_master_parmoon_amd import _abi
    lib = _abi.load_pmg()
    ctx = C.c_void_p()
    p = _abi.default_params()
    st = lib.pmg_ctx_create(0, C.byref(p), C.byref(ctx))
    assert st in (2, 1)  # PMG_ERR_DEVICE (runtime error) or INVALID_ARGUMENT (no such device)
    assert not ctx.value


def test_invalid_params_rejected_before_device():
    from to_merge_master_parmoon_amd import _abi
    lib = _abi.load_pmg()
    ctx = C.c_void_p()
    p = _abi.default_params(sc_gmres_restart=0, sc_lin_maxit_saddle=0)
    p.precision = 7
    assert lib.pmg_ctx_create(0, C.byref(p), C.byref(ctx)) == 1
    assert not ctx.value


def test_smoother_choices_checked_before_device():
    """Smoothers 1-4 exist (cell Vanka in the Gauss-Seidel / colored order, nodal Vanka with
    the diagonal / full local system); the nodal ones run in FP64 only, anything else is
    PMG_ERR_UNSUPPORTED (10) before a device is touched."""
    from to_merge_master_parmoon_amd import _abi
    lib = _abi.load_pmg()
    for sm, prec in ((0, 0), (5, 0), (3, 1), (4, 2)):
        ctx = C.c_void_p()
        p = _abi.default_params(sc_smoother_saddle=sm, precision=prec)
        assert lib.pmg_ctx_create(0, C.byref(p), C.byref(ctx)) == 10, (sm, prec)
        assert not ctx.value


def test_refine_steps_checked_before_device():
    """bf16_refine_steps: 0..3 (else PMG_ERR_INVALID_ARGUMENT), and only with the stored
    inverses of the mixed forms (FP64 or patch_inverse = 0: PMG_ERR_UNSUPPORTED)."""
    from to_merge_master_parmoon_amd import _abi
    lib = _abi.load_pmg()
    for kw, want in ((dict(bf16_refine_steps=4, precision=2), 1), (dict(bf16_refine_steps=-1, precision=2), 1),
                     (dict(bf16_refine_steps=1, precision=0), 10),
                     (dict(bf16_refine_steps=1, precision=2, patch_inverse=0), 10)):
        ctx = C.c_void_p()
        p = _abi.default_params(**kw)
        assert lib.pmg_ctx_create(0, C.byref(p), C.byref(ctx)) == want, kw
        assert not ctx.value
