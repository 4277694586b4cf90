"""Level dumps at the library boundary (pmg_level_read / pmg_level_write, pmg_vector_read
/ pmg_vector_write): one directory per level with the CSR structures and Entries arrays,
the GlobalNumbers / BeginIndex maps and the parent/child data as raw arrays (doubles as
ParMooN's SaveData writes them), plus a manifest of the sizes.  A hierarchy read from dumps
exposes ``descs`` (coarsest first) like the generated problems, so ``build_hierarchy``
and the batch context take it unchanged."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _abi


class LevelDump:
    """One level read from a dump directory; the descriptor's arrays live until close()."""

    def __init__(self, path: str):
        self._lib = _abi.load_pmg()
        self.desc = _abi.PmgLevelDesc()
        h = C.c_void_p()
        st = self._lib.pmg_level_read(os.fsencode(path), C.byref(self.desc), C.byref(h))
        if st != 0:
            raise _abi.PmgError(st, f"pmg_level_read {path}")
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            self._lib.pmg_level_read_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DumpedHierarchy:
    """Levels read from the directories ``level0`` (coarsest) ... ``level{L-1}`` of a root."""

    def __init__(self, root: str):
        self.levels = []
        k = 0
        while os.path.isdir(os.path.join(root, f"level{k}")):
            self.levels.append(LevelDump(os.path.join(root, f"level{k}")))
            k += 1
        if not self.levels:
            raise FileNotFoundError(f"no level0 directory under {root}")
        self.descs = [lv.desc for lv in self.levels]
        self.n_levels = len(self.levels)

    def close(self):
        for lv in self.levels:
            lv.close()


def write_level(path: str, desc) -> None:
    os.makedirs(path, exist_ok=True)
    st = _abi.load_pmg().pmg_level_write(os.fsencode(path), C.byref(desc))
    if st != 0:
        raise _abi.PmgError(st, f"pmg_level_write {path}")


def write_hierarchy(root: str, descs) -> None:
    for k, d in enumerate(descs):
        write_level(os.path.join(root, f"level{k}"), d)


def read_vector(path: str, n: int) -> np.ndarray:
    x = np.empty(n)
    st = _abi.load_pmg().pmg_vector_read(os.fsencode(path), x.ctypes.data_as(_abi.c_double_p), n)
    if st != 0:
        raise _abi.PmgError(st, f"pmg_vector_read {path}")
    return x


def write_vector(path: str, x: np.ndarray) -> None:
    x = np.ascontiguousarray(x, dtype=np.float64)
    st = _abi.load_pmg().pmg_vector_write(os.fsencode(path), x.ctypes.data_as(_abi.c_double_p), x.size)
    if st != 0:
        raise _abi.PmgError(st, f"pmg_vector_write {path}")
