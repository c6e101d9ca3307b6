"""ORACLE binding -- test infrastructure only (tests/, __graft_entry__.smoke(), bench.py cpu_baseline).

Loads oracle/liboracle.so (the CPU restatement of the reference's CPU backend) and, when built,
oracle/_ref/libref_kernels.so (the reference's own kernel headers compiled in place). Both expose
the gsk_backend_ops table, so the library's solver orchestration can drive them.
"""

from __future__ import annotations

import ctypes as C
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_LIB = os.path.join(HERE, "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libref_kernels.so")

_libs: dict = {}


def build(reference: str = "/root/reference") -> None:
    """Compile the oracle (and oracle/_ref, the reference's own kernels, when the reference tree exists)."""
    subprocess.run(["make", "-s", "-C", HERE, f"REFERENCE={reference}"], check=True)


def _load(path: str, prefix: str):
    if path in _libs:
        return _libs[path]
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} is not built (run make -C {HERE})")
    L = C.CDLL(path)
    getattr(L, f"{prefix}_create").restype = C.c_void_p
    getattr(L, f"{prefix}_create").argtypes = [C.c_void_p, C.c_int]
    getattr(L, f"{prefix}_destroy").argtypes = [C.c_void_p]
    getattr(L, f"{prefix}_fill_ops").argtypes = [C.c_void_p, C.c_void_p]
    if prefix == "orc":
        L.orc_last_error.restype = C.c_char_p
        L.orc_set_num_threads.argtypes = [C.c_int]
    _libs[path] = L
    return L


def ref_available() -> bool:
    return os.path.exists(REF_LIB)


class CpuBackend:
    """A CPU backend for the tests: kind = "oracle" (restatement) or "ref" (reference kernels)."""

    def __init__(self, setup, kind: str = "oracle", precision: int = -1, ops_type=None):
        path, prefix = (ORACLE_LIB, "orc") if kind == "oracle" else (REF_LIB, "ref")
        self._L = _load(path, prefix)
        self._prefix = prefix
        view = setup.view()
        self._setup = setup
        h = getattr(self._L, f"{prefix}_create")(C.byref(view), precision)
        if not h:
            msg = self._L.orc_last_error().decode() if prefix == "orc" else "create failed"
            raise RuntimeError(msg)
        self._h = h
        self.ops = ops_type() if ops_type is not None else None
        if self.ops is not None:
            st = getattr(self._L, f"{prefix}_fill_ops")(C.c_void_p(h), C.byref(self.ops))
            if st != 0:
                raise RuntimeError("fill_ops failed")

    def close(self):
        if getattr(self, "_h", None):
            getattr(self._L, f"{self._prefix}_destroy")(C.c_void_p(self._h))
            self._h = None

    def __del__(self):
        self.close()


def set_threads(n: int) -> int:
    L = _load(ORACLE_LIB, "orc")
    return L.orc_set_num_threads(n)
