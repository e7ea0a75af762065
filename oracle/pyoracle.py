"""ctypes binding of the CPU oracle (oracle/lssph_oracle.c).

TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker.  The oracle library is built by
oracle/Makefile into oracle/build/.
"""
from __future__ import annotations

import ctypes
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "build" / "liblssph_oracle.so"
NATIVE_LIB_PATH = HERE / "build" / "liblssph_oracle_native.so"

_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int32)
_libs: dict = {}


def load(native: bool = False):
    path = NATIVE_LIB_PATH if native else LIB_PATH
    if path in _libs:
        return _libs[path]
    if not path.exists():
        raise FileNotFoundError(f"{path} not built (run make -C oracle)")
    lib = ctypes.CDLL(str(path))
    vp = ctypes.c_void_p
    lib.orc_create.restype = vp
    lib.orc_create.argtypes = [vp, ctypes.c_int32, ctypes.c_int32]
    lib.orc_destroy.argtypes = [vp]
    lib.orc_error.restype = ctypes.c_char_p
    lib.orc_error.argtypes = [vp]
    lib.orc_set_threads.argtypes = [ctypes.c_int32]
    lib.orc_set_state.argtypes = [vp, _I] + [_D] * 7
    lib.orc_get_state.argtypes = [vp, _I] + [_D] * 7
    for name in ("orc_init_vm", "orc_init_cell", "orc_update_cell", "orc_ghost",
                 "orc_calculate_rhs", "orc_rk2", "orc_rk4", "orc_pst"):
        getattr(lib, name).argtypes = [vp]
        getattr(lib, name).restype = ctypes.c_int32
    lib.orc_get_rhs.argtypes = [vp] + [_D] * 4
    lib.orc_check_steady_state.argtypes = [vp, _I, _D]
    lib.orc_get_vm.argtypes = [vp, _I] + [_D] * 6
    lib.orc_get_cell_index.argtypes = [vp, _I]
    lib.orc_neighbors.argtypes = [vp, _I, _I, ctypes.c_int64, ctypes.POINTER(ctypes.c_int64)]
    lib.orc_neighbors.restype = ctypes.c_int32
    lib.orc_kernel_w.restype = ctypes.c_double
    lib.orc_kernel_w.argtypes = [ctypes.c_int32, ctypes.c_double, ctypes.c_double]
    lib.orc_kernel_dw.restype = ctypes.c_double
    lib.orc_kernel_dw.argtypes = [ctypes.c_int32, ctypes.c_double, ctypes.c_double]
    lib.orc_dgesv.argtypes = [ctypes.c_int32, ctypes.c_int32, _D, _I, _D]
    lib.orc_dgesv.restype = ctypes.c_int32
    _libs[path] = lib
    return lib


def _dp(a):
    return None if a is None else a.ctypes.data_as(_D)


def _ip(a):
    return a.ctypes.data_as(_I)


class OracleSim:
    """One oracle instance; method names follow the reference subroutines."""

    def __init__(self, params, num_int: int, num_ext: int, native: bool = False):
        self.lib = load(native)
        self.params = params
        self.num_int, self.num_ext = num_int, num_ext
        self.has_tem = params.target_problem == 4
        self.h = self.lib.orc_create(ctypes.byref(params), num_int, num_ext)
        if not self.h:
            raise MemoryError("orc_create failed")

    def close(self):
        if self.h:
            self.lib.orc_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _check(self, rc):
        if rc != 0:
            raise RuntimeError(f"oracle error {rc}: {self.lib.orc_error(self.h).decode()}")

    @staticmethod
    def set_threads(n: int):
        load().orc_set_threads(n)

    def set_state(self, st):
        arr = [np.ascontiguousarray(a, dtype=np.float64) for a in (st.x, st.y, st.u, st.v, st.rho, st.pre)]
        tem = None if st.tem is None else np.ascontiguousarray(st.tem, dtype=np.float64)
        pt = np.ascontiguousarray(st.ptype, dtype=np.int32)
        self._check(self.lib.orc_set_state(self.h, _ip(pt), *(_dp(a) for a in arr), _dp(tem)))

    def get_state(self) -> dict:
        n = self.num_int + self.num_ext
        out = {"ptype": np.zeros(n, np.int32)}
        for k in ("x", "y", "u", "v", "rho", "pre"):
            out[k] = np.zeros(n)
        out["tem"] = np.zeros(n) if self.has_tem else None
        self.lib.orc_get_state(self.h, _ip(out["ptype"]), *(_dp(out[k]) for k in
                                                           ("x", "y", "u", "v", "rho", "pre", "tem")))
        return out

    def init(self, st):
        """set_state + init_virtual_markers + init_cell/update_cell."""
        self.set_state(st)
        self._check(self.lib.orc_init_vm(self.h))
        self._check(self.lib.orc_init_cell(self.h))

    def update_cell(self):
        self._check(self.lib.orc_update_cell(self.h))

    def ghost(self):
        self._check(self.lib.orc_ghost(self.h))

    def calculate_rhs(self) -> dict:
        self._check(self.lib.orc_calculate_rhs(self.h))
        n = self.num_int
        out = {k: np.zeros(n) for k in ("u", "v", "rho", "tem")}
        self.lib.orc_get_rhs(self.h, *(_dp(out[k]) for k in ("u", "v", "rho", "tem")))
        if not self.has_tem:
            del out["tem"]
        return out

    def rk2(self):
        self._check(self.lib.orc_rk2(self.h))

    def rk4(self):
        self._check(self.lib.orc_rk4(self.h))

    def pst(self):
        self._check(self.lib.orc_pst(self.h))

    def step(self, n: int = 1):
        """n iterations of main.f90:119-134 (RK, then PST unless diffusion)."""
        for _ in range(n):
            (self.rk2 if self.params.rk == 2 else self.rk4)()
            if self.params.target_problem != 1:
                self.pst()

    def check_steady_state(self):
        flag = np.zeros(1, np.int32)
        df = np.zeros(1)
        self._check(self.lib.orc_check_steady_state(self.h, _ip(flag), _dp(df)))
        return int(flag[0]), float(df[0])

    def get_vm(self) -> dict:
        n = self.num_ext
        out = {"ptype": np.zeros(n, np.int32)}
        for k in ("x", "y", "u", "v", "pre", "tem"):
            out[k] = np.zeros(n)
        self.lib.orc_get_vm(self.h, _ip(out["ptype"]), *(_dp(out[k]) for k in
                                                        ("x", "y", "u", "v", "pre", "tem")))
        if not self.has_tem:
            out["tem"] = None
        return out

    def cell_index(self) -> np.ndarray:
        out = np.zeros(self.num_int + self.num_ext, np.int32)
        self.lib.orc_get_cell_index(self.h, _ip(out))
        return out

    def neighbors(self):
        counts = np.zeros(self.num_int, np.int32)
        cap = np.int64(self.num_int) * 128
        lst = np.zeros(int(cap), np.int32)
        total = ctypes.c_int64()
        self._check(self.lib.orc_neighbors(self.h, _ip(counts), _ip(lst), cap, ctypes.byref(total)))
        return counts, lst[: total.value].copy()


def kernel_w(kernel_type: int, r: float, h: float) -> float:
    return load().orc_kernel_w(kernel_type, r, h)


def kernel_dw(kernel_type: int, r: float, h: float) -> float:
    return load().orc_kernel_dw(kernel_type, r, h)


def dgesv(a: np.ndarray, b: np.ndarray):
    """(info, x) with LAPACK DGESV semantics; a is (n, n), b is (n, nrhs)."""
    n = a.shape[0]
    af = np.asfortranarray(a, dtype=np.float64).copy(order="F")
    bf = np.asfortranarray(b.reshape(n, -1), dtype=np.float64).copy(order="F")
    ipiv = np.zeros(n, np.int32)
    info = load().orc_dgesv(n, bf.shape[1], af.ctypes.data_as(_D), _ip(ipiv), bf.ctypes.data_as(_D))
    return int(info), bf
