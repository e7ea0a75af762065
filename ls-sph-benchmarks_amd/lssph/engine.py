"""ctypes binding of lib/liblssph_hip.so (include/lssph.h).

`Simulation` owns one device handle.  Its methods carry the names of the
reference subroutines they replace (init_cell, update_cell, ghost,
calculate_RHS, RK2, RK4, PST, check_steady_state; see the C header for the
file:line of each) and raise LssphError where the C call returns an error.
There is no CPU fallback: without the built library or a GPU every call fails.
"""
from __future__ import annotations

import ctypes
from pathlib import Path
from typing import Sequence

import numpy as np

from .config import Param, Params
from .particles import State

PKG = Path(__file__).resolve().parents[1]
LIB_PATH = PKG / "lib" / "liblssph_hip.so"

LSSPH_FP64 = 0
LSSPH_MIXED = 1

_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int32)
_lib = None


class LssphError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"lssph error {code}: {msg}")
        self.code = code


class Stats(ctypes.Structure):
    _fields_ = [("fp64_fallbacks", ctypes.c_int64), ("list_overflows", ctypes.c_int64),
                ("max_neighbors", ctypes.c_int32), ("max_cell_count", ctypes.c_int32),
                ("escaped", ctypes.c_int32), ("pad", ctypes.c_int32)]


class KernelTimes(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in ("lsa_ms", "pst_ms", "cells_ms", "ghost_ms", "other_ms")] + [
        (n, ctypes.c_int64) for n in ("lsa_launches", "pst_launches", "cells_launches", "ghost_launches",
                                      "other_launches")]


EXPORTS = ("lssph_create", "lssph_destroy", "lssph_last_error", "lssph_upload", "lssph_download",
           "lssph_init_cell", "lssph_update_cell", "lssph_ghost", "lssph_calculate_rhs",
           "lssph_download_rhs", "lssph_rk2", "lssph_rk4", "lssph_pst", "lssph_step",
           "lssph_check_steady_state", "lssph_download_vm", "lssph_download_cell_index",
           "lssph_neighbors", "lssph_stats", "lssph_reset_stats", "lssph_synchronize",
           "lssph_set_tuning", "lssph_set_skin", "lssph_set_scan", "lssph_stream", "lssph_profile", "lssph_kernel_times",
           "lssph_pair_lists", "lssph_probe", "lssph_set_isolation", "lssph_instance_status",
           "lssph_probe_static", "lssph_set_wave_particles", "lssph_set_cell_split")


def load_library(path: Path | None = None):
    """Loads the HIP library (fails loudly when it is not built)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise FileNotFoundError(f"{p} is not built; run `make -C ls-sph-benchmarks_amd` or "
                                "__graft_entry__.build()")
    lib = ctypes.CDLL(str(p))
    vp = ctypes.c_void_p
    i32 = ctypes.c_int32
    lib.lssph_create.argtypes = [ctypes.POINTER(Params), i32, _I, _I, ctypes.POINTER(vp)]
    lib.lssph_destroy.argtypes = [vp]
    lib.lssph_last_error.restype = ctypes.c_char_p
    lib.lssph_last_error.argtypes = [vp]
    lib.lssph_upload.argtypes = [vp, _I] + [_D] * 7
    lib.lssph_download.argtypes = [vp, _I] + [_D] * 7
    for name in ("lssph_init_cell", "lssph_update_cell", "lssph_ghost", "lssph_calculate_rhs",
                 "lssph_rk2", "lssph_rk4", "lssph_pst", "lssph_synchronize", "lssph_reset_stats"):
        getattr(lib, name).argtypes = [vp]
    lib.lssph_download_rhs.argtypes = [vp] + [_D] * 4
    lib.lssph_step.argtypes = [vp, i32]
    lib.lssph_check_steady_state.argtypes = [vp, _I, _D]
    lib.lssph_download_vm.argtypes = [vp, _I] + [_D] * 6
    lib.lssph_download_cell_index.argtypes = [vp, _I]
    lib.lssph_neighbors.argtypes = [vp, _I, _I, ctypes.c_int64, ctypes.POINTER(ctypes.c_int64)]
    lib.lssph_stats.argtypes = [vp, ctypes.POINTER(Stats)]
    lib.lssph_set_tuning.argtypes = [vp, i32, i32, i32]
    lib.lssph_set_skin.argtypes = [vp, ctypes.c_double]
    lib.lssph_set_scan.argtypes = [vp, i32]
    lib.lssph_stream.restype = vp
    lib.lssph_stream.argtypes = [vp]
    lib.lssph_profile.argtypes = [vp, i32]
    lib.lssph_kernel_times.argtypes = [vp, ctypes.POINTER(KernelTimes)]
    lib.lssph_pair_lists.argtypes = [vp, i32, _I, _I, _I]
    lib.lssph_probe.argtypes = [vp, i32, i32, i32, i32, _D, _D, _D]
    lib.lssph_probe_static.argtypes = [vp, i32, i32, i32, i32, _D, _D, _D]
    lib.lssph_set_isolation.argtypes = [vp, i32]
    lib.lssph_set_wave_particles.argtypes = [vp, i32]
    lib.lssph_set_cell_split.argtypes = [vp, i32]
    lib.lssph_instance_status.argtypes = [vp, _I, _I]
    for name in EXPORTS:
        if name not in ("lssph_last_error", "lssph_stream"):
            getattr(lib, name).restype = i32
    if path is None:
        _lib = lib
    return lib


LSSPH_E_CAPACITY = -6  # lssph.h: caller buffer too small


def _dp(a):
    return None if a is None else a.ctypes.data_as(_D)


def _ip(a):
    return None if a is None else a.ctypes.data_as(_I)


class Simulation:
    """Device-resident LS-SPH state of one box or an ensemble of boxes.

    params: a Param (one instance) or a list of Params / lssph_params_t;
    states: matching particle States.  Upload order: the fluid blocks of all
    instances, then their wall blocks."""

    def __init__(self, params, num_int: Sequence[int], num_ext: Sequence[int], precision: int | None = None,
                 skin: float | None = None, cell_split: int | None = None):
        self.lib = load_library()
        plist = params if isinstance(params, (list, tuple)) else [params]
        cparams = [p.to_c(precision) if isinstance(p, Param) else p for p in plist]
        if precision is not None:
            for c in cparams:
                c.precision = precision
        self.n_inst = len(cparams)
        arr = (Params * self.n_inst)(*cparams)
        self.params = cparams
        self.num_int = [int(n) for n in num_int]
        self.num_ext = [int(n) for n in num_ext]
        if len(self.num_int) != self.n_inst or len(self.num_ext) != self.n_inst:
            raise ValueError(f"{self.n_inst} parameter sets but {len(self.num_int)} fluid and "
                             f"{len(self.num_ext)} wall counts")
        self.n_int = sum(self.num_int)
        self.n_ext = sum(self.num_ext)
        self.n_total = self.n_int + self.n_ext
        self.has_tem = cparams[0].target_problem == 4
        ni = np.asarray(self.num_int, np.int32)
        ne = np.asarray(self.num_ext, np.int32)
        h = ctypes.c_void_p()
        rc = self.lib.lssph_create(arr, self.n_inst, _ip(ni), _ip(ne), ctypes.byref(h))
        if rc != 0:
            raise LssphError(rc, self.lib.lssph_last_error(None).decode())
        self.h = h
        if skin is not None:
            self.set_skin(skin)
        if cell_split is not None:
            self.set_cell_split(cell_split)

    # -- lifecycle ------------------------------------------------------------
    def close(self):
        if getattr(self, "h", None):
            self.lib.lssph_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc):
        if rc != 0:
            raise LssphError(rc, self.lib.lssph_last_error(self.h).decode())

    # -- transfers --------------------------------------------------------------
    def upload(self, st) -> None:
        """st: a State, a list of States (one per instance), or a dict of arrays
        already in upload order."""
        if isinstance(st, (list, tuple)):
            st = concat_states(st)
        get = (lambda k: st[k]) if isinstance(st, dict) else (lambda k: getattr(st, k))
        arrs = [np.ascontiguousarray(get(k), dtype=np.float64) for k in ("x", "y", "u", "v", "rho", "pre")]
        tem = get("tem") if self.has_tem else None
        tem = None if tem is None else np.ascontiguousarray(tem, dtype=np.float64)
        pt = np.ascontiguousarray(get("ptype"), dtype=np.int32)
        for name, a in zip(("ptype", "x", "y", "u", "v", "rho", "pre", "tem"), [pt] + arrs + [tem]):
            if a is not None and a.size != self.n_total:
                raise ValueError(f"state {name} has {a.size} entries, handle expects {self.n_total}")
        self._check(self.lib.lssph_upload(self.h, _ip(pt), *(_dp(a) for a in arrs), _dp(tem)))

    def download(self) -> dict:
        n = self.n_total
        out = {"ptype": np.zeros(n, np.int32)}
        for k in ("x", "y", "u", "v", "rho", "pre"):
            out[k] = np.zeros(n)
        out["tem"] = np.zeros(n) if self.has_tem else None
        self._check(self.lib.lssph_download(self.h, _ip(out["ptype"]), *(_dp(out[k]) for k in
                                                                         ("x", "y", "u", "v", "rho", "pre",
                                                                          "tem"))))
        return out

    def download_vm(self) -> dict:
        m = self.n_ext
        out = {"ptype": np.zeros(m, np.int32)}
        for k in ("x", "y", "u", "v", "pre"):
            out[k] = np.zeros(m)
        out["tem"] = np.zeros(m) if self.has_tem else None
        self._check(self.lib.lssph_download_vm(self.h, _ip(out["ptype"]), *(_dp(out[k]) for k in
                                                                           ("x", "y", "u", "v", "pre", "tem"))))
        return out

    def cell_index(self) -> np.ndarray:
        out = np.zeros(self.n_total, np.int32)
        self._check(self.lib.lssph_download_cell_index(self.h, _ip(out)))
        return out

    # -- reference subroutines ----------------------------------------------------
    def init_cell(self):
        self._check(self.lib.lssph_init_cell(self.h))

    def update_cell(self):
        self._check(self.lib.lssph_update_cell(self.h))

    def ghost(self):
        self._check(self.lib.lssph_ghost(self.h))

    def calculate_RHS(self) -> dict:
        self._check(self.lib.lssph_calculate_rhs(self.h))
        out = {k: np.zeros(self.n_int) for k in ("u", "v", "rho", "tem")}
        self._check(self.lib.lssph_download_rhs(self.h, *(_dp(out[k]) for k in ("u", "v", "rho", "tem"))))
        if not self.has_tem:
            del out["tem"]
        return out

    def RK2(self):
        self._check(self.lib.lssph_rk2(self.h))

    def RK4(self):
        self._check(self.lib.lssph_rk4(self.h))

    def PST(self):
        self._check(self.lib.lssph_pst(self.h))

    def step(self, nsteps: int = 1):
        """nsteps of the main.f90 loop body (RK, then PST unless diffusion),
        replayed from a captured HIP graph."""
        self._check(self.lib.lssph_step(self.h, int(nsteps)))

    def eager_step(self, nsteps: int = 1):
        """Same as step() but launched call by call (no graph)."""
        p = self.params[0]
        for _ in range(nsteps):
            (self.RK2 if p.rk == 2 else self.RK4)()
            if p.target_problem != 1:
                self.PST()

    def check_steady_state(self):
        flags = np.zeros(self.n_inst, np.int32)
        df = np.zeros(self.n_inst)
        self._check(self.lib.lssph_check_steady_state(self.h, _ip(flags), _dp(df)))
        return flags, df

    def neighbors(self):
        counts = np.zeros(self.n_int, np.int32)
        total = ctypes.c_int64()
        rc = self.lib.lssph_neighbors(self.h, _ip(counts), None, 0, ctypes.byref(total))
        if rc != LSSPH_E_CAPACITY:  # the sizing call (no list) expects only "capacity too small"
            self._check(rc)
        lst = np.zeros(max(total.value, 1), np.int32)
        self._check(self.lib.lssph_neighbors(self.h, _ip(counts), _ip(lst), total.value, ctypes.byref(total)))
        return counts, lst[: total.value]

    def pair_lists(self, cap: int = 128):
        """The pair kernels' neighbor lists as their walk visits them:
        (counts, list[n_int, cap] of upload indices, scan path per particle)."""
        counts = np.zeros(self.n_int, np.int32)
        path = np.zeros(self.n_int, np.int32)
        lst = np.zeros((self.n_int, cap), np.int32)
        self._check(self.lib.lssph_pair_lists(self.h, int(cap), _ip(counts), _ip(lst), _ip(path)))
        return counts, lst, path

    PROBE_FIELDS = {"u": 0, "v": 1, "rho": 2, "pre": 3, "tem": 4}

    def probe(self, x, y, field: str = "tem", q: int = 6, inst: int = 0) -> np.ndarray:
        """Type-B LS fit of a field at probe points (upload coordinates):
        an (n, 3) array of value, d/dx, d/dy."""
        x = np.ascontiguousarray(x, np.float64)
        y = np.ascontiguousarray(y, np.float64)
        if x.shape != y.shape or x.ndim != 1:
            raise ValueError("probe: x and y must be 1-D arrays of one length")
        out = np.zeros((x.size, 3))
        self._check(self.lib.lssph_probe(self.h, int(inst), self.PROBE_FIELDS[field], int(q), int(x.size),
                                         _dp(x), _dp(y), _dp(out)))
        return out

    def probe_static(self, x, y, field: str = "tem", q: int = 6, inst: int = 0) -> np.ndarray:
        """probe() through the allocation-free entry point (at most 16384 probes)."""
        x = np.ascontiguousarray(x, np.float64)
        y = np.ascontiguousarray(y, np.float64)
        if x.shape != y.shape or x.ndim != 1:
            raise ValueError("probe: x and y must be 1-D arrays of one length")
        out = np.zeros((x.size, 3))
        self._check(self.lib.lssph_probe_static(self.h, int(inst), self.PROBE_FIELDS[field], int(q), int(x.size),
                                                _dp(x), _dp(y), _dp(out)))
        return out

    def set_isolation(self, on: bool = True):
        """Ensembles: retire a diverged instance instead of failing the handle."""
        self._check(self.lib.lssph_set_isolation(self.h, 1 if on else 0))

    RETIRED_SKIN = 1     # lssph.h LSSPH_RETIRED_SKIN
    RETIRED_ESCAPED = 2  # lssph.h LSSPH_RETIRED_ESCAPED

    def instance_status(self, reasons: bool = False):
        """Per instance: 0 while it runs, else the step during which it was
        retired; with reasons=True also the cause (RETIRED_SKIN, RETIRED_ESCAPED)."""
        out = np.zeros(self.n_inst, np.int32)
        why = np.zeros(self.n_inst, np.int32)
        self._check(self.lib.lssph_instance_status(self.h, _ip(out), _ip(why)))
        return (out, why) if reasons else out

    def stats(self) -> Stats:
        s = Stats()
        self._check(self.lib.lssph_stats(self.h, ctypes.byref(s)))
        return s

    def reset_stats(self):
        self._check(self.lib.lssph_reset_stats(self.h))

    def synchronize(self):
        self._check(self.lib.lssph_synchronize(self.h))

    def set_tuning(self, block: int = 0, list_len: int = 0, reorder_interval: int = 0):
        self._check(self.lib.lssph_set_tuning(self.h, block, list_len, reorder_interval))

    def set_skin(self, skin: float):
        """Cell-list skin as a fraction of h_eff (0: rebuild after every pass)."""
        self._check(self.lib.lssph_set_skin(self.h, ctypes.c_double(skin)))

    def set_cell_split(self, columns: int):
        """Cell columns per stencil reach in x (1-4): the stencil is then 2 columns + 1
        cells wide and tests fewer candidates; the neighbor sets do not change."""
        self._check(self.lib.lssph_set_cell_split(self.h, int(columns)))

    def set_wave_particles(self, ppw: int):
        """Particles per wavefront of the SPH_MODEL 6 kernels (1, 2; 0: default)."""
        self._check(self.lib.lssph_set_wave_particles(self.h, int(ppw)))

    def set_scan(self, staged: bool):
        """Wave-staged candidate scan on (default) or off; results are bitwise identical."""
        self._check(self.lib.lssph_set_scan(self.h, 1 if staged else 0))

    def profile(self, enable: bool = True):
        self._check(self.lib.lssph_profile(self.h, 1 if enable else 0))

    def kernel_times(self) -> KernelTimes:
        t = KernelTimes()
        self._check(self.lib.lssph_kernel_times(self.h, ctypes.byref(t)))
        return t

    @property
    def stream(self) -> int:
        return self.lib.lssph_stream(self.h)


def concat_states(states: Sequence[State]) -> dict:
    """Upload-order arrays of an ensemble: all fluid blocks, then all walls."""
    keys = ("ptype", "x", "y", "u", "v", "rho", "pre")
    out = {}
    for k in keys:
        out[k] = np.concatenate([getattr(s, k)[: s.num_int] for s in states] +
                                [getattr(s, k)[s.num_int:] for s in states])
    if states[0].tem is not None:
        out["tem"] = np.concatenate([s.tem[: s.num_int] for s in states] + [s.tem[s.num_int:] for s in states])
    else:
        out["tem"] = None
    return out


def split_states(arrs: dict, num_int: Sequence[int], num_ext: Sequence[int]) -> list:
    """Inverse of concat_states: per-instance dicts in the reference order."""
    out = []
    fi, wi = 0, int(sum(num_int))
    for ni, ne in zip(num_int, num_ext):
        d = {}
        for k, a in arrs.items():
            if a is None:
                d[k] = None
                continue
            d[k] = np.concatenate([a[fi:fi + ni], a[wi:wi + ne]])
        out.append(d)
        fi += ni
        wi += ne
    return out


def simulation_from_state(param: Param, st: State, precision: int | None = None,
                          skin: float | None = None, cell_split: int | None = None) -> Simulation:
    """Creates a handle for one box, uploads st and builds the cell list."""
    sim = Simulation(param, [st.num_int], [st.num_ext], precision=precision, skin=skin, cell_split=cell_split)
    sim.upload(st)
    sim.init_cell()
    return sim
