"""Configuration: config.h macros -> Param (derived values) -> lssph_params_t.

The reference configures every run with compile-time macros in config.h
(config.h:24-213) and derives the remaining parameters in init_param
(source/setup/init_param_mod.f90:24-201).  Here the same macro names are read
at run time (from a dict or from a config.h file) and the same derivations
are made, so a reference config.h drives this package unchanged.
"""
from __future__ import annotations

import ctypes
import dataclasses
import math
import re
from pathlib import Path

PI = math.acos(-1.0)

# Macro name -> (attribute, type, default).  Defaults are the reference's
# shipped config.h values (config.h:24-212).
MACROS = {
    "TARGET_PROBLEM": ("target_problem", int, 2),
    "U_BOUNDARY_TOP": ("u_bc_top", int, 1),
    "U_BOUNDARY_BOTTOM": ("u_bc_bottom", int, 1),
    "U_BOUNDARY_LEFT": ("u_bc_left", int, 1),
    "U_BOUNDARY_RIGHT": ("u_bc_right", int, 1),
    "U_TOP": ("u_top", float, 1.0),
    "TEM_BOUNDARY_TOP": ("tem_bc_top", int, 1),
    "TEM_BOUNDARY_BOTTOM": ("tem_bc_bottom", int, 1),
    "TEM_BOUNDARY_LEFT": ("tem_bc_left", int, 2),
    "TEM_BOUNDARY_RIGHT": ("tem_bc_right", int, 2),
    "TEM_TOP": ("tem_top", float, 0.0),
    "TEM_BOTTOM": ("tem_bottom", float, 1.0e-2),
    "TEM_LEFT": ("tem_left", float, 0.0),
    "TEM_RIGHT": ("tem_right", float, 0.0),
    "SAVE_NAME": ("save_name", str, "run"),
    "READ_NAME": ("read_name", str, "new"),
    "OMP_THREADS": ("omp_threads", int, 8),
    "LEN_X": ("len_x", float, 1.0),
    "LEN_Y": ("len_y", float, 1.0),
    "NUM_X": ("num_x", int, 100),
    "NUM_Y": ("num_y", int, 100),
    "RK": ("rk", int, 4),
    "START_STEP": ("start_step", int, 1),
    "END_STEP": ("end_step", int, 10000),
    "WRITE_STEP": ("write_step", int, 100),
    "THRESHOLD": ("threshold", float, 1.0e-10),
    "COE_CFL": ("coe_cfl", float, 0.5),
    "COE_DIF": ("coe_dif", float, 0.2),
    "ZETA_RSST": ("zeta_rsst", float, 1.0),
    "XI_VIM": ("xi_vim", float, 1.0),
    "RHO_REF": ("rho_ref", float, 1.0),
    "VIS_REF": ("vis_ref", float, 1.0e-2),
    "K_REF": ("k_ref", float, 1.0e2),
    "K_TH_REF": ("k_th_ref", float, 1.0e-2),
    "CP_REF": ("cp_ref", float, 1.0e3),
    "ALPHA_REF": ("alpha_ref", float, 7.1e-3),
    "GRAVITY": ("gravity", float, 10.0),
    "POS_PERT": ("pos_pert", float, 0.0),
    "TG_A": ("tg_a", int, 1),
    "TG_B": ("tg_b", int, 1),
    "COE_H": ("coe_h", float, 1.2),
    "PST_C": ("pst_c", float, 1.5),
    "COE_DELTA_SPH": ("coe_delta_sph", float, 0.0),
    "KERNEL_TYPE": ("kernel_type", int, 3),
    "SPH_MODEL": ("sph_model", int, 3),
    "WALL_MODEL": ("wall_model", int, 2),
    # Extensions of this package (no reference macro): periodic box, precision,
    # diffusion-test scaling, explicit time step, mixed-precision fallback.
    "PERIODIC": ("periodic", int, 0),
    "PRECISION": ("precision", int, 0),
    "DIFF_COEF": ("diff_coef", float, 1.0),
    "DIFF_SOURCE": ("diff_source", float, 1.0),
    "DT": ("dt_override", float, 0.0),
    "COND_MAX": ("cond_max", float, 1.0e4),
}


def _fortran_float(text: str) -> float:
    return float(text.strip().lower().replace("d", "e"))


def _coerce(kind, value):
    if kind is float:
        return _fortran_float(value) if isinstance(value, str) else float(value)
    if kind is int:
        return int(value)
    return str(value).strip().strip('"')


@dataclasses.dataclass
class Config:
    """User-defined configuration: one attribute per config.h macro."""

    target_problem: int = 2
    u_bc_top: int = 1
    u_bc_bottom: int = 1
    u_bc_left: int = 1
    u_bc_right: int = 1
    u_top: float = 1.0
    tem_bc_top: int = 1
    tem_bc_bottom: int = 1
    tem_bc_left: int = 2
    tem_bc_right: int = 2
    tem_top: float = 0.0
    tem_bottom: float = 1.0e-2
    tem_left: float = 0.0
    tem_right: float = 0.0
    save_name: str = "run"
    read_name: str = "new"
    omp_threads: int = 8
    len_x: float = 1.0
    len_y: float = 1.0
    num_x: int = 100
    num_y: int = 100
    rk: int = 4
    start_step: int = 1
    end_step: int = 10000
    write_step: int = 100
    threshold: float = 1.0e-10
    coe_cfl: float = 0.5
    coe_dif: float = 0.2
    zeta_rsst: float = 1.0
    xi_vim: float = 1.0
    rho_ref: float = 1.0
    vis_ref: float = 1.0e-2
    k_ref: float = 1.0e2
    k_th_ref: float = 1.0e-2
    cp_ref: float = 1.0e3
    alpha_ref: float = 7.1e-3
    gravity: float = 10.0
    pos_pert: float = 0.0
    tg_a: int = 1
    tg_b: int = 1
    coe_h: float = 1.2
    pst_c: float = 1.5
    coe_delta_sph: float = 0.0
    kernel_type: int = 3
    sph_model: int = 3
    wall_model: int = 2
    periodic: int = 0
    precision: int = 0
    diff_coef: float = 1.0
    diff_source: float = 1.0
    dt_override: float = 0.0
    cond_max: float = 1.0e4

    @classmethod
    def from_macros(cls, macros: dict) -> "Config":
        cfg = cls()
        for name, value in macros.items():
            if name not in MACROS:
                raise KeyError(f"unknown config macro {name}")
            attr, kind, _ = MACROS[name]
            setattr(cfg, attr, _coerce(kind, value))
        return cfg

    def replace(self, **kw) -> "Config":
        return dataclasses.replace(self, **kw)


_DEFINE = re.compile(r"^\s*#\s*define\s+([A-Z_][A-Z0-9_]*)\s+(.+?)\s*$")


def read_config_h(path: str | Path) -> dict:
    """Macro values of a config.h file.  #include "x.h" is followed relative to
    the file; #ifndef guards keep the first definition (a case header that
    defines overrides before including a defaults header works as cpp)."""
    path = Path(path)
    macros: dict = {}
    skip_depth = 0
    for line in path.read_text().splitlines():
        line = re.sub(r"/\*.*?\*/", "", line)
        s = line.strip()
        if s.startswith("#ifndef"):
            name = s.split()[1]
            skip_depth = skip_depth + 1 if (skip_depth or name in macros) else 0
            continue
        if s.startswith("#endif"):
            skip_depth = max(skip_depth - 1, 0)
            continue
        if skip_depth:
            continue
        inc = re.match(r'^#\s*include\s+"([^"]+)"', s)
        if inc:
            for k, v in read_config_h(path.parent / inc.group(1)).items():
                macros.setdefault(k, v)
            continue
        m = _DEFINE.match(s)
        if m and m.group(1) in MACROS:
            macros.setdefault(m.group(1), m.group(2))
    return macros


def _pow_int(t: float, n: int) -> float:
    r = 1.0
    for _ in range(n):
        r *= t
    return r


def kernel_w(kernel_type: int, r: float, h: float) -> float:
    """W(r, h) in 2-D, kernel_functions_mod.f90:26-81."""
    q = r / h
    if kernel_type == 1:
        if 0.0 <= q <= 1.0:
            w = _pow_int(2.0 - q, 3) - 4.0 * _pow_int(1.0 - q, 3)
        elif 1.0 < q <= 2.0:
            w = _pow_int(2.0 - q, 3)
        else:
            w = 0.0
        return 5.0 / (14.0 * PI * (h * h)) * w
    if kernel_type == 2:
        if 0.0 <= q <= 1.0:
            w = _pow_int(3.0 - q, 5) - 6.0 * _pow_int(2.0 - q, 5) + 15.0 * _pow_int(1.0 - q, 5)
        elif 1.0 < q <= 2.0:
            w = _pow_int(3.0 - q, 5) - 6.0 * _pow_int(2.0 - q, 5)
        elif 2.0 < q <= 3.0:
            w = _pow_int(3.0 - q, 5)
        else:
            w = 0.0
        return 7.0 / (478.0 * PI * (h * h)) * w
    if kernel_type == 3:
        w = _pow_int(1.0 - 0.5 * q, 4) * (1.0 + 2.0 * q) if 0.0 <= q <= 2.0 else 0.0
        return 7.0 / (4.0 * PI * (h * h)) * w
    if kernel_type == 4:
        w = (_pow_int(1.0 - 0.5 * q, 6) * (1.0 + 3.0 * q + 35.0 * (q * q) / 12.0)
             if 0.0 <= q <= 2.0 else 0.0)
        return 9.0 / (4.0 * PI * (h * h)) * w
    if kernel_type == 5:
        w = (_pow_int(1.0 - 0.5 * q, 8) * (1.0 + 4.0 * q + 25.0 * (q * q) / 4.0 + 4.0 * (q * q * q))
             if 0.0 <= q <= 2.0 else 0.0)
        return 39.0 / (14.0 * PI * (h * h)) * w
    raise ValueError(f"KERNEL_TYPE must be 1..5, got {kernel_type}")


class Params(ctypes.Structure):
    """lssph_params_t of include/lssph.h (field order is the ABI)."""

    _fields_ = [(n, ctypes.c_double) for n in (
        "len_x", "len_y", "dx", "sp_mass", "h", "h_eff", "w_ave", "wl_thick",
        "rho_ref", "vis_ref", "sound_speed", "thermal_dif",
        "alpha_ref", "gravity", "tem_ave",
        "tem_top", "tem_bottom", "tem_left", "tem_right", "u_top",
        "zeta_rsst", "xi_vim", "coe_delta_sph", "pst_c", "dt", "threshold",
        "diff_coef", "diff_source", "cond_max")] + [(n, ctypes.c_int32) for n in (
            "target_problem", "kernel_type", "sph_model", "wall_model", "rk",
            "u_bc_top", "u_bc_bottom", "u_bc_left", "u_bc_right",
            "tem_bc_top", "tem_bc_bottom", "tem_bc_left", "tem_bc_right",
            "num_x", "num_y", "num_wl", "periodic", "precision")]


@dataclasses.dataclass
class Param:
    """Param_type (global_types.f90:22-73): the config plus derived values."""

    cfg: Config
    total_step: int
    num_wl: int
    dx: float
    sp_mass: float
    h: float
    h_eff: float
    w_ave: float
    wl_thick: float
    delta_tem: float
    tem_ave: float
    sound_speed: float
    kinematic_vis: float
    thermal_dif: float
    re: float
    ra: float
    pr: float
    dt: float
    dt_cfl: float
    dt_vis: float
    dt_th: float
    dt_cfl_relax: float
    dt_vis_relax: float

    def __getattr__(self, name):
        # user-defined values read through (param%len_x, ...)
        if name != "cfg":
            return getattr(self.cfg, name)
        raise AttributeError(name)

    def to_c(self, precision: int | None = None) -> Params:
        c = self.cfg
        p = Params()
        for name in ("len_x", "len_y", "rho_ref", "vis_ref", "alpha_ref", "gravity",
                     "tem_top", "tem_bottom", "tem_left", "tem_right", "u_top",
                     "zeta_rsst", "xi_vim", "coe_delta_sph", "pst_c", "threshold",
                     "diff_coef", "diff_source", "cond_max"):
            setattr(p, name, getattr(c, name))
        for name in ("dx", "sp_mass", "h", "h_eff", "w_ave", "wl_thick", "sound_speed",
                     "thermal_dif", "tem_ave", "dt"):
            setattr(p, name, getattr(self, name))
        for name in ("target_problem", "kernel_type", "sph_model", "wall_model", "rk",
                     "u_bc_top", "u_bc_bottom", "u_bc_left", "u_bc_right",
                     "tem_bc_top", "tem_bc_bottom", "tem_bc_left", "tem_bc_right",
                     "num_x", "num_y", "periodic"):
            setattr(p, name, getattr(c, name))
        p.num_wl = self.num_wl
        p.precision = c.precision if precision is None else precision
        return p


def init_param(cfg: Config) -> Param:
    """Derived parameters exactly as init_param_mod.f90:92-199 computes them.

    Periodic boxes (extension) have no wall layers: num_wl = 0, wl_thick = 0.
    A positive cfg.dt_override replaces the derived time step."""
    check_param(cfg)
    total_step = cfg.end_step if cfg.read_name.strip() == "new" else cfg.end_step - cfg.start_step
    dx = cfg.len_x / float(cfg.num_x)
    sp_mass = cfg.rho_ref * (dx * dx)
    h = cfg.coe_h * dx
    h_eff = 3.0 * h if cfg.kernel_type == 2 else 2.0 * h
    w_ave = kernel_w(cfg.kernel_type, dx, h)
    num_wl = 0 if cfg.periodic else int(math.ceil(h_eff / dx))
    wl_thick = float(num_wl) * dx
    if cfg.target_problem == 4:
        delta_tem = cfg.tem_bottom - cfg.tem_top
        tem_ave = 0.5 * (cfg.tem_bottom + cfg.tem_top)
    else:
        delta_tem = 0.0
        tem_ave = 0.0
    sound_speed = math.sqrt(cfg.k_ref / cfg.rho_ref)
    kinematic_vis = 1.0 if cfg.target_problem == 1 else cfg.vis_ref / cfg.rho_ref
    thermal_dif = cfg.k_th_ref / (cfg.rho_ref * cfg.cp_ref)
    if cfg.target_problem == 2:
        re = 1.0 * cfg.len_x / kinematic_vis
    else:
        re = cfg.u_top * cfg.len_x / kinematic_vis
    ra = cfg.alpha_ref * delta_tem * cfg.gravity * (cfg.len_y * cfg.len_y * cfg.len_y) / (
        kinematic_vis * thermal_dif)
    pr = kinematic_vis / thermal_dif
    dt_cfl = cfg.coe_cfl * (dx / sound_speed)
    dt_vis = cfg.coe_dif * ((dx * dx) / kinematic_vis)
    dt_th = cfg.coe_dif * ((dx * dx) / thermal_dif)
    dt_cfl_relax = dt_cfl * cfg.zeta_rsst * cfg.xi_vim
    dt_vis_relax = dt_vis * (cfg.xi_vim * cfg.xi_vim)
    if cfg.target_problem == 1:
        dt = dt_vis
    elif cfg.target_problem == 4:
        dt = min(dt_cfl_relax, dt_vis_relax, dt_th)
    else:
        dt = min(dt_cfl_relax, dt_vis_relax)
    if cfg.dt_override > 0.0:
        dt = cfg.dt_override
    return Param(cfg=cfg, total_step=total_step, num_wl=num_wl, dx=dx, sp_mass=sp_mass,
                 h=h, h_eff=h_eff, w_ave=w_ave, wl_thick=wl_thick, delta_tem=delta_tem,
                 tem_ave=tem_ave, sound_speed=sound_speed, kinematic_vis=kinematic_vis,
                 thermal_dif=thermal_dif, re=re, ra=ra, pr=pr, dt=dt, dt_cfl=dt_cfl,
                 dt_vis=dt_vis, dt_th=dt_th, dt_cfl_relax=dt_cfl_relax,
                 dt_vis_relax=dt_vis_relax)


class ConfigError(ValueError):
    """Raised where the reference's check_param stops with `error stop`."""


def check_param(cfg: Config) -> None:
    """Input validation of check_param_mod.f90:22-284 (same conditions)."""
    def fail(msg):
        raise ConfigError(msg + " Please check 'config.h'.")

    if not 1 <= cfg.target_problem <= 4:
        fail("TARGET_PROBLEM must be between 1 and 4.")
    if not 1 <= cfg.kernel_type <= 5:
        fail("KERNEL_TYPE must be between 1 and 5.")
    if not 1 <= cfg.sph_model <= 6:
        fail("SPH_MODEL must be between 1 and 6.")
    if not 1 <= cfg.wall_model <= 3:
        fail("WALL_MODEL must be between 1 and 3.")
    ub = (cfg.u_bc_top, cfg.u_bc_bottom, cfg.u_bc_left, cfg.u_bc_right)
    if not cfg.periodic:
        if cfg.target_problem == 1 and any(b not in (1, 2) for b in ub):
            fail("Diffusion Equation Test: U_BOUNDARY_* must be 1 or 2.")
        if cfg.target_problem == 2 and any(b != 1 for b in ub):
            fail("Taylor-Green Vortex: U_BOUNDARY_* must be 1.")
        if cfg.target_problem == 3 and ub != (3, 2, 2, 2):
            fail("Lid-driven Cavity Flow: U_BOUNDARY_TOP must be 3, all other U_BOUNDARY_* 2.")
        if cfg.target_problem == 4 and (any(b not in (1, 2) for b in ub) or (
                cfg.tem_bc_top, cfg.tem_bc_bottom, cfg.tem_bc_left, cfg.tem_bc_right) != (1, 1, 2, 2)):
            fail("Boussinesq Convection: U_BOUNDARY_* must be 1 or 2, TEM_BOUNDARY_TOP/BOTTOM 1, LEFT/RIGHT 2.")
    if cfg.target_problem == 1 and not cfg.periodic and (cfg.len_x != 1.0 or cfg.len_y != 1.0):
        fail("For Diffusion Equation Test, LEN_X and LEN_Y must be exactly 1.0.")
    if cfg.target_problem == 2 and (cfg.tg_a <= 0 or cfg.tg_b <= 0):
        fail("TG_A and TG_B must be positive.")
    if cfg.len_x <= 0.0 or cfg.len_y <= 0.0:
        fail("LEN_X and LEN_Y must be positive.")
    if cfg.num_x <= 0 or cfg.num_y <= 0:
        fail("NUM_X and NUM_Y must be positive.")
    if cfg.start_step <= 0 or cfg.end_step <= 0 or cfg.write_step <= 0:
        fail("START_STEP, END_STEP and WRITE_STEP must be positive.")
    if cfg.start_step >= cfg.end_step:
        fail("START_STEP must be less than END_STEP.")
    if not (0.0 < cfg.coe_cfl < 1.0 and 0.0 < cfg.coe_dif < 1.0):
        fail("COE_CFL and COE_DIF must be between 0 and 1.")
    if cfg.zeta_rsst < 1.0 or cfg.xi_vim < 1.0:
        fail("ZETA_RSST and XI_VIM must be 1.0 or greater.")
    if min(cfg.rho_ref, cfg.vis_ref, cfg.k_ref, cfg.k_th_ref, cfg.cp_ref, cfg.alpha_ref) <= 0.0:
        fail("Physical properties must be positive.")
    if cfg.gravity < 0.0:
        fail("GRAVITY must be non-negative.")
    if cfg.target_problem == 1 and not 0.0 <= cfg.pos_pert < 0.5:
        fail("POS_PERT must be between 0.0 and 0.5")
    if cfg.coe_h <= 0.0:
        fail("COE_H must be positive.")
    if cfg.pst_c < 0.0 or cfg.coe_delta_sph < 0.0:
        fail("PST_C and COE_DELTA_SPH must be non-negative.")
    if abs(cfg.len_x / float(cfg.num_x) - cfg.len_y / float(cfg.num_y)) > 1.0e-8:
        fail("Particle spacing must be isotropic (dx == dy).")
    if cfg.precision not in (0, 1):
        fail("PRECISION must be 0 (FP64) or 1 (mixed).")
