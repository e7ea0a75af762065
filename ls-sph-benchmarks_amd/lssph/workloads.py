"""Benchmark and test workloads as concrete configurations.

Each builder returns (param, state) ready for Simulation.upload():
  taylor_green_box   reference Taylor-Green vortex (TARGET_PROBLEM 2) in the
                     closed box with free-slip ghost walls
  taylor_green_periodic  the same flow on a doubly periodic box (extension)
  heat_periodic      periodic heat equation f = sin 2pi x sin 2pi y (extension
                     of the diffusion test, TARGET_PROBLEM 1, no source)
  cavity             lid-driven cavity (TARGET_PROBLEM 3)
  boussinesq         bottom-heated convection (TARGET_PROBLEM 4)
  cavity_ensemble    many cavities of uneven sizes for one handle
Jitter comes from the SplitMix64 stream of the given seed.
"""
from __future__ import annotations

import math

import numpy as np

from .config import PI, Config, init_param
from .particles import init_particle


def taylor_green_box(n: int = 64, re: float = 100.0, rk: int = 2, coe_h: float = 1.55,
                     coe_cfl: float = 0.64, jitter: float = 0.0, seed: int = 2026, precision: int = 0,
                     kernel_type: int = 3, tg: int = 2):
    cfg = Config(target_problem=2, num_x=n, num_y=n, rk=rk, coe_h=coe_h, coe_cfl=coe_cfl,
                 vis_ref=1.0 / re, k_ref=100.0, tg_a=tg, tg_b=tg, kernel_type=kernel_type,
                 wall_model=2, precision=precision)
    param = init_param(cfg)
    return param, init_particle(param, jitter, seed)


def taylor_green_periodic(n: int = 2048, re: float = 1000.0, rk: int = 4, coe_h: float = 1.55,
                          coe_cfl: float = 0.5, jitter: float = 0.2, seed: int = 2029, precision: int = 0,
                          kernel_type: int = 3, tg: int = 2):
    cfg = Config(target_problem=2, num_x=n, num_y=n, rk=rk, coe_h=coe_h, coe_cfl=coe_cfl,
                 vis_ref=1.0 / re, k_ref=100.0, tg_a=tg, tg_b=tg, kernel_type=kernel_type,
                 periodic=1, precision=precision)
    param = init_param(cfg)
    return param, init_particle(param, jitter, seed)


def heat_periodic(n: int = 512, kappa: float = 0.01, rk: int = 4, coe_h: float = 1.55, coe_dif: float = 0.1,
                  jitter: float = 0.1, seed: int = 2028, precision: int = 0):
    """dt = coe_dif * dx^2 / kappa; RHS = kappa * Laplacian (no source)."""
    dx = 1.0 / n
    cfg = Config(target_problem=1, num_x=n, num_y=n, rk=rk, coe_h=coe_h, coe_dif=coe_dif,
                 periodic=1, diff_coef=kappa, diff_source=0.0, precision=precision,
                 dt_override=coe_dif * (dx * dx) / kappa)
    param = init_param(cfg)
    return param, init_particle(param, jitter, seed, heat_mode=True)


def cavity(n: int = 64, re: float = 100.0, rk: int = 2, coe_h: float = 1.2, precision: int = 0,
           jitter: float = 0.0, seed: int = 0):
    cfg = Config(target_problem=3, num_x=n, num_y=n, rk=rk, coe_h=coe_h, vis_ref=1.0 / re, u_top=1.0,
                 u_bc_top=3, u_bc_bottom=2, u_bc_left=2, u_bc_right=2, wall_model=2, precision=precision)
    param = init_param(cfg)
    return param, init_particle(param, jitter, seed)


def boussinesq(n: int = 128, ra: float = 1.0e4, pr: float = 0.71, rk: int = 4, coe_h: float = 1.55,
               wall_model: int = 3, jitter: float = 0.2, seed: int = 2030, precision: int = 0,
               k_ref: float = 100.0):
    """nu and kappa from Ra = alpha dT g L^3 / (nu kappa) and Pr = nu / kappa
    (init_param_mod.f90:162-167) with alpha = 7.1e-3, g = 10, dT = 1."""
    alpha, g, dtem = 7.1e-3, 10.0, 1.0
    kappa = math.sqrt(alpha * dtem * g / (ra * pr))
    nu = pr * kappa
    cfg = Config(target_problem=4, num_x=n, num_y=n, rk=rk, coe_h=coe_h, vis_ref=nu, k_th_ref=kappa,
                 cp_ref=1.0, rho_ref=1.0, alpha_ref=alpha, gravity=g, tem_top=0.0, tem_bottom=1.0,
                 u_bc_top=2, u_bc_bottom=2, u_bc_left=2, u_bc_right=2, tem_bc_top=1, tem_bc_bottom=1,
                 tem_bc_left=2, tem_bc_right=2, wall_model=wall_model, k_ref=k_ref, precision=precision)
    param = init_param(cfg)
    return param, init_particle(param, jitter, seed)


def cavity_ensemble(count: int = 64, sizes=(32, 64, 128, 256), re_range=(100.0, 1000.0), rk: int = 2,
                    coe_h: float = 1.2, seed: int = 2030, precision: int = 1):
    """count cavities; resolution drawn uniformly from sizes and Re
    log-uniformly from re_range, both from the SplitMix64 stream of seed."""
    from .particles import splitmix64_uniform
    r = splitmix64_uniform(seed, 2 * count)
    params, states, meta = [], [], []
    for k in range(count):
        n = sizes[min(int(r[2 * k] * len(sizes)), len(sizes) - 1)]
        re = 10.0 ** (math.log10(re_range[0]) + r[2 * k + 1] * (math.log10(re_range[1]) - math.log10(re_range[0])))
        p, st = cavity(n, re, rk=rk, coe_h=coe_h, precision=precision)
        params.append(p)
        states.append(st)
        meta.append({"n": n, "re": re})
    return params, states, meta


def tg_exact(param, x, y, t):
    """Analytic Taylor-Green velocity at time t (taylor_green.py:35,113-123):
    amplitude exp(-sigma t), sigma = nu ((a pi / Lx)^2 + (b pi / Ly)^2)."""
    c = param.cfg
    a, b = float(c.tg_a), float(c.tg_b)
    lx, ly = c.len_x, c.len_y
    sigma = param.kinematic_vis * ((a * PI / lx) ** 2 + (b * PI / ly) ** 2)
    amp = math.exp(-sigma * t)
    ratio = (a * ly) / (b * lx)
    xx = x - param.wl_thick
    yy = y - param.wl_thick
    u = amp * np.sin(a * PI * xx / lx) * np.cos(b * PI * yy / ly)
    v = -amp * ratio * np.cos(a * PI * xx / lx) * np.sin(b * PI * yy / ly)
    return u, v


def tg_l2_error(param, st: dict, t: float, n_int: int) -> float:
    """L2 = sqrt(mean((u-u_ex)^2 + (v-v_ex)^2)) over the fluid particles."""
    u_ex, v_ex = tg_exact(param, st["x"][:n_int], st["y"][:n_int], t)
    return float(np.sqrt(np.mean((st["u"][:n_int] - u_ex) ** 2 + (st["v"][:n_int] - v_ex) ** 2)))
