"""Quantities the reference's analysis scripts report, computed on the device
state of a handle.

nusselt: mean Nusselt numbers of the bottom and top walls of a convection
box, as analysis/benchmarks/boussinesq_conv.py:80-108 computes them from the
last output file: dT/dy by a type-B least-squares fit (lssph_probe, the fit
of analysis/common/lssph_b.py:21-144 over fluid and wall particles within
h_eff) at num_x points x = linspace(0, L_x, num_x) on each wall, local
Nu = (dT/dy) / (-delta_T / L_y), averaged per wall.
"""
from __future__ import annotations

import numpy as np


def probe_order(target_problem: int, sph_model: int) -> int:
    """Basis size of the analysis fit (lssph_b.py:43-53); the library's
    probe supports q = 3, 6 and 10 (SPH_MODEL 1-4)."""
    table = {3: (6, 3), 4: (10, 6)}
    m = max(sph_model, 3)
    if m not in table:
        raise NotImplementedError(f"SPH_MODEL {sph_model}: the analysis fit has more than 10 terms")
    return table[m][0] if target_problem == 4 else table[m][1]


def nusselt(sim, param, inst: int = 0):
    """(Nu_bottom, Nu_top, (local_bottom, local_top)) of instance inst."""
    c = param.cfg
    if c.target_problem != 4:
        raise ValueError("the Nusselt number needs the temperature field (TARGET_PROBLEM 4)")
    q = probe_order(c.target_problem, c.sph_model)
    wl = param.wl_thick
    x = np.linspace(0.0, c.len_x, c.num_x) + wl
    grad_ref = -param.delta_tem / c.len_y
    local = []
    for y_wall in (0.0, c.len_y):
        d = sim.probe(x, np.full(c.num_x, y_wall + wl), field="tem", q=q, inst=inst)
        local.append(d[:, 2] / grad_ref)
    return float(np.mean(local[0])), float(np.mean(local[1])), tuple(local)


def cavity_profiles(sim, param, inst: int = 0):
    """Centerline velocity profiles of a lid-driven cavity, as
    analysis/benchmarks/cavity_flow.py:88-106 takes them from the last output
    file: u at x = L_x / 2 over y = linspace(0, L_y, num_y) and v at
    y = L_y / 2 over x = linspace(0, L_x, num_x), each the constant term of
    the type-B fit (lssph_probe) over fluid and wall particles within h_eff.
    Returns ((y, u), (x, v)) in physical coordinates."""
    c = param.cfg
    q = probe_order(c.target_problem, c.sph_model)
    wl = param.wl_thick
    y = np.linspace(0.0, c.len_y, c.num_y)
    x = np.linspace(0.0, c.len_x, c.num_x)
    u = sim.probe(np.full(c.num_y, 0.5 * c.len_x + wl), y + wl, field="u", q=q, inst=inst)[:, 0]
    v = sim.probe(x + wl, np.full(c.num_x, 0.5 * c.len_y + wl), field="v", q=q, inst=inst)[:, 0]
    return (y, u), (x, v)


def vrms(state: dict, num_int: int) -> float:
    """sqrt(mean(u^2 + v^2)) over the fluid particles (cavity_flow.py:74-84,
    boussinesq_conv.py:64-68)."""
    u = state["u"][:num_int]
    v = state["v"][:num_int]
    return float(np.sqrt(np.mean(u * u + v * v)))
