"""Restatement of the reference analysis' wall-gradient fit and Nusselt number.

TEST INFRASTRUCTURE ONLY: the checker of the library's probe fit
(lssph_probe) and of lssph.analysis.nusselt; only tests/ import it.

  probe_b   analysis/common/lssph_b.py:21-144: type-B least-squares fit of a
            field at target points over every particle (fluid and wall)
            within h_eff (KD-tree ball query, r <= h_eff), weights
            W(r, h) m / rho_j, basis [1, x/h, y/h, x^2/2h^2, xy/h^2,
            y^2/2h^2, ...] of the offsets x_j - x_target; returns
            (d_0, d_1 / h, d_2 / h), NaN below q neighbors or for a
            singular system.
  nusselt   analysis/benchmarks/boussinesq_conv.py:80-108: dT/dy by probe_b
            at num_x points x = linspace(0, L_x, num_x) on the bottom and
            top walls (physical coordinates, the wall thickness subtracted),
            local Nu = (dT/dy) / (-delta_T / L_y), averaged per wall.
"""
from __future__ import annotations

import numpy as np
from scipy.spatial import KDTree

from .pyoracle import kernel_w


def _q_of(target_problem: int, sph_model: int) -> int:
    """lssph_b.py:43-53."""
    table = {3: (6, 3), 4: (10, 6), 5: (15, 10), 6: (21, 15)}
    m = max(sph_model, 3)
    return table[m][0] if target_problem == 4 else table[m][1]


def _basis(dx_h: np.ndarray, dy_h: np.ndarray, q: int) -> np.ndarray:
    cols = [np.ones_like(dx_h), dx_h, dy_h]
    if q >= 6:
        cols += [0.5 * dx_h ** 2, dx_h * dy_h, 0.5 * dy_h ** 2]
    if q >= 10:
        cols += [(1.0 / 6.0) * dx_h ** 3, 0.5 * dx_h ** 2 * dy_h, 0.5 * dx_h * dy_h ** 2, (1.0 / 6.0) * dy_h ** 3]
    if q > 10:
        raise NotImplementedError("q = 15 and 21 are not restated")
    return np.stack(cols, axis=1)


def probe_b(h: float, h_eff: float, sp_mass: float, kernel_type: int, q: int, tx, ty, xj, yj, fj, rhoj):
    """lssph_b.py:21-144 for one field; (n, 3) array."""
    tx = np.atleast_1d(np.asarray(tx, np.float64))
    ty = np.atleast_1d(np.asarray(ty, np.float64))
    tree = KDTree(np.column_stack((xj, yj)))
    idx = tree.query_ball_point(np.column_stack((tx, ty)), r=h_eff)
    out = np.full((tx.size, 3), np.nan)
    wfun = np.vectorize(lambda r: kernel_w(kernel_type, float(r), h))
    for i, nb in enumerate(idx):
        if len(nb) < q:
            continue
        nb = np.asarray(nb)
        dx = xj[nb] - tx[i]
        dy = yj[nb] - ty[i]
        r = np.sqrt(dx * dx + dy * dy)
        a = _basis(dx / h, dy / h, q)
        wv = wfun(r) * (sp_mass / rhoj[nb])
        m = a.T @ (wv[:, None] * a)
        b = a.T @ (wv * fj[nb])
        try:
            d = np.linalg.solve(m, b)
        except np.linalg.LinAlgError:
            continue
        out[i, 0] = d[0]
        out[i, 1:] = d[1:3] / h
    return out


def nusselt(param, st: dict, num_total: int):
    """boussinesq_conv.py:80-108 on a state dict (upload-order arrays of one
    box): (Nu_bottom, Nu_top) averages and the local values."""
    c = param.cfg
    wl = param.wl_thick
    x = st["x"][:num_total] - wl
    y = st["y"][:num_total] - wl
    q = _q_of(c.target_problem, c.sph_model)
    tx = np.linspace(0.0, c.len_x, c.num_x)
    grad_ref = -param.delta_tem / c.len_y
    loc = []
    for ty in (np.zeros(c.num_x), np.full(c.num_x, c.len_y)):
        d = probe_b(param.h, param.h_eff, param.sp_mass, c.kernel_type, q, tx, ty, x, y,
                    st["tem"][:num_total], st["rho"][:num_total])
        loc.append(d[:, 2] / grad_ref)
    return float(np.mean(loc[0])), float(np.mean(loc[1])), loc
