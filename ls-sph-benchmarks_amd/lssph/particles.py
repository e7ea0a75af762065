"""Host-side particle setup: lattice, particle types, initial conditions.

Restates the reference's one-time host setup so that the GPU path receives
the same particle state the Fortran reference builds:
  - setup_closed_box (source/setup/setup_closed_box_mod.f90:23-237): a
    lattice offset by dx/2 that covers the fluid box plus num_wl wall layers,
    built by repeated addition, fluid particles first (row-major, x fastest),
    then the ghost-wall particles with wall types 1..8;
  - the initial conditions of impose_initial_conditions_mod.f90:24-145.
Extensions (no reference counterpart): a doubly periodic box without walls,
seeded SplitMix64 jitter, and the periodic heat-equation initial field.
"""
from __future__ import annotations

import dataclasses

import numpy as np

from .config import PI, Param


@dataclasses.dataclass
class State:
    """SP_type arrays (global_types.f90:78-85) in the reference order."""

    ptype: np.ndarray
    x: np.ndarray
    y: np.ndarray
    u: np.ndarray
    v: np.ndarray
    rho: np.ndarray
    pre: np.ndarray
    tem: np.ndarray | None
    num_int: int
    num_ext: int

    @property
    def num_total(self) -> int:
        return self.num_int + self.num_ext

    def copy(self) -> "State":
        return State(ptype=self.ptype.copy(), x=self.x.copy(), y=self.y.copy(),
                     u=self.u.copy(), v=self.v.copy(), rho=self.rho.copy(),
                     pre=self.pre.copy(), tem=None if self.tem is None else self.tem.copy(),
                     num_int=self.num_int, num_ext=self.num_ext)


def _lattice(n_x: int, n_y: int, dx: float):
    """x_sys/y_sys of setup_closed_box_mod.f90:72-82: the first row by
    repeated addition of dx from dx/2, every later row by adding dx to the
    row below."""
    xs = np.empty(n_x)
    xs[0] = dx / 2.0
    for i in range(1, n_x):
        xs[i] = xs[i - 1] + dx
    ys = np.empty(n_y)
    ys[0] = dx / 2.0
    for j in range(1, n_y):
        ys[j] = ys[j - 1] + dx
    x = np.tile(xs, n_y)
    y = np.repeat(ys, n_x)
    return x, y


def setup_closed_box(param: Param) -> State:
    """Positions and particle types (setup_closed_box_mod.f90:23-237).
    Fields are zero except wall density = rho_ref (:209)."""
    c = param.cfg
    nxf = c.num_x + 2 * param.num_wl
    nyf = c.num_y + 2 * param.num_wl
    xs, ys = _lattice(nxf, nyf, param.dx)
    x_min = 0.0 + param.wl_thick
    x_max = c.len_x + param.wl_thick
    y_min = 0.0 + param.wl_thick
    y_max = c.len_y + param.wl_thick
    fluid = (x_min < xs) & (x_max > xs) & (y_min < ys) & (y_max > ys)
    xi, yi = xs[fluid], ys[fluid]
    xe, ye = xs[~fluid], ys[~fluid]
    n_int, n_ext = xi.size, xe.size
    pt_e = np.zeros(n_ext, np.int32)
    below, above = ye < y_min, ye > y_max
    left, right = xe < x_min, xe > x_max
    pt_e[below] = 1
    pt_e[below & left] = 5
    pt_e[below & right] = 6
    pt_e[above] = 2
    pt_e[above & left] = 7
    pt_e[above & right] = 8
    mid = ~below & ~above
    pt_e[mid & left] = 3
    pt_e[mid & right] = 4
    if np.any(pt_e == 0):
        raise RuntimeError("Fatal Error in setup: unclassified wall particle")
    n = n_int + n_ext
    st = State(ptype=np.concatenate([np.zeros(n_int, np.int32), pt_e]),
               x=np.concatenate([xi, xe]), y=np.concatenate([yi, ye]),
               u=np.zeros(n), v=np.zeros(n), rho=np.zeros(n), pre=np.zeros(n),
               tem=np.zeros(n) if c.target_problem == 4 else None,
               num_int=n_int, num_ext=n_ext)
    st.rho[n_int:] = c.rho_ref
    return st


def setup_periodic_box(param: Param) -> State:
    """Extension: num_x x num_y fluid particles on the dx/2-offset lattice of
    a doubly periodic box, no walls."""
    c = param.cfg
    x, y = _lattice(c.num_x, c.num_y, param.dx)
    n = x.size
    return State(ptype=np.zeros(n, np.int32), x=x, y=y, u=np.zeros(n), v=np.zeros(n),
                 rho=np.full(n, c.rho_ref), pre=np.zeros(n),
                 tem=np.zeros(n) if c.target_problem == 4 else None, num_int=n, num_ext=0)


def splitmix64_uniform(seed: int, count: int) -> np.ndarray:
    """count uniforms in [0, 1) from the SplitMix64 stream of `seed`
    (Steele, Lea & Flood 2014): z_k = mix(seed + (k+1) * 0x9E3779B97F4A7C15)."""
    with np.errstate(over="ignore"):
        k = np.arange(1, count + 1, dtype=np.uint64)
        z = np.uint64(seed) + k * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def jitter(st: State, param: Param, amplitude: float, seed: int) -> None:
    """Uniform jitter of +-amplitude*dx per coordinate of every fluid
    particle from the SplitMix64 stream of `seed` (x from even draws, y from
    odd draws).  Periodic boxes wrap back into [0, L)."""
    r = splitmix64_uniform(seed, 2 * st.num_int)
    n = st.num_int
    st.x[:n] += amplitude * param.dx * (2.0 * r[0::2] - 1.0)
    st.y[:n] += amplitude * param.dx * (2.0 * r[1::2] - 1.0)
    if param.cfg.periodic:
        st.x[:n] = np.mod(st.x[:n], param.cfg.len_x)
        st.y[:n] = np.mod(st.y[:n], param.cfg.len_y)


def impose_tg_vortex(st: State, param: Param) -> None:
    """Taylor-Green initial state, impose_initial_conditions_mod.f90:61-96."""
    c = param.cfg
    a, b = float(c.tg_a), float(c.tg_b)
    lx, ly = c.len_x, c.len_y
    ratio = (a * ly) / (b * lx)
    n = st.num_int
    x = st.x[:n] - param.wl_thick
    y = st.y[:n] - param.wl_thick
    st.u[:n] = +1.0 * np.sin(a * PI * x / lx) * np.cos(b * PI * y / ly)
    st.v[:n] = -(1.0 * ratio * np.cos(a * PI * x / lx) * np.sin(b * PI * y / ly))
    st.pre[:n] = 0.25 * c.rho_ref * 1.0 * (np.cos(2.0 * a * PI * x / lx)
                                          + (ratio * ratio) * np.cos(2.0 * b * PI * y / ly))
    st.rho[:n] = c.rho_ref + st.pre[:n] / (param.sound_speed * param.sound_speed)


def impose_cavity(st: State, param: Param) -> None:
    """Fluid at rest, impose_initial_conditions_mod.f90:102-113."""
    st.rho[:] = param.cfg.rho_ref
    st.pre[:] = 0.0
    st.u[:] = 0.0
    st.v[:] = 0.0


def impose_boussinesq(st: State, param: Param) -> None:
    """Linear temperature profile plus a bottom perturbation,
    impose_initial_conditions_mod.f90:119-143."""
    c = param.cfg
    impose_cavity(st, param)
    n = st.num_int
    x = st.x[:n] - param.wl_thick
    y = st.y[:n] - param.wl_thick
    t = -(param.delta_tem / c.len_y) * y + c.tem_bottom
    low = y <= 0.1 * c.len_y
    t[low] = t[low] + 1.0e-2 * param.delta_tem * np.cos(PI * x[low] / c.len_x)
    st.tem[:n] = t


def impose_diffusion(st: State, param: Param) -> None:
    """Diffusion test start (impose_initial_conditions_mod.f90:24-55) without
    its clock-seeded jitter; use jitter() for a reproducible perturbation."""
    st.rho[:] = param.cfg.rho_ref
    st.u[:] = 0.0
    st.pre[:] = 0.0
    st.v[:] = 0.0


def impose_heat_mode(st: State, param: Param) -> None:
    """Extension (periodic heat equation): f = sin(2 pi x) sin(2 pi y) on a
    unit box, which decays as exp(-8 pi^2 kappa t)."""
    impose_diffusion(st, param)
    n = st.num_int
    st.u[:n] = np.sin(2.0 * PI * st.x[:n] / param.cfg.len_x) * np.sin(2.0 * PI * st.y[:n] / param.cfg.len_y)


def init_particle(param: Param, jitter_amp: float = 0.0, seed: int = 0,
                  heat_mode: bool = False) -> State:
    """init_particle (init_particle_mod.f90:46-86) for a new run: geometry,
    optional seeded jitter (applied before the initial field so the field is
    evaluated at the jittered positions), then the problem's initial state."""
    c = param.cfg
    st = setup_periodic_box(param) if c.periodic else setup_closed_box(param)
    if jitter_amp > 0.0:
        jitter(st, param, jitter_amp, seed)
    if c.target_problem == 1:
        (impose_heat_mode if heat_mode else impose_diffusion)(st, param)
    elif c.target_problem == 2:
        impose_tg_vortex(st, param)
    elif c.target_problem == 3:
        impose_cavity(st, param)
    else:
        impose_boussinesq(st, param)
    return st
