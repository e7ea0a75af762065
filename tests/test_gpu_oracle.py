"""HIP library against the oracle on cases the reference cannot run
(periodic boxes, ensembles) and on the mixed-precision path (MI355X)."""
import numpy as np
import pytest

from helpers import rel_err
from lssph import workloads
from lssph.engine import Simulation, simulation_from_state, split_states
from oracle.pyoracle import OracleSim

pytestmark = pytest.mark.gpu
TOL = 1e-12


def _pair(param, st, precision=0):
    print(f"[gpu] n_int={st.num_int} n_ext={st.num_ext} periodic={param.cfg.periodic} "
          f"target={param.cfg.target_problem} precision={precision}", flush=True)
    o = OracleSim(param.to_c(), st.num_int, st.num_ext)
    o.init(st)
    s = simulation_from_state(param, st, precision=precision)
    return o, s


def _cmp(a, b, fields, n=None):
    return {k: rel_err(a[k][:n], b[k][:n]) for k in fields}


@pytest.mark.parametrize("rk", [2, 4])
def test_periodic_taylor_green_matches_oracle(rk):
    param, st = workloads.taylor_green_periodic(n=40, re=100.0, rk=rk, jitter=0.2, seed=7)
    o, s = _pair(param, st)
    with s:
        co, lo = o.neighbors()
        cs, ls = s.neighbors()
        assert np.array_equal(co, cs) and np.array_equal(lo, ls)
        r_o, r_s = o.calculate_rhs(), s.calculate_RHS()
        assert max(_cmp(r_s, r_o, ("u", "v", "rho")).values()) <= TOL
        for _ in range(3):
            o.step(1)
            s.step(1)
        e = _cmp(s.download(), o.get_state(), ("x", "y", "u", "v", "rho", "pre"))
        assert max(e.values()) <= TOL, e


@pytest.mark.parametrize("n", [10, 11])
def test_smallest_periodic_box_matches_oracle(n):
    """3 cells per direction (L / h_eff = 3.2 and 3.5): every stencil wraps
    and every particle sees images across both boundaries."""
    param, st = workloads.taylor_green_periodic(n=n, re=100.0, rk=4, jitter=0.2, seed=13)
    o, s = _pair(param, st)
    with s:
        co, lo = o.neighbors()
        cs, ls = s.neighbors()
        assert np.array_equal(co, cs) and np.array_equal(lo, ls)
        r_o, r_s = o.calculate_rhs(), s.calculate_RHS()
        assert max(_cmp(r_s, r_o, ("u", "v", "rho")).values()) <= TOL
        for _ in range(5):
            o.step(1)
            s.step(1)
        e = _cmp(s.download(), o.get_state(), ("x", "y", "u", "v", "rho", "pre"))
        assert max(e.values()) <= TOL, e


def test_periodic_heat_matches_oracle():
    param, st = workloads.heat_periodic(n=48, kappa=0.01, jitter=0.1, seed=9)
    o, s = _pair(param, st)
    with s:
        for _ in range(5):
            o.step(1)
            s.step(1)
        e = rel_err(s.download()["u"], o.get_state()["u"])
        assert e <= TOL, e


def test_taylor_green_64_full_run_matches_oracle():
    """The plain-method configuration at full length: 64^2 regular lattice,
    Re 100, RK2 with shifting, dt = 1e-3 (COE_CFL 0.64, K_REF 100), 100
    steps, FP64.  The GPU trajectory matches the oracle's, and so does the
    Taylor-Green L2 velocity error against the analytic decay."""
    param, st = workloads.taylor_green_periodic(n=64, re=100.0, rk=2, coe_cfl=0.64, jitter=0.0, seed=0)
    assert abs(param.dt - 1e-3) <= 1e-15, param.dt
    o, s = _pair(param, st)
    with s:
        s.step(100)
        got = s.download()
    o.step(100)
    ref = o.get_state()
    e = _cmp(got, ref, ("x", "y", "u", "v", "rho", "pre"))
    print(f"[gpu] TG 64^2, 100 steps: {e}", flush=True)
    assert max(e.values()) <= 1e-10, e
    t = 100 * param.dt
    l2_g = workloads.tg_l2_error(param, got, t, st.num_int)
    l2_o = workloads.tg_l2_error(param, ref, t, st.num_int)
    assert abs(l2_g - l2_o) <= 1e-9 * l2_o, (l2_g, l2_o)


def test_boussinesq_1024_mixed_close_to_fp64():
    """The convection configuration at full size (1024^2, jitter 0.2 dx,
    Ra 1e6, WALL_MODEL 3, RK4 + PST): after 20 steps the mixed path's
    fields are within 1e-6 (norm-wise relative) of the FP64 path's, and no
    particle needed the FP64 re-solve."""
    out, stats = [], []
    for prec in (0, 1):
        param, st = workloads.boussinesq(n=1024, ra=1e6, jitter=0.2, seed=2030, precision=prec)
        print(f"[gpu] Boussinesq 1024^2 precision={prec}", flush=True)
        with simulation_from_state(param, st, precision=prec) as s:
            s.ghost()
            s.step(20)
            out.append(s.download())
            stats.append(s.stats())
    n = st.num_int
    e = _cmp(out[1], out[0], ("x", "y", "u", "v", "rho", "pre", "tem"), n)
    print(f"[gpu] mixed vs fp64: {e}", flush=True)
    assert max(e.values()) <= 1e-6, e
    assert stats[1].fp64_fallbacks == 0 and stats[1].list_overflows == 0


def test_periodic_heat_512_decays_at_the_analytic_rate():
    """The diffusion configuration at full size (512^2, jitter 0.1 dx, kappa
    0.01, RK4, dt = 0.1 dx^2 / kappa, 200 steps, FP64) against closed forms:
    the first RHS is kappa times the Laplacian of sin 2pi x sin 2pi y,
    -8 pi^2 kappa f, and the field stays that mode with amplitude
    exp(-8 pi^2 kappa t).  Measured: Laplacian within 2.1e-4, decay rate
    within 6.3e-5, off-mode residual 3.5e-8; the bounds below leave 5-30x
    room, while a wrong factor or sign is off by order one."""
    import math
    param, st = workloads.heat_periodic(n=512, kappa=0.01, rk=4, jitter=0.1, seed=2028)
    n = st.num_int
    f0 = st.u[:n].copy()
    lam = 8.0 * math.pi ** 2 * 0.01
    print(f"[gpu] periodic heat 512^2, 200 steps", flush=True)
    with simulation_from_state(param, st, precision=0) as s:
        r0 = s.calculate_RHS()["u"][:n]
        s.step(200)
        u = s.download()["u"][:n]
    assert np.linalg.norm(r0 + lam * f0) / np.linalg.norm(lam * f0) <= 1e-3
    amp = float(np.dot(u, f0) / np.dot(f0, f0))
    g = math.exp(-lam * 200 * param.dt)
    assert abs(math.log(amp) / math.log(g) - 1.0) <= 1e-3, (amp, g)
    assert np.linalg.norm(u - amp * f0) / np.linalg.norm(f0) <= 1e-6


def test_reorder_does_not_change_results():
    param, st = workloads.taylor_green_periodic(n=32, re=100.0, rk=4, jitter=0.2, seed=11)
    with simulation_from_state(param, st) as a, simulation_from_state(param, st) as b:
        a.set_tuning(reorder_interval=1)
        b.set_tuning(reorder_interval=1000)
        a.step(6)
        b.step(6)
        da, db = a.download(), b.download()
        for k in ("x", "y", "u", "v", "rho", "pre"):
            assert np.array_equal(da[k], db[k]), k


def test_eager_steps_reorder_without_changing_results():
    """Eager RK + PST calls also re-sort the records every reorder_interval
    steps; the results are bitwise those without re-sorts."""
    param, st = workloads.taylor_green_periodic(n=32, re=100.0, rk=4, jitter=0.2, seed=11)
    with simulation_from_state(param, st) as a, simulation_from_state(param, st) as b:
        a.set_tuning(reorder_interval=2)
        b.set_tuning(reorder_interval=1000)
        a.eager_step(7)
        b.eager_step(7)
        da, db = a.download(), b.download()
        for k in ("x", "y", "u", "v", "rho", "pre"):
            assert np.array_equal(da[k], db[k]), k


@pytest.mark.parametrize("precision", [0, 1])
def test_staged_scan_matches_per_lane_scan_bitwise(precision):
    """The wave-staged and the per-lane candidate scans build the same lists
    in the same order, so every field is bitwise equal.  At 256^2 most waves
    lie in one cell row and stage their candidates; waves at the periodic
    edges and across cell rows scan per lane, so both paths run in each
    pass."""
    param, st = workloads.taylor_green_periodic(n=256, re=1000.0, rk=4, jitter=0.2, seed=17)
    print(f"[gpu] staged vs per-lane scan, TG 256^2 precision={precision}", flush=True)
    with simulation_from_state(param, st, precision=precision) as a, \
            simulation_from_state(param, st, precision=precision) as b:
        b.set_scan(False)
        a.step(3)
        b.step(3)
        da, db = a.download(), b.download()
        for k in ("x", "y", "u", "v", "rho", "pre"):
            assert np.array_equal(da[k], db[k]), k


def test_staged_scan_matches_per_lane_scan_bitwise_walls_and_ensembles():
    """As above for an ensemble of walled cavities: waves holding wall slots
    or two instances scan per lane, full interior waves stage."""
    params, states, meta = workloads.cavity_ensemble(count=3, sizes=(64, 96), seed=9, precision=1)
    print(f"[gpu] staged vs per-lane scan, ensemble sizes={[m['n'] for m in meta]}", flush=True)
    ni = [s.num_int for s in states]
    ne = [s.num_ext for s in states]
    out = []
    for staged in (True, False):
        with Simulation(params, ni, ne, precision=1) as sim:
            sim.set_scan(staged)
            sim.upload(states)
            sim.init_cell()
            sim.ghost()
            sim.step(3)
            out.append(sim.download())
    for k in ("x", "y", "u", "v", "rho", "pre"):
        assert np.array_equal(out[0][k], out[1][k]), k


def test_list_overflow_path_matches():
    """A short on-chip list forces the direct candidate walk; results agree."""
    param, st = workloads.taylor_green_periodic(n=32, re=100.0, rk=2, jitter=0.2, seed=12)
    with simulation_from_state(param, st) as a, simulation_from_state(param, st) as b:
        b.set_tuning(list_len=8)
        a.step(2)
        b.step(2)
        assert b.stats().list_overflows > 0
        da, db = a.download(), b.download()
        for k in ("x", "y", "u", "v", "rho", "pre"):
            assert rel_err(db[k], da[k]) <= TOL, k


def test_tuning_rejects_a_block_the_kept_list_does_not_fit():
    """The quintic spline (h_eff = 4.65 dx) gets an 86-entry list; block 256
    with that list needs 87 * 256 * 4 B > 64 KiB, so the call fails and the
    handle keeps its settings (bitwise the same steps as an untouched one);
    block 256 with a 63-entry list fits.  A null output array is rejected."""
    import ctypes
    from lssph.engine import LssphError
    param, st = workloads.taylor_green_periodic(n=32, re=100.0, rk=2, jitter=0.2, seed=12, kernel_type=2)
    with simulation_from_state(param, st) as a, simulation_from_state(param, st) as b:
        with pytest.raises(LssphError, match="64 KiB"):
            b.set_tuning(block=256)
        with pytest.raises(LssphError, match="64 KiB"):
            b.set_tuning(block=256, list_len=64)
        a.step(2)
        b.step(2)
        da, db = a.download(), b.download()
        for k in ("x", "y", "u", "v", "rho", "pre"):
            assert np.array_equal(da[k], db[k]), k
        b.set_tuning(block=256, list_len=63)
        a.step(1)
        b.step(1)
        da, db = a.download(), b.download()
        for k in ("x", "y", "u", "v", "rho", "pre"):
            assert rel_err(db[k], da[k]) <= TOL, k
        assert b.lib.lssph_download_cell_index(b.h, None) != 0
        assert b"null" in b.lib.lssph_last_error(b.h)


def test_misuse_returns_error_codes_and_leaves_the_handle_working():
    """Calls out of order return LSSPH_E_STATE (-3) and bad values
    LSSPH_E_ARG (-1), each with a message; none changes the handle, which
    then steps bitwise like one that saw no bad call."""
    from lssph.engine import LssphError
    fields = ("ptype", "x", "y", "u", "v", "rho", "pre")
    param, st = workloads.taylor_green_periodic(n=24, re=100.0, rk=4, jitter=0.2, seed=21)

    def code(f, *a):
        with pytest.raises(LssphError) as e:
            f(*a)
        assert str(e.value).split(": ", 1)[1]
        return e.value.code

    with Simulation(param, [st.num_int], [st.num_ext]) as s, simulation_from_state(param, st) as ref:
        assert code(s.step, 1) == -3  # before lssph_upload
        assert code(s.set_cell_split, 0) == -1 and code(s.set_cell_split, 5) == -1
        assert code(s.calculate_RHS) == -3
        assert code(s.download) == -3
        s.upload(st)
        assert code(s.step, 1) == -3  # before lssph_init_cell
        s.init_cell()
        assert code(s.RK2) == -1  # the handle was created for RK4
        assert code(s.step, -1) == -1
        assert code(s.set_skin, 0.1) == -3  # the cell grid is fixed at upload
        assert code(s.set_cell_split, 3) == -3
        assert s.lib.lssph_set_scan(s.h, 2) == -1
        assert code(s.set_tuning, 32) == -1
        bad = {k: getattr(st, k).copy() for k in fields}
        bad["rho"][3] = 0.0
        assert code(s.upload, bad) == -1
        bad = {k: getattr(st, k).copy() for k in fields}
        bad["ptype"][0] = 1  # a wall type in the fluid block
        assert code(s.upload, bad) == -1
        s.step(3)
        ref.step(3)
        a, b = s.download(), ref.download()
        for k in fields[1:]:
            assert np.array_equal(a[k], b[k]), k


def test_upload_checks_every_field_size():
    """Every array of an upload must hold n_total entries; a short one is
    refused before the library reads it, and the handle stays usable."""
    param, st = workloads.taylor_green_periodic(n=16, re=100.0, rk=2, jitter=0.2, seed=3)
    with simulation_from_state(param, st) as s:
        bad = {k: getattr(st, k) for k in ("ptype", "x", "y", "u", "v", "rho", "pre")}
        bad["rho"] = bad["rho"][:-1]
        with pytest.raises(ValueError, match="rho"):
            s.upload(bad)
        counts, lst = s.neighbors()
        assert counts.size == st.num_int and lst.size == counts.sum() > 0


def test_mixed_precision_close_to_fp64():
    param, st = workloads.taylor_green_periodic(n=64, re=1000.0, rk=4, jitter=0.2, seed=13)
    with simulation_from_state(param, st, precision=0) as f64, simulation_from_state(param, st, precision=1) as mix:
        f64.step(20)
        mix.step(20)
        a, b = f64.download(), mix.download()
        t = 20 * param.dt
        l2_64 = workloads.tg_l2_error(param, a, t, st.num_int)
        l2_mx = workloads.tg_l2_error(param, b, t, st.num_int)
        assert abs(l2_mx - l2_64) <= 0.01 * l2_64, (l2_64, l2_mx)
        for k in ("u", "v"):
            assert rel_err(b[k], a[k]) < 1e-4, k


@pytest.mark.parametrize("box", ["periodic", "walls"])
@pytest.mark.parametrize("kernel_type", [1, 2, 4, 5])
def test_mixed_precision_close_to_fp64_other_kernels(kernel_type, box):
    """The mixed path's single-precision kernel shapes (cubic spline,
    quintic spline, Wendland C4, C6; Wendland C2 is covered above) against
    the FP64 path, which the reference's fixtures pin for each.  After 10
    steps the measured norm-wise differences are at most 4.9e-8 (u, v) and
    1.2e-8 (rho) over the eight cases; the bounds leave about 20x room, while
    a wrong shape, derivative or normalisation constant is off by far more."""
    if box == "periodic":
        param, st = workloads.taylor_green_periodic(n=64, re=1000.0, rk=4, jitter=0.2, seed=19,
                                                    kernel_type=kernel_type)
    else:
        param, st = workloads.taylor_green_box(n=64, re=100.0, rk=2, jitter=0.2, seed=19, kernel_type=kernel_type)
    print(f"[gpu] mixed vs FP64, kernel_type={kernel_type} {box}", flush=True)
    out = []
    for prec in (0, 1):
        with simulation_from_state(param, st, precision=prec) as s:
            if box == "walls":
                s.ghost()
            s.step(10)
            out.append(s.download())
    n = st.num_int
    e = {k: rel_err(out[1][k][:n], out[0][k][:n]) for k in ("u", "v", "rho")}
    print(f"[gpu] mixed vs FP64: {e}", flush=True)
    assert e["u"] < 1e-6 and e["v"] < 1e-6 and e["rho"] < 3e-7, e


def test_mixed_fp64_fallback_path_matches_fp64():
    """A condition limit below every pivot ratio sends every mixed solve to
    its FP64 fallback: k_lsa's list-driven FP64 relaunch and the shifting's
    in-kernel LU on the FP32 sums (coefficients of the squares rescaled).
    The fallbacks are counted, and the velocities agree with the FP64 path
    at least as closely as the mixed path's (the RHS passes are FP64 there;
    the positions, which the FP32 shift sums set, agree equally)."""
    from lssph import config as lcfg
    param, st = workloads.taylor_green_periodic(n=64, re=1000.0, rk=4, jitter=0.2, seed=14)
    param_low = lcfg.init_param(param.cfg.replace(cond_max=1.0))
    with simulation_from_state(param, st, precision=0) as f64, \
            simulation_from_state(param_low, st, precision=1) as fb, \
            simulation_from_state(param, st, precision=1) as mix:
        for s in (f64, fb, mix):
            s.step(3)
        a, b, c = f64.download(), fb.download(), mix.download()
        assert fb.stats().fp64_fallbacks >= 4 * 3 * st.num_int
        assert mix.stats().fp64_fallbacks == 0
        errs = {k: (rel_err(b[k], a[k]), rel_err(c[k], a[k])) for k in ("x", "y", "u", "v", "rho", "pre")}
        print(f"[gpu] fallback vs fp64, mixed vs fp64: {errs}", flush=True)
        for k, (e_fb, e_mix) in errs.items():
            assert e_fb <= 1e-6, (k, e_fb)
        # the RHS passes are FP64 there; only the shifting sums in FP32
        for k in ("u", "v"):
            assert errs[k][0] <= errs[k][1], (k, errs[k])


def test_cavity_ensemble_matches_oracle_per_instance():
    params, states, meta = workloads.cavity_ensemble(count=5, sizes=(12, 16, 20), seed=5, precision=0)
    print(f"[gpu] ensemble sizes={[m['n'] for m in meta]}", flush=True)
    ni = [s.num_int for s in states]
    ne = [s.num_ext for s in states]
    with Simulation(params, ni, ne, precision=0) as sim:
        sim.upload(states)
        sim.init_cell()
        sim.ghost()
        sim.step(4)
        out = split_states(sim.download(), ni, ne)
    for p, st, got in zip(params, states, out):
        o = OracleSim(p.to_c(), st.num_int, st.num_ext)
        o.init(st)
        o.ghost()
        o.step(4)
        ref = o.get_state()
        e = _cmp(got, ref, ("x", "y", "u", "v", "rho", "pre"))
        assert max(e.values()) <= TOL, e


def test_boussinesq_ensemble_matches_oracle_per_instance():
    """Three convection boxes of different sizes and Rayleigh numbers in one
    handle (temperature field, adiabatic and fixed-temperature walls,
    WALL_MODEL 3): each instance matches the oracle run alone."""
    cases = [(16, 1e4, 31), (24, 1e5, 32), (20, 1e6, 33)]
    params, states = [], []
    for n, ra, seed in cases:
        p, st = workloads.boussinesq(n=n, ra=ra, jitter=0.2, seed=seed, precision=0)
        params.append(p)
        states.append(st)
    print(f"[gpu] Boussinesq ensemble {cases}", flush=True)
    ni = [s.num_int for s in states]
    ne = [s.num_ext for s in states]
    with Simulation(params, ni, ne, precision=0) as sim:
        sim.upload(states)
        sim.init_cell()
        sim.ghost()
        sim.step(3)
        out = split_states(sim.download(), ni, ne)
    for p, st, got in zip(params, states, out):
        o = OracleSim(p.to_c(), st.num_int, st.num_ext)
        o.init(st)
        o.ghost()
        o.step(3)
        e = _cmp(got, o.get_state(), ("x", "y", "u", "v", "rho", "pre", "tem"))
        assert max(e.values()) <= TOL, e


@pytest.mark.parametrize("kind", ["cavity", "boussinesq"])
def test_mixed_ensemble_equals_separate_handles_bitwise(kind):
    """On the mixed path an instance's arithmetic does not depend on the
    others in its handle (waves spanning two instances take the per-lane
    copy with the same operations): five cavities, or three convection
    boxes with their temperature field, stepped together equal each stepped
    in a handle of its own, bit for bit."""
    if kind == "cavity":
        params, states, meta = workloads.cavity_ensemble(count=5, sizes=(12, 16, 20), seed=6, precision=1)
    else:
        pairs = [workloads.boussinesq(n=n, ra=ra, jitter=0.2, seed=sd, precision=1)
                 for n, ra, sd in ((16, 1e4, 41), (24, 1e5, 42), (20, 1e6, 43))]
        params, states = [p for p, _ in pairs], [st for _, st in pairs]
    print(f"[gpu] mixed {kind} ensemble n_int={[st.num_int for st in states]}", flush=True)
    ni = [s.num_int for s in states]
    ne = [s.num_ext for s in states]
    with Simulation(params, ni, ne, precision=1) as sim:
        sim.upload(states)
        sim.init_cell()
        sim.ghost()
        sim.step(4)
        out = split_states(sim.download(), ni, ne)
    for p, st, got in zip(params, states, out):
        with simulation_from_state(p, st, precision=1) as one:
            one.ghost()
            one.step(4)
            ref = one.download()
        for k in ("x", "y", "u", "v", "rho", "pre") + (("tem",) if kind == "boussinesq" else ()):
            assert np.array_equal(got[k], ref[k]), k


@pytest.mark.parametrize("ppw", [2, 3])
def test_q20_ensemble_equals_separate_handles_bitwise(ppw):
    """SPH_MODEL 6 with lane groups (two particles per wavefront) or row
    groups (three): the groups of one wavefront may hold particles of
    different instances, each with its own constants and cells.  Four
    cavities (h_eff = 3.1 dx, about 30 neighbors) stepped together equal each
    stepped in a handle of its own, bit for bit."""
    from lssph import config as lcfg
    params, states, meta = workloads.cavity_ensemble(count=4, sizes=(12, 16, 20), seed=7, precision=0,
                                                     coe_h=1.55)
    params = [lcfg.init_param(p.cfg.replace(sph_model=6)) for p in params]
    print(f"[gpu] q = 20 ensemble, {ppw} particles per wavefront, n_int={[st.num_int for st in states]}",
          flush=True)
    ni = [s.num_int for s in states]
    ne = [s.num_ext for s in states]
    with Simulation(params, ni, ne, precision=0) as sim:
        sim.set_wave_particles(ppw)
        sim.upload(states)
        sim.init_cell()
        sim.ghost()
        sim.step(2)
        out = split_states(sim.download(), ni, ne)
    for p, st, got in zip(params, states, out):
        with simulation_from_state(p, st, precision=0) as one:
            one.set_wave_particles(ppw)
            one.ghost()
            one.step(2)
            ref = one.download()
        for k in ("x", "y", "u", "v", "rho", "pre"):
            assert np.array_equal(got[k], ref[k]), k


def test_heat_ensemble_steady_check_per_instance():
    """Diffusion boxes of different sizes and diffusivities in one handle:
    fields and the per-instance steady-state maxima max|f - f_prev| equal
    those of each box in a handle of its own, bit for bit."""
    pairs = [workloads.heat_periodic(n=n, kappa=k, jitter=0.1, seed=sd)
             for n, k, sd in ((24, 0.01, 51), (32, 0.02, 52), (40, 0.005, 53))]
    params, states = [p for p, _ in pairs], [st for _, st in pairs]
    print(f"[gpu] heat ensemble n_int={[st.num_int for st in states]}", flush=True)
    ni = [s.num_int for s in states]
    ne = [s.num_ext for s in states]
    with Simulation(params, ni, ne, precision=0) as sim:
        sim.upload(states)
        sim.init_cell()
        sim.step(3)
        flags, df = sim.check_steady_state()
        out = split_states(sim.download(), ni, ne)
    assert flags.shape == (3,) and np.all(df > 0.0)
    for k, (p, st) in enumerate(pairs):
        with simulation_from_state(p, st, precision=0) as one:
            one.step(3)
            f1, d1 = one.check_steady_state()
            ref = one.download()
        assert flags[k] == f1[0] and df[k] == d1[0], (k, df[k], d1[0])
        assert np.array_equal(out[k]["u"], ref["u"]), k


def test_large_ensemble_ghost_path_matches_oracle():
    """22 cavities of 256^2 carry 68 K virtual markers.  The ghost launcher
    picks its lanes per marker from the wall model, never from the marker
    count, so the order of a marker's sums, hence its wall values, does not
    depend on the size of the ensemble: the first and last instance equal
    the same cavity in a handle of its own bit for bit, and match the
    oracle."""
    params, states, meta = workloads.cavity_ensemble(count=22, sizes=(256,), seed=8, precision=0)
    ni = [s.num_int for s in states]
    ne = [s.num_ext for s in states]
    assert sum(ne) >= 65536
    print(f"[gpu] large ensemble: {len(states)} cavities, {sum(ne)} markers", flush=True)
    with Simulation(params, ni, ne, precision=0) as sim:
        sim.upload(states)
        sim.init_cell()
        sim.ghost()
        sim.step(2)
        out = split_states(sim.download(), ni, ne)
    for k in (0, len(states) - 1):
        p, st = params[k], states[k]
        with simulation_from_state(p, st, precision=0) as one:
            one.ghost()
            one.step(2)
            alone = one.download()
        for f in ("x", "y", "u", "v", "rho", "pre"):
            assert np.array_equal(out[k][f], alone[f]), (k, f)
        o = OracleSim(p.to_c(), st.num_int, st.num_ext)
        o.init(st)
        o.ghost()
        o.step(2)
        e = _cmp(out[k], o.get_state(), ("x", "y", "u", "v", "rho", "pre"))
        assert max(e.values()) <= TOL, (k, e)


def test_boussinesq_wall_model3_matches_oracle():
    param, st = workloads.boussinesq(n=24, ra=1e4, jitter=0.2, seed=3)
    o, s = _pair(param, st)
    with s:
        o.ghost()
        s.ghost()
        for _ in range(3):
            o.step(1)
            s.step(1)
        e = _cmp(s.download(), o.get_state(), ("x", "y", "u", "v", "rho", "pre", "tem"))
        assert max(e.values()) <= TOL, e


@pytest.mark.parametrize("case", ["periodic", "cavity"])
def test_incremental_cell_builds_keep_reference_order(case):
    """After steps (incremental builds after every stage and shift) the
    neighbor lists, in summation order, equal the oracle's (which rebuilds
    its cell lists from scratch) and those of a full rebuild."""
    if case == "periodic":
        param, st = workloads.taylor_green_periodic(n=48, re=100.0, rk=4, jitter=0.3, seed=21)
    else:
        param, st = workloads.cavity(n=24, re=100.0, rk=2)
    o, s = _pair(param, st)
    with s:
        o.ghost()
        s.ghost()
        for _ in range(4):
            o.step(1)
            s.step(1)
        co, lo = o.neighbors()
        cs, ls = s.neighbors()
        assert np.array_equal(co, cs) and np.array_equal(lo, ls)
        np.testing.assert_array_equal(s.cell_index(), o.cell_index())
        s.update_cell()
        cf, lf = s.neighbors()
        assert np.array_equal(cf, cs) and np.array_equal(lf, ls)
        e = _cmp(s.download(), o.get_state(), ("x", "y", "u", "v", "rho", "pre"))
        assert max(e.values()) <= TOL, e


@pytest.mark.parametrize("rk", [2, 4])
def test_far_moves_between_builds(rk):
    """A uniform flow fast enough to carry particles one to three cells per
    stage, with the cells rebuilt after every pass (skin 0): the incremental
    build must take those particles from the stray list (and select members
    one by one in cells with many arrivals); the states and neighbor lists
    still equal the oracle's."""
    param, st = workloads.taylor_green_periodic(n=24, re=100.0, rk=rk, jitter=0.3, seed=23)
    cw = 1.0 / int(1.0 / param.h_eff)
    speed = 3.0 * cw / param.dt
    st.u[:] = speed
    st.v[:] = 0.7 * speed
    o = OracleSim(param.to_c(), st.num_int, st.num_ext)
    o.init(st)
    s = simulation_from_state(param, st, skin=0.0)
    with s:
        for _ in range(2):
            o.step(1)
            s.step(1)
            co, lo = o.neighbors()
            cs, ls = s.neighbors()
            assert np.array_equal(co, cs) and np.array_equal(lo, ls)
        e = _cmp(s.download(), o.get_state(), ("x", "y", "u", "v", "rho", "pre"))
        assert max(e.values()) <= TOL, e


@pytest.mark.parametrize("rk", [2, 4])
def test_periodic_wraps_between_builds(rk):
    """A uniform diagonal flow at 0.3 of the skin per step carries particles
    across both periodic edges while the cell list built at the start of the
    step is still in use: the image shift of each pair is taken from the
    positions, so states and neighbor lists stay the oracle's."""
    param, st = workloads.taylor_green_periodic(n=24, re=100.0, rk=rk, jitter=0.3, seed=29)
    st.x[:] = np.mod(st.x + 0.018, 1.0)  # a column of particles near each edge
    st.y[:] = np.mod(st.y + 0.017, 1.0)
    step = 0.3 * 0.08 * param.h_eff
    st.u[:] = 0.8 * step / param.dt
    st.v[:] = 0.6 * step / param.dt
    o = OracleSim(param.to_c(), st.num_int, st.num_ext)
    o.init(st)
    with simulation_from_state(param, st) as s:
        for _ in range(6):
            o.step(1)
            s.step(1)
            co, lo = o.neighbors()
            cs, ls = s.neighbors()
            assert np.array_equal(co, cs) and np.array_equal(lo, ls)
        got = s.download()
        n = st.num_int
        assert np.any(got["x"][:n] < st.x[:n] - 0.5) and np.any(got["y"][:n] < st.y[:n] - 0.5)
        e = _cmp(got, o.get_state(), ("x", "y", "u", "v", "rho", "pre"))
        assert max(e.values()) <= TOL, e


def test_skin_guard_reports_moves_beyond_the_skin():
    """With the default skin, the same fast flow moves particles farther than
    half the skin within a step: the library reports it instead of using a
    stale cell list."""
    from lssph.engine import LssphError
    param, st = workloads.taylor_green_periodic(n=24, re=100.0, rk=4, jitter=0.3, seed=23)
    speed = 3.0 / int(1.0 / param.h_eff) / param.dt
    st.u[:] = speed
    st.v[:] = 0.7 * speed
    with simulation_from_state(param, st) as s:
        s.step(1)
        with pytest.raises(LssphError, match="skin"):
            s.synchronize()


@pytest.mark.parametrize("skin", [0.0, 0.3])
def test_results_do_not_depend_on_the_skin(skin):
    """The cell-list skin only changes which candidates are tested; the
    trajectory matches the oracle either way."""
    param, st = workloads.taylor_green_periodic(n=40, re=100.0, rk=4, jitter=0.2, seed=31)
    o = OracleSim(param.to_c(), st.num_int, st.num_ext)
    o.init(st)
    with simulation_from_state(param, st, skin=skin) as s:
        for _ in range(3):
            o.step(1)
            s.step(1)
        e = _cmp(s.download(), o.get_state(), ("x", "y", "u", "v", "rho", "pre"))
        assert max(e.values()) <= TOL, e


@pytest.mark.parametrize("model", [1, 2, 4, 5, 6])
def test_periodic_other_sph_models_match_oracle(model):
    """Classical SPH (SPH_MODEL 1/2, with the q = 2 shifting) and LS-SPH with
    the 3rd- to 5th-order Taylor bases (q = 9, 14 per thread; q = 20 a lane
    group per particle) in the periodic box: neighbor lists, the first RHS
    and three RK4 + PST steps against the oracle."""
    from lssph import config as lcfg
    base, st = workloads.taylor_green_periodic(n=32, re=100.0, rk=4, jitter=0.2, seed=40 + model)
    param = lcfg.init_param(base.cfg.replace(sph_model=model))
    # RHS_rho = -rho div(u) nearly cancels in this flow, which magnifies the
    # last-bit differences of the derivatives (~2e-15 in RHS_u) by ~100 and,
    # for q = 9 / 14 / 20, by cond(M) ~ 1e2 / 3e3 / 4e7 on top
    tol = {1: TOL, 2: TOL, 4: 1e-11, 5: 1e-11, 6: 2e-9}[model]
    o, s = _pair(param, st)
    with s:
        r_o, r_s = o.calculate_rhs(), s.calculate_RHS()
        e = _cmp(r_s, r_o, ("u", "v", "rho"))
        assert max(e.values()) <= tol, e
        for _ in range(3):
            o.step(1)
            s.step(1)
        e = _cmp(s.download(), o.get_state(), ("x", "y", "u", "v", "rho", "pre"))
        assert max(e.values()) <= tol, e


@pytest.mark.parametrize("model", [2, 4])
def test_mixed_request_runs_fp64_for_other_models(model):
    """LSSPH_MIXED applies to SPH_MODEL 3; the other models compute in FP64
    whatever the request, bit for bit."""
    from lssph import config as lcfg
    base, st = workloads.taylor_green_periodic(n=24, re=100.0, rk=2, jitter=0.2, seed=50 + model)
    param = lcfg.init_param(base.cfg.replace(sph_model=model))
    with simulation_from_state(param, st, precision=0) as a, simulation_from_state(param, st, precision=1) as b:
        a.step(2)
        b.step(2)
        da, db = a.download(), b.download()
        for k in ("x", "y", "u", "v", "rho", "pre"):
            assert np.array_equal(da[k], db[k]), k


@pytest.mark.parametrize("model,ppw", [(6, 1), (6, 2), (6, 3)])
@pytest.mark.parametrize("box", ["periodic", "walls", "bous", "heat", "quintic"])
def test_wave_particles_do_not_change_results(model, ppw, box):
    """SPH_MODEL 6 with 1, 2 or 3 particles per wavefront (lane groups of 64
    or 32 lanes, or row groups of 20): the same results bit for bit as the
    default setting, over two RK4 + PST steps (walls: the ghost fill in
    between), for the flow, convection (temperature column) and heat
    (one field, no shifting) operators, and with the quintic spline, whose
    support (h_eff = 4.65 dx, about 68 neighbors) takes the sums over four
    chunks of neighbors."""
    from lssph import config as lcfg
    if box == "periodic":
        base, st = workloads.taylor_green_periodic(n=28, re=100.0, rk=4, jitter=0.2, seed=70 + model)
    elif box == "walls":
        base, st = workloads.taylor_green_box(n=24, re=100.0, rk=4, jitter=0.2, seed=70 + model)
    elif box == "bous":
        base, st = workloads.boussinesq(n=24, ra=1e5, jitter=0.2, seed=70 + model)
    elif box == "quintic":
        base, st = workloads.taylor_green_periodic(n=32, re=100.0, rk=4, jitter=0.2, seed=80 + model,
                                                   kernel_type=2)
    else:
        base, st = workloads.heat_periodic(n=28, kappa=0.01, jitter=0.1, seed=70 + model)
    param = lcfg.init_param(base.cfg.replace(sph_model=model))
    print(f"[gpu] SPH_MODEL {model}, {ppw} particles per wavefront, {box}", flush=True)
    out = []
    for setting in (0, ppw):
        with simulation_from_state(param, st, precision=0) as s:
            s.set_wave_particles(setting)
            if box in ("walls", "bous"):
                s.ghost()
            rhs = s.calculate_RHS()
            s.step(2)
            out.append((rhs, s.download()))
    keys = ("u",) if box == "heat" else ("u", "v", "rho")
    for k in keys:
        assert np.array_equal(out[0][0][k], out[1][0][k]), k
    fields = ("x", "y", "u", "v", "rho", "pre") + (("tem",) if box == "bous" else ())
    for k in fields:
        assert np.array_equal(out[0][1][k], out[1][1][k]), k


def test_wave_particles_rejects_bad_values():
    from lssph import config as lcfg
    from lssph.engine import LssphError
    base, st = workloads.taylor_green_periodic(n=16, re=100.0, rk=4, jitter=0.0, seed=0)
    param = lcfg.init_param(base.cfg.replace(sph_model=6))
    with simulation_from_state(param, st, precision=0) as s:
        for bad in (4, -1):  # 1, 2 or 3 (0: the default)
            with pytest.raises(LssphError):
                s.set_wave_particles(bad)


@pytest.mark.parametrize("split", [1, 3, 4])
@pytest.mark.parametrize("case", ["periodic", "smallest", "cavity"])
def test_cell_split_matches_oracle(split, case):
    """Internal cells 1 / split of the stencil reach wide in x (the default
    is 2, which every other test runs): the stencil covers 2 split + 1
    columns, wraps across periodic edges on split columns per side, and the
    incremental builds move particles between narrower cells.  Neighbor
    lists in summation order, the first RHS, and the states after steps
    against the oracle; a full rebuild gives the same lists."""
    if case == "periodic":
        param, st = workloads.taylor_green_periodic(n=40, re=100.0, rk=4, jitter=0.3, seed=41)
    elif case == "smallest":  # 3 cell rows, 3 split columns: every stencil wraps
        param, st = workloads.taylor_green_periodic(n=10, re=100.0, rk=4, jitter=0.2, seed=42)
    else:
        param, st = workloads.cavity(n=24, re=100.0, rk=2)
    print(f"[gpu] cell split {split}, {case}", flush=True)
    o = OracleSim(param.to_c(), st.num_int, st.num_ext)
    o.init(st)
    with simulation_from_state(param, st, cell_split=split) as s:
        o.ghost()
        s.ghost()
        co, lo = o.neighbors()
        cs, ls = s.neighbors()
        assert np.array_equal(co, cs) and np.array_equal(lo, ls)
        r_o, r_s = o.calculate_rhs(), s.calculate_RHS()
        assert max(_cmp(r_s, r_o, ("u", "v", "rho")).values()) <= TOL
        for _ in range(4):
            o.step(1)
            s.step(1)
        co, lo = o.neighbors()
        cs, ls = s.neighbors()
        assert np.array_equal(co, cs) and np.array_equal(lo, ls)
        s.update_cell()
        cf, lf = s.neighbors()
        assert np.array_equal(cf, cs) and np.array_equal(lf, ls)
        e = _cmp(s.download(), o.get_state(), ("x", "y", "u", "v", "rho", "pre"))
        assert max(e.values()) <= TOL, e


@pytest.mark.parametrize("split", [1, 3])
def test_cell_split_far_moves_and_wraps(split):
    """Particles carried one to three cells per stage with the cells rebuilt
    after every pass (skin 0), across both periodic edges: with narrower
    cells more of them leave the stencil of their old cell and go through
    the stray list; states and neighbor lists stay the oracle's."""
    param, st = workloads.taylor_green_periodic(n=24, re=100.0, rk=4, jitter=0.3, seed=43)
    cw = 1.0 / int(1.0 / param.h_eff)
    speed = 2.0 * cw / param.dt
    st.u[:] = speed
    st.v[:] = -0.6 * speed
    o = OracleSim(param.to_c(), st.num_int, st.num_ext)
    o.init(st)
    print(f"[gpu] cell split {split}, far moves", flush=True)
    with simulation_from_state(param, st, skin=0.0, cell_split=split) as s:
        for _ in range(2):
            o.step(1)
            s.step(1)
            co, lo = o.neighbors()
            cs, ls = s.neighbors()
            assert np.array_equal(co, cs) and np.array_equal(lo, ls)
        e = _cmp(s.download(), o.get_state(), ("x", "y", "u", "v", "rho", "pre"))
        assert max(e.values()) <= TOL, e
