"""Ensemble isolation (lssph_set_isolation): an instance whose particles
leave its cell grid or outrun the cell-list skin -- a diverged run -- is
retired, and the other instances of the handle run on exactly as they would
alone.  Without isolation the same move fails the handle (LSSPH_E_DOMAIN),
the behaviour of a single box."""
import numpy as np
import pytest

from lssph import workloads
from lssph.engine import LssphError, Simulation, simulation_from_state, split_states

pytestmark = pytest.mark.gpu


def _ensemble(spoil):
    pairs = [workloads.taylor_green_periodic(n=n, re=1000.0, rk=4, jitter=0.2, seed=sd, precision=1)
             for n, sd in ((24, 51), (32, 52), (28, 53))]
    params = [p for p, _ in pairs]
    states = [st.copy() for _, st in pairs]
    spoil(states[1])
    return params, states


def _blow_up(st):
    st.u[:] *= 1.0e4  # moves far beyond the skin in the first stage


def _nan(st):
    st.u[7] = np.nan  # a non-finite velocity: non-finite positions after the first stage


@pytest.mark.parametrize("spoil", [_blow_up, _nan], ids=["fast", "nan"])
def test_diverged_instance_is_retired_and_the_others_run_on(spoil):
    params, states = _ensemble(spoil)
    ni = [s.num_int for s in states]
    ne = [s.num_ext for s in states]
    print(f"[gpu] isolation: ensemble {ni}, instance 1 spoiled ({spoil.__name__})", flush=True)
    with Simulation(params, ni, ne, precision=1) as sim:
        sim.set_isolation(True)
        sim.upload(states)
        sim.init_cell()
        sim.step(4)
        sim.eager_step(2)
        status, reason = sim.instance_status(reasons=True)
        out = split_states(sim.download(), ni, ne)
    assert list(status) == [0, 1, 0], status
    assert reason[1] in (Simulation.RETIRED_SKIN, Simulation.RETIRED_ESCAPED) and reason[0] == reason[2] == 0
    for k in (0, 2):
        with simulation_from_state(params[k], states[k], precision=1) as one:
            one.step(4)
            one.eager_step(2)
            ref = one.download()
        for f in ("x", "y", "u", "v", "rho", "pre"):
            assert np.array_equal(out[k][f], ref[f]), (k, f)


def test_without_isolation_a_diverged_instance_fails_the_handle():
    params, states = _ensemble(_blow_up)
    ni = [s.num_int for s in states]
    ne = [s.num_ext for s in states]
    with Simulation(params, ni, ne, precision=1) as sim:
        sim.upload(states)
        sim.init_cell()
        with pytest.raises(LssphError) as e:
            sim.step(2)
            sim.download()
        assert e.value.code == -4


def test_without_isolation_a_nan_instance_is_reported_and_stops_advancing():
    """Without isolation a non-finite position is an LSSPH_E_DOMAIN error of
    the handle, reported by the next synchronising call, and the instance
    stops advancing at the pass that made it: its particles keep their cells
    instead of being binned, step after step, into one cell whose scans and
    rebuilds would grow with the square of its size.  Steps after the error
    cost what steps of a healthy handle cost."""
    import time
    params, states = _ensemble(_nan)
    ni = [s.num_int for s in states]
    ne = [s.num_ext for s in states]
    with Simulation(params, ni, ne, precision=1) as sim:
        sim.upload(states)
        sim.init_cell()
        sim.eager_step(2)
        sim.stats()  # a synchronising call that does not check the domain counters
        t0 = time.perf_counter()
        sim.eager_step(20)
        sim.stats()
        frozen_ms = (time.perf_counter() - t0) * 1e3 / 20
        with pytest.raises(LssphError) as e:
            sim.download()
        assert e.value.code == -4
    healthy, hstates = _ensemble(lambda st: None)
    with Simulation(healthy, ni, ne, precision=1) as sim:
        sim.upload(hstates)
        sim.init_cell()
        sim.eager_step(2)
        sim.stats()
        t0 = time.perf_counter()
        sim.eager_step(20)
        sim.stats()
        healthy_ms = (time.perf_counter() - t0) * 1e3 / 20
    print(f"[gpu] ms per eager step: after the error {frozen_ms:.3f}, healthy {healthy_ms:.3f}", flush=True)
    assert frozen_ms <= 2.0 * healthy_ms + 1.0


def test_isolation_restarts_its_count_and_a_healthy_ensemble_retires_nothing():
    params, states = _ensemble(lambda st: None)
    ni = [s.num_int for s in states]
    ne = [s.num_ext for s in states]
    with Simulation(params, ni, ne, precision=1) as sim:
        sim.upload(states)
        sim.init_cell()
        sim.set_isolation(True)
        sim.step(3)
        assert list(sim.instance_status()) == [0, 0, 0]
        sim.set_isolation(False)
        sim.step(1)
        sim.download()
