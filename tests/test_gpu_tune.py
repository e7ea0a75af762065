"""The autotuner on the GPU, on small workloads: the setting it picks runs,
and the settings that only change the layout on the GPU (threads per block,
list length, particles per wavefront) give the same results bit for bit;
the skin, which changes the summation order within stencil rows, moves the
mixed path's FP32 sums by their rounding only: the results of two skins
differ by no more than the mixed path's own error against an FP64 run of
the same ensemble (twice it, per field; measured 4.7e-7 of max|v| after 6
cavity steps, where the mixed run's own error is of the same order)."""
import numpy as np
import pytest

from helpers import rel_err
from lssph import config as lcfg
from lssph import tune, workloads
from lssph.engine import Simulation, simulation_from_state, split_states

pytestmark = pytest.mark.gpu


def _run(make, setting, steps):
    sim = make(setting.skin)
    try:
        sim.set_tuning(block=setting.block, list_len=setting.list_len)
        if setting.wave_particles:
            sim.set_wave_particles(setting.wave_particles)
        sim.step(steps)
        return sim.download()
    finally:
        sim.close()


def test_autotune_on_an_ensemble_and_its_pick_leaves_results_unchanged():
    params, states, _ = workloads.cavity_ensemble(count=4, sizes=(24, 32, 48), seed=11, precision=1)
    params64, _, _ = workloads.cavity_ensemble(count=4, sizes=(24, 32, 48), seed=11, precision=0)
    ni = [s.num_int for s in states]
    ne = [s.num_ext for s in states]

    def make(skin, precision=1):
        sim = Simulation(params if precision else params64, ni, ne, precision=precision, skin=skin)
        sim.upload(states)
        sim.init_cell()
        sim.ghost()
        return sim

    print("[gpu] autotune, 4-cavity ensemble", flush=True)
    r = tune.autotune(make, steps=2, blocks=(64, 128, 256), skins=(0.06, 0.08, 0.12))
    print(f"[gpu] tuned: {r.as_dict()['best']}", flush=True)
    assert not r.best.error and np.isfinite(r.best.ms_per_step)
    ok = [s for s in r.table if not s.error]
    assert len(ok) >= 6
    ref = _run(make, tune.Setting(128, 0, 0.08), 6)
    fp64 = _run(lambda skin: make(skin, precision=0), tune.Setting(128, 0, 0.08), 6)
    own = {k: rel_err(ref[k], fp64[k]) for k in ("x", "y", "u", "v", "rho", "pre")}
    print(f"[gpu] mixed against FP64 after 6 steps: {own}", flush=True)
    for s in ok:
        got = _run(make, s, 6)
        for k in ("x", "y", "u", "v", "rho", "pre"):
            if s.skin == 0.08:
                assert np.array_equal(got[k], ref[k]), (s, k)
            else:
                assert rel_err(got[k], ref[k]) <= 2.0 * own[k] + 1e-12, (s, k, own[k])


def test_autotune_particles_per_wavefront_of_sph_model_6():
    base, st = workloads.taylor_green_periodic(n=40, re=100.0, rk=4, jitter=0.2, seed=12)
    param = lcfg.init_param(base.cfg.replace(sph_model=6))

    def make(skin):
        return simulation_from_state(param, st, precision=0, skin=skin)

    print("[gpu] autotune, SPH_MODEL 6, particles per wavefront 1, 2 and 3", flush=True)
    r = tune.autotune(make, steps=1, blocks=(128,), skins=(0.08,), wave_particles=(1, 2, 3),
                      default=(128, 0, 0.08, 1))
    print(f"[gpu] tuned: {r.as_dict()}", flush=True)
    assert {s.wave_particles for s in r.table if not s.error} == {1, 2, 3}
    a = _run(make, tune.Setting(128, 0, 0.08, 1), 2)
    for ppw in (2, 3):
        b = _run(make, tune.Setting(128, 0, 0.08, ppw), 2)
        for k in ("x", "y", "u", "v", "rho", "pre"):
            assert np.array_equal(a[k], b[k]), (ppw, k)
