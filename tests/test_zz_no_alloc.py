"""lssph_probe_static, the entry point that allocates nothing, against the
allocating lssph_probe it is a variant of: the same kernel on the same
state gives the same numbers bit for bit, on two calls in a row (the second
with fewer probes, so stale staging contents would show), and a call beyond
its staging buffer is refused."""
import numpy as np
import pytest

from lssph import workloads
from lssph.engine import LssphError, simulation_from_state

pytestmark = pytest.mark.gpu


def _points(param, n, seed):
    rng = np.random.default_rng(seed)
    wl = param.wl_thick
    x = wl + rng.uniform(0.0, param.cfg.len_x, n)
    y = wl + rng.uniform(0.0, param.cfg.len_y, n)
    x[: n // 4] = wl + np.linspace(0.0, param.cfg.len_x, n // 4)  # on the bottom wall
    y[: n // 4] = wl
    return x, y


def test_static_probe_equals_the_allocating_probe_on_two_calls():
    param, st = workloads.boussinesq(n=256, ra=1e6, jitter=0.2, seed=71, precision=1)
    print("[gpu] lssph_probe_static vs lssph_probe, 256^2 convection box", flush=True)
    with simulation_from_state(param, st, precision=1) as s:
        s.ghost()
        s.step(10)
        for n, field, q, seed in ((4096, "tem", 6, 1), (300, "u", 3, 2)):
            x, y = _points(param, n, seed)
            ref = s.probe(x, y, field=field, q=q)
            got = s.probe_static(x, y, field=field, q=q)
            assert got.shape == ref.shape
            assert np.array_equal(got, ref, equal_nan=True), (n, field)
            assert np.all(np.isfinite(got))


def test_static_probe_refuses_more_probes_than_its_buffer():
    param, st = workloads.boussinesq(n=32, ra=1e5, jitter=0.0, seed=72)
    with simulation_from_state(param, st, precision=0) as s:
        x = np.full(16385, 0.5)
        with pytest.raises(LssphError) as e:
            s.probe_static(x, x)
        assert e.value.code == -1 and "16384" in str(e.value)
        out = s.probe_static(x[:16384], x[:16384])  # the maximum itself is accepted
        assert np.all(np.isfinite(out))
