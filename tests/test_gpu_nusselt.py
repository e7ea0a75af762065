"""The library's probe fit (lssph_probe) and Nusselt number
(lssph.analysis.nusselt) against closed forms -- quadratic temperature
fields, which the q = 6 type-B fit reproduces exactly, so the local Nusselt
numbers of both walls are known (see tests/test_oracle_nusselt.py) -- and
against the oracle's restatement of the reference analysis on a convecting
state."""
import numpy as np
import pytest

from lssph import analysis, workloads
from lssph.engine import simulation_from_state
from oracle import analysis as oracle_analysis

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("a,b", [(0.0, 0.0), (0.3, 0.0), (-0.4, 0.5)])
@pytest.mark.parametrize("precision", [0, 1])
def test_quadratic_profiles_give_the_closed_form(a, b, precision):
    param, st = workloads.boussinesq(n=48, ra=1e5, jitter=0.2, seed=61, precision=precision)
    xp = st.x - param.wl_thick
    yp = st.y - param.wl_thick
    st = st.copy()
    st.tem[:] = 1.0 - yp + a * yp * (1.0 - yp) + b * xp * yp
    print(f"[gpu] Nu closed form a={a} b={b} precision={precision}", flush=True)
    with simulation_from_state(param, st, precision=precision) as s:
        nb, nt, loc = analysis.nusselt(s, param)
    assert abs(nb - (1.0 - a - 0.5 * b)) < 1e-11, nb
    assert abs(nt - (1.0 + a - 0.5 * b)) < 1e-11, nt
    x = np.linspace(0.0, 1.0, param.cfg.num_x)
    assert np.allclose(loc[0], 1.0 - a - b * x, rtol=0, atol=1e-10)


def test_probe_value_and_gradients_of_other_fields():
    """u, v, p set to quadratics in (x, y): the fit returns their values and
    both gradient components at interior and wall points, for q = 3 (linear
    fields) and 6."""
    param, st = workloads.cavity(n=40, re=400.0, jitter=0.2, seed=62)
    st = st.copy()
    xp, yp = st.x, st.y
    st.u[:] = 0.3 + 0.5 * xp - 0.2 * yp
    st.v[:] = 1.0 - xp * xp + 0.5 * xp * yp
    st.pre[:] = 2.0 * yp * yp - xp
    px = np.array([0.3, 0.55, param.wl_thick, 0.8 + param.wl_thick])
    py = np.array([0.4, param.wl_thick + 0.5, 0.7, param.wl_thick])
    with simulation_from_state(param, st, precision=1) as s:
        u3 = s.probe(px, py, field="u", q=3)
        v6 = s.probe(px, py, field="v", q=6)
        p10 = s.probe(px, py, field="pre", q=10)
    assert np.allclose(u3[:, 0], 0.3 + 0.5 * px - 0.2 * py, rtol=0, atol=1e-12)
    assert np.allclose(u3[:, 1], 0.5, atol=1e-11) and np.allclose(u3[:, 2], -0.2, atol=1e-11)
    assert np.allclose(v6[:, 0], 1.0 - px * px + 0.5 * px * py, rtol=0, atol=1e-12)
    assert np.allclose(v6[:, 1], -2 * px + 0.5 * py, atol=1e-10) and np.allclose(v6[:, 2], 0.5 * px, atol=1e-10)
    assert np.allclose(p10[:, 1], -1.0, atol=1e-9) and np.allclose(p10[:, 2], 4.0 * py, atol=1e-9)


def test_nusselt_matches_the_oracle_on_a_convecting_state():
    param, st = workloads.boussinesq(n=96, ra=1e6, jitter=0.2, seed=63)
    print("[gpu] Nu of a 96^2 convection box after 300 steps vs the oracle", flush=True)
    with simulation_from_state(param, st, precision=0) as s:
        s.ghost()
        s.step(300)
        nb, nt, loc = analysis.nusselt(s, param)
        state = s.download()
    ob, ot, oloc = oracle_analysis.nusselt(param, state, st.num_total)
    print(f"[gpu] Nu GPU ({nb:.12f}, {nt:.12f}) oracle ({ob:.12f}, {ot:.12f})", flush=True)
    assert abs(nb - ob) <= 1e-10 * abs(ob) and abs(nt - ot) <= 1e-10 * abs(ot)
    assert np.allclose(loc[0], oloc[0], rtol=1e-9, atol=1e-12)


def test_probe_outside_the_box_gives_nan_and_bad_arguments_fail():
    from lssph.engine import LssphError
    param, st = workloads.boussinesq(n=24, ra=1e5, jitter=0.0, seed=64)
    with simulation_from_state(param, st, precision=0) as s:
        out = s.probe(np.array([50.0]), np.array([50.0]))
        assert np.all(np.isnan(out))
        with pytest.raises(LssphError):
            s.probe(np.array([0.5]), np.array([0.5]), q=5)
        with pytest.raises(LssphError):
            s.probe(np.array([0.5]), np.array([0.5]), inst=1)
