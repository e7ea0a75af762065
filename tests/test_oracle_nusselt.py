"""The oracle's restatement of the analysis' wall-gradient fit and Nusselt
number (oracle/analysis.py) against closed forms.  The q = 6 type-B fit
reproduces any quadratic field exactly, so with T set on every particle
(fluid and wall) from a quadratic in (x, y) the local Nusselt numbers are
known at each wall point:
  T = 1 - y:                     Nu = 1 on both walls (pure conduction);
  T = 1 - y + a y (1 - y) + b x y:  dT/dy = -1 + a (1 - 2y) + b x, so the wall
      means over x = linspace(0, 1, n) are Nu_bottom = 1 - a - b/2 and
      Nu_top = 1 + a - b/2."""
import numpy as np
import pytest

from lssph import workloads
from oracle import analysis


def _state(n=24, a=0.0, b=0.0, seed=3):
    param, st = workloads.boussinesq(n=n, ra=1e5, jitter=0.2, seed=seed)
    xp = st.x - param.wl_thick
    yp = st.y - param.wl_thick
    tem = 1.0 - yp + a * yp * (1.0 - yp) + b * xp * yp
    d = {"x": st.x, "y": st.y, "tem": tem, "rho": st.rho * (1.0 + 0.01 * np.sin(7 * xp))}
    return param, st, d


def test_conduction_profile_gives_nu_one():
    param, st, d = _state()
    nb, nt, loc = analysis.nusselt(param, d, st.num_total)
    assert abs(nb - 1.0) < 1e-12 and abs(nt - 1.0) < 1e-12, (nb, nt)
    assert np.all(np.abs(loc[0] - 1.0) < 1e-11) and np.all(np.abs(loc[1] - 1.0) < 1e-11)


@pytest.mark.parametrize("a,b", [(0.3, 0.0), (0.0, 0.8), (-0.4, 0.5)])
def test_quadratic_profiles_give_the_closed_form(a, b):
    param, st, d = _state(a=a, b=b)
    nb, nt, loc = analysis.nusselt(param, d, st.num_total)
    assert abs(nb - (1.0 - a - 0.5 * b)) < 1e-11, (nb, a, b)
    assert abs(nt - (1.0 + a - 0.5 * b)) < 1e-11, (nt, a, b)
    x = np.linspace(0.0, 1.0, param.cfg.num_x)
    assert np.allclose(loc[0], 1.0 - a - b * x, rtol=0, atol=1e-10)


def test_value_and_x_gradient_of_the_fit():
    """The fit's constant term and d/dx at interior points of a quadratic."""
    param, st, d = _state(a=0.3, b=0.8)
    xp, yp = 0.37, 0.61
    out = analysis.probe_b(param.h, param.h_eff, param.sp_mass, 3, 6, [xp + param.wl_thick],
                           [yp + param.wl_thick], st.x, st.y, d["tem"], d["rho"])[0]
    t = 1.0 - yp + 0.3 * yp * (1.0 - yp) + 0.8 * xp * yp
    assert abs(out[0] - t) < 1e-12 and abs(out[1] - 0.8 * yp) < 1e-11
    assert abs(out[2] - (-1.0 + 0.3 * (1.0 - 2 * yp) + 0.8 * xp)) < 1e-11


def test_too_few_neighbors_give_nan():
    param, st, d = _state()
    far = 10.0
    out = analysis.probe_b(param.h, param.h_eff, param.sp_mass, 3, 6, [far], [far], st.x, st.y, d["tem"], d["rho"])
    assert np.all(np.isnan(out))
