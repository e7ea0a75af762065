"""The reference's scheme (the oracle, pinned per step to the compiled
reference) cannot run five of the 64 lid-driven cavities of
workloads.cavity_ensemble for 2000 steps: the 32^2 instances at Re 535-750
(cell Reynolds numbers 17-23) diverge between steps 1700 and 2000, while
the 32^2 instances just below that range run through.  The GPU ensemble
test retires exactly these instances (ensemble isolation).  The compiled
reference itself, built from oracle/ref_cases/cavity32_re722.h (instance 5),
grows its velocity maximum from 0.94 to 4.7 between steps 1600 and 1800
and crashes before step 2000."""
import numpy as np
import pytest

from lssph import workloads
from oracle.pyoracle import OracleSim

DIVERGING = [5, 20, 45, 48, 49]
RUNNING_32 = [53, 0, 51]  # the 32^2 instances with the next-highest Re (456, 390, 358)


def _runs_through(p, st, steps=2000):
    o = OracleSim(p.to_c(), st.num_int, st.num_ext, native=True)
    o.init(st)
    try:
        for s in range(0, steps, 100):
            o.step(100)
            g = o.get_state()
            if not np.all(np.isfinite(g["u"][:st.num_int])) or np.max(np.abs(g["u"][:st.num_int])) > 3.0:
                return False, s + 100
    except RuntimeError:
        return False, s
    return True, steps


@pytest.mark.parametrize("k", DIVERGING + RUNNING_32)
def test_oracle_divergence_of_the_small_high_re_cavities(k):
    OracleSim.set_threads(8)
    params, states, meta = workloads.cavity_ensemble(count=64, precision=0)
    assert meta[k]["n"] == 32
    ok, step = _runs_through(params[k], states[k])
    print(f"instance {k}: Re {meta[k]['re']:.1f}, {'runs through' if ok else f'diverges by step {step}'}")
    assert ok == (k in RUNNING_32), (k, step)
    if not ok:
        assert step >= 1700
