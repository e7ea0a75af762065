"""HIP library against the compiled reference's golden fixtures (MI355X).

The FP64 GPU path must match the reference Fortran to 1e-12 norm-wise
relative per field, per pass and per step (1e-10 for q = 20, helpers.CASE_TOL);
neighbor sets and cell indices must be bit-exact (same sets, same summation
order).  The cases cover SPH_MODEL 1-6.
"""
import numpy as np
import pytest

from helpers import CASES as ALL_CASES, Golden, case_tol, max_rel, rel_err

pytestmark = pytest.mark.gpu
CASES = ALL_CASES


def _sim(g: Golden, st=None, precision=0):
    from lssph.engine import simulation_from_state
    print(f"[gpu] case={g.name} n_total={g.n_total} precision={precision}", flush=True)
    return simulation_from_state(g.param, st if st is not None else g.input_state(), precision=precision)


@pytest.mark.parametrize("case", CASES)
def test_initial_ghost_markers_cells(case):
    g = Golden(case)
    with _sim(g) as s:
        assert np.array_equal(s.cell_index(), g["cell0"])
        s.ghost()
        errs = max_rel(s.download(), g, "state0")
        assert max(errs.values()) <= case_tol(case), errs
        vm = s.download_vm()
        assert np.array_equal(vm["ptype"], g["vm_ptype"])
        for k in ("x", "y", "u", "v", "pre") + (("tem",) if g.has_tem else ()):
            assert rel_err(vm[k], g[f"vm_{k}"]) <= case_tol(case), k


@pytest.mark.parametrize("case", CASES)
def test_neighbor_sets_bit_exact(case):
    g = Golden(case)
    with _sim(g) as s:
        counts, lst = s.neighbors()
        assert np.array_equal(counts, g["nbr_counts"])
        assert np.array_equal(lst, g["nbr_list"])


@pytest.mark.parametrize("case", CASES)
def test_first_rhs(case):
    g = Golden(case)
    with _sim(g) as s:
        s.ghost()
        r = s.calculate_RHS()
        for k in r:
            if g.has(f"rhs0_{k}"):
                assert rel_err(r[k], g[f"rhs0_{k}"]) <= case_tol(case), k


@pytest.mark.parametrize("case", [c for c in CASES if not c.startswith("diff")])
def test_one_pst_pass(case):
    g = Golden(case)
    with _sim(g) as s:
        s.ghost()
        s.PST()
        errs = max_rel(s.download(), g, "pst0", fields=("x", "y", "u", "v", "rho", "pre"))
        assert max(errs.values()) <= case_tol(case), errs


@pytest.mark.parametrize("case", CASES)
def test_trajectory_eager(case):
    g = Golden(case)
    with _sim(g) as s:
        s.ghost()
        s.eager_step(1)
        e1 = max_rel(s.download(), g, "step_1")
        assert max(e1.values()) <= case_tol(case), e1
        s.eager_step(g.steps - 1)
        el = max_rel(s.download(), g, "step_last")
        assert max(el.values()) <= case_tol(case), el


@pytest.mark.parametrize("case", CASES)
def test_trajectory_graph_matches_eager_bitwise(case):
    g = Golden(case)
    with _sim(g) as a, _sim(g) as b:
        a.ghost()
        b.ghost()
        a.eager_step(g.steps)
        b.step(g.steps)
        da, db = a.download(), b.download()
        for k in g.fields():
            assert np.array_equal(da[k], db[k]), k
        el = max_rel(db, g, "step_last")
        assert max(el.values()) <= case_tol(case), el


@pytest.mark.parametrize("case", CASES)
def test_per_step_from_reference_state(case):
    g = Golden(case)
    with _sim(g, g.state("step_1")) as s:
        s.eager_step(1)
        e = max_rel(s.download(), g, "step_2")
        assert max(e.values()) <= case_tol(case), e


def test_diffusion_steady_flags():
    g = Golden("diff_rk4")
    with _sim(g) as s:
        s.ghost()
        flags = []
        for _ in range(g.steps):
            s.eager_step(1)
            flags.append(int(s.check_steady_state()[0][0]))
        assert flags == list(g["steady"])


def _ls_q5_cases():
    return [c for c in CASES if Golden(c).cfg.sph_model == 3]


@pytest.mark.parametrize("case", _ls_q5_cases())
def test_mixed_path_tracks_the_reference_trajectory(case):
    """The mixed path (SPH_MODEL 3 only) from the reference's own input
    state, over the reference's run (6 or 10 steps), against its final
    state.  Measured maxima over the nine cases: x, y 4.4e-9; u, v 1.9e-7;
    rho 2.5e-8; pre 3.7e-6 (the EOS amplifies the density's error); tem
    1.1e-10.  The bounds leave at least 13x room.  No particle takes the FP64
    re-solve, so the single-precision sums are what is measured."""
    g = Golden(case)
    with _sim(g, precision=1) as s:
        s.ghost()
        s.step(g.steps)
        e = max_rel(s.download(), g, "step_last")
        fb = s.stats().fp64_fallbacks
    print(f"[gpu] mixed vs reference, {case}: {e} fallbacks={fb}", flush=True)
    bound = {"x": 1e-7, "y": 1e-7, "pre": 5e-5}
    for k, v in e.items():
        assert v <= bound.get(k, 3e-6), (k, e)
    assert fb == 0
