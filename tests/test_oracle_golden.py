"""The oracle against the compiled reference (golden fixtures, CPU only).

Every fixture was written by the reference Fortran itself (built from its own
sources by oracle/build_ref.sh and driven by oracle/ref_harness.f90), so these
tests pin the oracle's restatement pass by pass: initial ghost fill, marker
values, cell indices, neighbor sets, the first RHS, one PST pass, and the
trajectory after 1, 2 and all steps.  Tolerance: 1e-12 norm-wise relative per
field (the reference's own 1-vs-8-thread spread is ~1e-14), 1e-9 for q = 20
(helpers.CASE_TOL).
"""
import numpy as np
import pytest

from helpers import CASES, Golden, case_tol, max_rel, rel_err
from oracle.pyoracle import OracleSim



def _oracle(g: Golden, st=None):
    o = OracleSim(g.param.to_c(), g.n_int, g.n_ext)
    o.init(st if st is not None else g.input_state())
    return o


@pytest.mark.parametrize("case", CASES)
def test_initial_ghost_and_markers(case):
    g = Golden(case)
    o = _oracle(g)
    o.ghost()
    s = o.get_state()
    errs = max_rel(s, g, "state0")
    assert max(errs.values()) <= case_tol(case), errs
    vm = o.get_vm()
    assert np.array_equal(vm["ptype"], g["vm_ptype"])
    for k in ("x", "y", "u", "v", "pre") + (("tem",) if g.has_tem else ()):
        assert rel_err(vm[k], g[f"vm_{k}"]) <= case_tol(case), k


@pytest.mark.parametrize("case", CASES)
def test_cells_and_neighbor_sets_exact(case):
    g = Golden(case)
    o = _oracle(g)
    assert np.array_equal(o.cell_index(), g["cell0"])
    o.ghost()
    counts, lst = o.neighbors()
    assert np.array_equal(counts, g["nbr_counts"])
    assert np.array_equal(lst, g["nbr_list"])  # same visit (summation) order


@pytest.mark.parametrize("case", CASES)
def test_first_rhs(case):
    g = Golden(case)
    o = _oracle(g)
    o.ghost()
    r = o.calculate_rhs()
    for k in r:
        if g.has(f"rhs0_{k}"):
            assert rel_err(r[k], g[f"rhs0_{k}"]) <= case_tol(case), k


@pytest.mark.parametrize("case", [c for c in CASES if not c.startswith("diff")])
def test_one_pst_pass(case):
    g = Golden(case)
    o = _oracle(g)
    o.ghost()
    o.pst()
    errs = max_rel(o.get_state(), g, "pst0", fields=("x", "y", "u", "v", "rho", "pre"))
    assert max(errs.values()) <= case_tol(case), errs


@pytest.mark.parametrize("case", CASES)
def test_trajectory(case):
    g = Golden(case)
    o = _oracle(g)
    o.ghost()
    o.step(1)
    e1 = max_rel(o.get_state(), g, "step_1")
    assert max(e1.values()) <= case_tol(case), e1
    o.step(g.steps - 1)
    el = max_rel(o.get_state(), g, "step_last")
    assert max(el.values()) <= case_tol(case), el


@pytest.mark.parametrize("case", CASES)
def test_per_step_from_reference_state(case):
    """Restart from the reference's state after step 1 (its wall values
    included, as the reference restart port does) and compare step 2."""
    g = Golden(case)
    o = _oracle(g, g.state("step_1"))
    o.step(1)
    e = max_rel(o.get_state(), g, "step_2")
    assert max(e.values()) <= case_tol(case), e


def test_diffusion_steady_flags():
    g = Golden("diff_rk4")
    o = _oracle(g)
    o.ghost()
    flags = []
    for _ in range(g.steps):
        o.step(1)
        flags.append(o.check_steady_state()[0])
    assert flags == list(g["steady"])


@pytest.mark.parametrize("case", CASES)
def test_params_bit_exact(case):
    """init_param of this package reproduces the reference's derived values."""
    g = Golden(case)
    p, r = g.param, g.ref_params
    assert p.num_wl == r["num_WL"]
    for ours, theirs in (("dx", "dx"), ("sp_mass", "SP_mass"), ("h", "h"), ("h_eff", "h_eff"),
                         ("w_ave", "W_ave"), ("wl_thick", "WL_thick"), ("sound_speed", "sound_speed"),
                         ("kinematic_vis", "kinematic_vis"), ("thermal_dif", "thermal_dif"),
                         ("dt_cfl", "dt_CFL"), ("dt_vis", "dt_vis"), ("dt_th", "dt_th"), ("dt", "dt"),
                         ("tem_ave", "tem_ave"), ("delta_tem", "delta_tem")):
        assert getattr(p, ours) == r[theirs], ours
