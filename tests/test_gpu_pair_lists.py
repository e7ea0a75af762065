"""The neighbor lists the pair kernels actually walk (lssph_pair_lists: the
operator's own scan and walk code, launch shape and on-chip layout) equal
the reference's neighbor sets N(i) = { j != i in the 3x3 stencil, r_ij^2 <=
h_eff^2 } (LSSPH_A_mod.f90:102-124) as sorted sets, for every scan path:
the per-particle scan, the wave-staged scan with positions kept on chip,
the wave-staged scan row by row, and the direct walk of an overflowed list;
periodic edges (image codes), walls and ensembles (waves spanning two
instances) included.  At the initial state the sets are compared with the
oracle's; after steps (positions moved, incremental cell builds) with the
library's own validation walk (lssph_neighbors, itself pinned to the oracle
elsewhere) on the same device state."""
import numpy as np
import pytest

from lssph import workloads
from lssph.engine import Simulation, simulation_from_state
from oracle.pyoracle import OracleSim

pytestmark = pytest.mark.gpu

CAP = 192


def _sets(counts, flat):
    off = np.concatenate([[0], np.cumsum(counts)])
    return [np.sort(flat[off[i]:off[i + 1]]) for i in range(len(counts))]


def _oracle_sets(param, st):
    o = OracleSim(param.to_c(), st.num_int, st.num_ext)
    o.init(st)
    c, lst = o.neighbors()
    return _sets(c, lst)


def _compare(sim, ref_sets):
    counts, lst, path = sim.pair_lists(cap=CAP)
    assert len(ref_sets) == counts.size
    bad = []
    for i, ref in enumerate(ref_sets):
        got = lst[i, :counts[i]]
        if np.any(got < 0) or not np.array_equal(np.sort(got), ref):
            bad.append(i)
    assert not bad, f"{len(bad)} particles differ, first {bad[:5]}"
    return {int(p): int(np.sum(path == p)) for p in np.unique(path)}


def _run(param, st, staged=True, list_len=0, steps=0, split=None):
    print(f"[gpu] pair lists n_int={st.num_int} n_ext={st.num_ext} periodic={param.cfg.periodic} "
          f"staged={staged} list_len={list_len} steps={steps} split={split}", flush=True)
    with simulation_from_state(param, st, precision=1, cell_split=split) as s:
        s.set_scan(staged)
        if list_len:
            s.set_tuning(list_len=list_len)
        if st.num_ext:
            s.ghost()
        if steps:
            s.step(steps)
            ref = _sets(*s.neighbors())
        else:
            ref = _oracle_sets(param, st)
        paths = _compare(s, ref)
    print(f"[gpu]   paths {paths}", flush=True)
    return paths


def test_staged_scan_lists_periodic():
    param, st = workloads.taylor_green_periodic(n=64, re=1000.0, rk=4, jitter=0.2, seed=21)
    paths = _run(param, st)
    assert paths.get(2, 0) > 0 and paths.get(0, 0) > 0  # staged waves and periodic-edge waves


def test_staged_row_by_row_and_overflow_lists():
    """A 30-entry list leaves 248 staging positions per wave: most waves
    stage their candidate rows one at a time, and the densest neighborhoods
    overflow the list and take the direct walk."""
    param, st = workloads.taylor_green_periodic(n=64, re=1000.0, rk=4, jitter=0.2, seed=22)
    paths = _run(param, st, list_len=30)
    assert paths.get(1, 0) > 0 and paths.get(3, 0) > 0


def test_per_particle_scan_lists():
    param, st = workloads.taylor_green_periodic(n=48, re=1000.0, rk=4, jitter=0.2, seed=23)
    paths = _run(param, st, staged=False)
    assert set(paths) == {0}


def test_overflow_walk_lists():
    param, st = workloads.taylor_green_periodic(n=40, re=1000.0, rk=4, jitter=0.2, seed=24)
    paths = _run(param, st, list_len=12)
    assert paths.get(3, 0) > 0.9 * st.num_int


def test_closed_box_lists_with_walls():
    param, st = workloads.cavity(n=40, re=400.0, jitter=0.1, seed=25)
    _run(param, st)


def test_lists_after_steps_and_incremental_builds():
    param, st = workloads.taylor_green_periodic(n=64, re=100.0, rk=4, jitter=0.2, seed=26)
    _run(param, st, steps=3)
    _run(param, st, steps=3, list_len=30)
    param, st = workloads.cavity(n=32, re=400.0, jitter=0.1, seed=27)
    _run(param, st, steps=4)


def test_ensemble_lists():
    """Three cavities in one handle: waves that span two instances scan per
    particle; upload indices are those of the ensemble's upload order."""
    params, states, _ = workloads.cavity_ensemble(count=3, sizes=(16, 24, 20), seed=8, precision=1)
    ni = [s.num_int for s in states]
    ne = [s.num_ext for s in states]
    ref = []
    f_off = np.concatenate([[0], np.cumsum(ni)])
    w_off = sum(ni) + np.concatenate([[0], np.cumsum(ne)])
    for k, (p, st) in enumerate(zip(params, states)):
        for s in _oracle_sets(p, st):
            ref.append(np.sort(np.where(s < st.num_int, f_off[k] + s, w_off[k] + s - st.num_int)))
    print(f"[gpu] pair lists of a 3-cavity ensemble {ni}", flush=True)
    with Simulation(params, ni, ne, precision=1) as sim:
        sim.upload(states)
        sim.init_cell()
        sim.ghost()
        paths = _compare(sim, ref)
    assert paths.get(0, 0) > 0


@pytest.mark.parametrize("split", [1, 3, 4])
def test_lists_with_other_cell_splits(split):
    """The default cells are half the stencil reach wide in x; narrower and
    square cells give the same sets on every scan path, periodic edges,
    walls and incremental builds included."""
    param, st = workloads.taylor_green_periodic(n=64, re=1000.0, rk=4, jitter=0.2, seed=28)
    paths = _run(param, st, split=split)
    assert paths.get(2, 0) > 0 and paths.get(0, 0) > 0
    _run(param, st, list_len=30, split=split, steps=2)
    param, st = workloads.cavity(n=32, re=400.0, jitter=0.1, seed=29)
    _run(param, st, split=split, steps=2)
