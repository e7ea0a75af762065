"""The autotuner's bookkeeping on the CPU, with a stand-in for the simulation
handle (no GPU): candidate order, rejected settings, a failing handle, and
the choice of the fastest setting."""
import pytest

from lssph import tune
from lssph.engine import LssphError


class FakeSim:
    """Steps cost a time that depends on the current (block, list_len, skin);
    set_tuning rejects block 256 with list_len 64 as lssph_set_tuning does
    beyond the LDS of a block; skin 0.02 fails while stepping."""

    log = []

    def __init__(self, skin, cell_split=0):
        self.skin = skin
        self.cell_split = cell_split
        self.block, self.list_len = 128, 0
        self.ppw = 0
        self.clock = 0.0

    def set_wave_particles(self, ppw):
        if ppw not in (0, 1, 2, 3):
            raise LssphError(-1, "lssph_set_wave_particles: 0, 1, 2 or 3")
        self.ppw = ppw

    def set_tuning(self, block=0, list_len=0, reorder_interval=0):
        if block == 256 and list_len == 64:
            raise LssphError(-1, "list_len * block exceeds 64 KiB of LDS")
        self.block = block or self.block
        self.list_len = list_len or self.list_len
        FakeSim.log.append((self.skin, self.block, self.list_len))

    def step(self, n=1):
        if self.skin == 0.02:
            raise LssphError(-4, "a particle moved farther than half the cell-list skin")
        cost = {64: 1.3, 128: 1.0, 256: 1.1}[self.block] * (1.0 + abs(self.skin - 0.08)) * (
            1.05 if self.list_len == 64 else 1.0) * (0.8 if self.ppw == 2 else 1.0) * (
            {0: 1.0, 1: 1.1, 3: 0.9}[self.cell_split])
        self.clock += cost * n

    def synchronize(self):
        pass

    def close(self):
        pass


@pytest.fixture
def fake_clock(monkeypatch):
    sims = []

    def make(skin, *cell_split):
        s = FakeSim(skin, *cell_split)
        sims.append(s)
        return s

    # time.perf_counter of the tuner reads the fake handle's clock
    monkeypatch.setattr(tune.time, "perf_counter", lambda: sims[-1].clock / 1e3 if sims else 0.0)
    FakeSim.log = []
    return make


def test_autotune_picks_the_fastest_setting(fake_clock):
    r = tune.autotune(fake_clock, steps=3, blocks=(64, 128, 256), list_lens=(64, 0), skins=(0.04, 0.08))
    assert (r.best.block, r.best.list_len, r.best.skin) == (128, 0, 0.08)
    assert r.best.ms_per_step == pytest.approx(1.0)
    assert (r.default.block, r.default.list_len, r.default.skin) == (128, 0, 0.08)
    # list length 0 (keep what lssph_create chose) is tried first on each handle
    for skin in (0.04, 0.08):
        lens = [ln for s, b, ln in FakeSim.log if s == skin]
        assert lens[:3] == [0, 0, 0] and set(lens[3:]) == {64}
    rejected = [s for s in r.table if s.error]
    assert [(s.block, s.list_len) for s in rejected] == [(256, 64), (256, 64)]
    assert len(r.table) == 12
    d = r.as_dict()
    assert d["best"]["list_len"] == "auto" and len(d["candidates"]) == 12


def test_autotune_skips_a_skin_that_fails_and_reports_it(fake_clock):
    r = tune.autotune(fake_clock, steps=2, blocks=(64, 128), skins=(0.02, 0.08))
    failed = [s for s in r.table if s.skin == 0.02]
    assert len(failed) == 1 and "skin" in failed[0].error
    assert r.best.skin == 0.08 and r.best.block == 128


def test_autotune_raises_when_nothing_runs(fake_clock):
    with pytest.raises(RuntimeError):
        tune.autotune(fake_clock, steps=2, blocks=(64,), skins=(0.02,))


def test_autotune_covers_particles_per_wavefront(fake_clock):
    """The SPH_MODEL 6 knob is one more dimension; an invalid value is
    recorded as rejected; the default entry matches on all four settings."""
    r = tune.autotune(fake_clock, steps=2, blocks=(128,), skins=(0.08,), wave_particles=(1, 2, 4),
                      default=(128, 0, 0.08, 1))
    assert r.best.wave_particles == 2
    assert r.default.wave_particles == 1 and r.default.ms_per_step == pytest.approx(1.0)
    bad = [s for s in r.table if s.error]
    assert [s.wave_particles for s in bad] == [4]
    assert r.as_dict()["best"]["wave_particles"] == 2


def test_autotune_covers_cell_splits(fake_clock):
    """The cell columns per stencil reach are fixed before the upload, like
    the skin: each split gets a fresh handle, made with make_sim(skin) for
    the library's default (0) and make_sim(skin, split) otherwise."""
    r = tune.autotune(fake_clock, steps=2, blocks=(128,), skins=(0.08,), splits=(0, 1, 3))
    assert [s.cell_split for s in r.table] == [0, 1, 3]
    assert r.best.cell_split == 3 and r.best.ms_per_step == pytest.approx(0.9)
    assert r.default.cell_split == 0 and r.default.ms_per_step == pytest.approx(1.0)
    assert r.as_dict()["best"]["cell_split"] == 3
