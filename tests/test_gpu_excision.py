"""GPU tests of the moving-excision block set and of the autotuner: blocks
removed and born inside knc_advance (add_remove_blocks, src/scheme.rs:
179-206) against the oracle's scheme_advance, per-block against batched
launches across block-set changes, and autotuned shapes against the
default shape (the shape must not change a single bit)."""
import numpy as np
import pytest

import knc
from util import gpu_state, max_rel_l1, oracle_from_setup

pytestmark = pytest.mark.gpu

ONE_STEP = 1e-12
MANY_STEPS = 1e-9


def _advance_pair(setup, calls, fold, **options):
    ctx = knc.context_from_setup(setup, excision=True, **options)
    sim = oracle_from_setup(setup)
    first = setup.blocks[0].index
    for _ in range(calls):
        dt_o, first = sim.scheme_advance(fold, setup.excision, setup.block_source, first)
        assert ctx.advance(fold) == dt_o
        assert ctx.block_indexes() == list(range(first, first + len(sim.radial_vertices)))
    assert ctx.time == sim.time and ctx.iteration == sim.iteration
    return ctx, sim


def test_jet_in_cloud_inner_excision_1000_steps():
    """IES at 0.5c removes the inner blocks as it passes them."""
    s = knc.setups.jet_in_cloud(nr=256, nq=64, nblocks=4, corr=4.0)
    ctx, sim = _advance_pair(s, 100, 10, launch_mode=knc.BATCHED)
    removed, added = ctx.excision_counts()
    assert removed >= 1 and added == 0
    assert len(ctx.block_indexes()) == 4 - removed
    assert max_rel_l1(gpu_state(ctx), sim.blocks()) <= MANY_STEPS


@pytest.mark.parametrize("mode", [knc.PER_BLOCK, knc.BATCHED])
def test_wind_shock_blocks_cycle_every_check(mode):
    """Every 5 steps the innermost block goes and block outer+1 is born from
    the wind model; uneven block sizes from {8, 16, 32, 64}."""
    s = knc.setups.wind_shock(nq=64, nblocks=16, choices=(8, 16, 32, 64), check_every=5)
    ctx, sim = _advance_pair(s, 3, 10, launch_mode=mode)
    assert ctx.excision_counts() == (5, 5)
    assert ctx.block_indexes() == list(range(5, 21))
    assert max_rel_l1(gpu_state(ctx), sim.blocks()) <= ONE_STEP


def test_wind_shock_batched_graph_equals_per_block():
    s = knc.setups.wind_shock(nq=64, nblocks=16, choices=(8, 16, 32, 64), check_every=7)
    a = knc.context_from_setup(s, excision=True, launch_mode=knc.PER_BLOCK)
    b = knc.context_from_setup(s, excision=True, launch_mode=knc.BATCHED, use_graph=True)
    for _ in range(4):
        assert a.advance(10) == b.advance(10)
    assert a.block_indexes() == b.block_indexes()
    for (x, m), (y, n) in zip(gpu_state(a), gpu_state(b)):
        assert np.array_equal(x, y) and np.array_equal(m, n)


def test_autotuned_shape_changes_no_bits():
    s = knc.setups.bomb_jet(nr=512, nq=256, nblocks=8, nshell=8)
    tuned = knc.context_from_setup(s, launch_mode=knc.BATCHED)
    plain = knc.context_from_setup(s, launch_mode=knc.BATCHED)
    r = tuned.autotune(trials=2, force=True)
    assert r["candidates"] >= 4 and not r["from_cache"]
    assert r["stage_threads"] in (64, 128, 256) and 4 <= r["rows_per_item"] <= 64
    assert 0.0 < r["ms_per_stage"] <= r["default_ms_per_stage"] * 1.5
    # tuning leaves the state and time untouched
    assert tuned.time == plain.time == 0.0
    for (x, m), (y, n) in zip(gpu_state(tuned), gpu_state(plain)):
        assert np.array_equal(x, y) and np.array_equal(m, n)
    for _ in range(2):
        assert tuned.advance(5) == plain.advance(5)
    for (x, m), (y, n) in zip(gpu_state(tuned), gpu_state(plain)):
        assert np.array_equal(x, y) and np.array_equal(m, n)
    again = knc.context_from_setup(s, launch_mode=knc.BATCHED).autotune()
    assert again["from_cache"] and (again["stage_threads"], again["rows_per_item"]) == \
        (r["stage_threads"], r["rows_per_item"])
