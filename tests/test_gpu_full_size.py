"""GPU parity at the benchmark sizes: the jet in cloud at 1024 x 256 zones
for 1000 steps with its moving inner excision surface, the kinetic bomb +
jet at 4096 x 1024 zones, and the wind shock over 64 uneven blocks with
excision every 50 steps, each against the oracle in the same process.
Bars: rel-L1 per field <= 1e-12 after one step, <= 1e-9 after many."""
import os

import pytest

import knc
from util import gpu_state, max_rel_l1, oracle_from_setup, rel_l1

pytestmark = [pytest.mark.gpu, pytest.mark.slow]

ONE_STEP = 1e-12
MANY_STEPS = 1e-9
THREADS = min(16, os.cpu_count() or 1)


def _pair(setup, folds, **options):
    """Advances a batched GPU context and the oracle by the given folds,
    both through scheme::advance with the setup's excision surfaces."""
    ctx = knc.context_from_setup(setup, excision=True, launch_mode=knc.BATCHED, **options)
    sim = oracle_from_setup(setup, THREADS)
    first = setup.blocks[0].index
    for fold in folds:
        dt, first = sim.scheme_advance(fold, setup.excision, setup.block_source, first)
        assert ctx.advance(fold) == dt
        assert ctx.block_indexes() == list(range(first, first + len(sim.radial_vertices)))
    assert ctx.time == sim.time and ctx.iteration == sim.iteration
    return ctx, sim


def test_jet_in_cloud_1000_steps_full_size():
    s = knc.setups.jet_in_cloud()
    ctx, sim = _pair(s, [10] * 100)
    removed, _ = ctx.excision_counts()
    assert removed >= 1
    err = rel_l1(gpu_state(ctx), sim.blocks())
    assert max(err.values()) <= MANY_STEPS, err


def test_bomb_jet_full_size_one_step_and_twenty():
    s = knc.setups.bomb_jet()
    ctx, sim = _pair(s, [1])
    err = rel_l1(gpu_state(ctx), sim.blocks())
    assert max(err.values()) <= ONE_STEP, err
    for _ in range(2):
        dt = sim.time_step()
        sim.advance(dt, 10)
        assert ctx.advance(10) == dt
    assert max_rel_l1(gpu_state(ctx), sim.blocks()) <= MANY_STEPS


def test_wind_shock_full_size_excision_every_50_steps():
    s = knc.setups.wind_shock()
    assert len(s.blocks) == 64
    ctx, sim = _pair(s, [50, 10])
    assert ctx.excision_counts() == (1, 1)
    err = rel_l1(gpu_state(ctx), sim.blocks())
    assert max(err.values()) <= MANY_STEPS, err


def test_jet_in_cloud_hllc_1000_steps_full_size():
    """The same mesh with HLLC, the solver of every reference setup
    (setups/*.yaml), for 1000 steps with the inner excision surface."""
    s = knc.setups.jet_in_cloud(riemann=knc.HLLC)
    ctx, sim = _pair(s, [10] * 100)
    err = rel_l1(gpu_state(ctx), sim.blocks())
    assert max(err.values()) <= MANY_STEPS, err
