"""GPU tests of KNC_GUESS_PREVIOUS: every stage starts Newton from the
pressure its zone's previous stage recovered.  Newton ends in a
guess-independent polish, so the state is the plain guess's to the last bit
-- fused, two-pass and per-block, across excision births (a new block has no
guess yet and starts from p = 0) -- while Newton takes fewer iterations."""
import numpy as np
import pytest

import knc
from util import gpu_state

pytestmark = pytest.mark.gpu


def _same(a, b):
    for (x, m), (y, n) in zip(gpu_state(a), gpu_state(b)):
        if not (np.array_equal(x, y) and np.array_equal(m, n)):
            return False
    return True


def _mean_iterations(st):
    h = st["hist"]
    return float((np.arange(h.size) * h).sum() / h.sum())


@pytest.mark.parametrize("form", ["batched", "two_pass", "per_block"])
def test_previous_guess_gives_the_plain_state(form):
    print(f"previous guess vs plain: bomb_jet 512x128, {form}", flush=True)
    s = knc.setups.bomb_jet(nr=512, nq=128, nblocks=8, nshell=8)
    opts = {"batched": dict(launch_mode=knc.BATCHED), "two_pass": dict(launch_mode=knc.BATCHED, stage_passes=2),
            "per_block": dict(launch_mode=knc.PER_BLOCK)}[form]
    plain = knc.context_from_setup(s, collect_stats=True, **opts)
    prev = knc.context_from_setup(s, collect_stats=True, c2p_guess=knc.GUESS_PREVIOUS, **opts)
    for _ in range(10):
        assert plain.advance(5) == prev.advance(5)
    assert _same(plain, prev)
    sp, sv = plain.c2p_stats(), prev.c2p_stats()
    assert sp["zones"] == sv["zones"] > 0
    assert sv["fallbacks"] == 0
    print(f"mean Newton evaluations per zone: plain {_mean_iterations(sp):.3f}, "
          f"previous {_mean_iterations(sv):.3f}", flush=True)
    assert _mean_iterations(sv) < _mean_iterations(sp) - 0.8


def test_previous_guess_across_births_and_removals():
    print("previous guess vs plain: wind_shock blocks cycling every 5 steps", flush=True)
    s = knc.setups.wind_shock(nq=64, nblocks=16, choices=(8, 16, 32, 64), check_every=5)
    plain = knc.context_from_setup(s, excision=True, launch_mode=knc.BATCHED)
    prev = knc.context_from_setup(s, excision=True, launch_mode=knc.BATCHED, c2p_guess=knc.GUESS_PREVIOUS,
                                  collect_stats=True)
    for _ in range(3):
        assert plain.advance(10) == prev.advance(10)
    assert prev.excision_counts() == (5, 5)
    assert plain.block_indexes() == prev.block_indexes()
    assert _same(plain, prev)
    assert prev.c2p_stats()["fallbacks"] == 0


def test_fold_learned_guess_across_births_and_removals():
    """KNC_GUESS_LEARNED_FOLD evaluates the network once per run of steps
    over one block set: with blocks born and removed every 5 steps each run
    is re-guessed, born blocks included, and the state is the plain one."""
    from knc import train_guess
    print("learned-per-fold guess vs plain: wind_shock blocks cycling every 5 steps", flush=True)
    s = knc.setups.wind_shock(nq=64, nblocks=16, choices=(8, 16, 32, 64), check_every=5)
    plain = knc.context_from_setup(s, excision=True, launch_mode=knc.BATCHED)
    fold = knc.context_from_setup(s, excision=True, launch_mode=knc.BATCHED, c2p_guess=knc.GUESS_LEARNED_FOLD,
                                  collect_stats=True)
    train_guess.load().attach(fold)
    for _ in range(3):
        assert plain.advance(10) == fold.advance(10)
    assert fold.excision_counts() == (5, 5)
    assert plain.block_indexes() == fold.block_indexes()
    assert _same(plain, fold)
    st = fold.c2p_stats()
    assert st["fallbacks"] <= 1e-3 * st["zones"]


def test_previous_guess_is_plain_for_euler():
    """Euler's c2p has no Newton solve: the mode changes nothing."""
    s = knc.setups.euler_blast(nr=128, nq=32)
    a = knc.context_from_setup(s, launch_mode=knc.BATCHED)
    b = knc.context_from_setup(s, launch_mode=knc.BATCHED, c2p_guess=knc.GUESS_PREVIOUS)
    for _ in range(3):
        assert a.advance(5) == b.advance(5)
    assert _same(a, b)


def test_previous_guess_jet_in_cloud_1000_steps_full_size():
    """The jet in cloud at 1024 x 256 zones in 8 blocks with its inner
    excision surface, 1000 steps: the previous-pressure guess gives the plain
    state bit for bit."""
    print("previous guess vs plain: jet_in_cloud 1024x256, 1000 steps", flush=True)
    s = knc.setups.jet_in_cloud()
    plain = knc.context_from_setup(s, excision=True, launch_mode=knc.BATCHED)
    prev = knc.context_from_setup(s, excision=True, launch_mode=knc.BATCHED, c2p_guess=knc.GUESS_PREVIOUS,
                                  collect_stats=True)
    for _ in range(100):
        assert plain.advance(10) == prev.advance(10)
    assert plain.block_indexes() == prev.block_indexes()
    assert plain.excision_counts()[0] >= 1
    assert _same(plain, prev)
    st = prev.c2p_stats()
    assert st["fallbacks"] == 0


def test_previous_guess_runs_are_deterministic():
    """Each stage reads the guess plane of its input slot and writes that of
    its output slot, so no zone's guess depends on the order in which
    workgroups run: two runs of the jet in cloud (1024 x 256, 100 steps with
    its excision), batched and per-block, give the same state and the same
    Newton iteration histogram, which follows the guesses zone by zone."""
    print("previous guess: jet_in_cloud 1024x256, 100 steps, run twice batched and once per block", flush=True)
    s = knc.setups.jet_in_cloud()
    runs = []
    for mode in (knc.BATCHED, knc.BATCHED, knc.PER_BLOCK):
        ctx = knc.context_from_setup(s, excision=True, launch_mode=mode, c2p_guess=knc.GUESS_PREVIOUS,
                                     collect_stats=True)
        for _ in range(10):
            ctx.advance(10)
        runs.append((gpu_state(ctx), ctx.c2p_stats()["hist"].copy(), ctx.block_indexes()))
        ctx.close()
    (s0, h0, b0) = runs[0]
    for (st, h, b) in runs[1:]:
        assert b == b0
        assert np.array_equal(h, h0)
        for (x, m), (y, n) in zip(st, s0):
            assert np.array_equal(x, y) and np.array_equal(m, n)
