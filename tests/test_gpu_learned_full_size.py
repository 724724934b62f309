"""The learned c2p guess as the benchmark runs it: the committed network
(knc/guess_net.npz, trained by knc.train_guess on 10^7 pairs, 2.1e6 of them
harvested from jet-in-cloud snapshots) on the kinetic bomb + jet mesh at
4096 x 1024 zones in 32 uneven blocks, evaluated once per RK step, against
the plain guess; and the harvest of training pairs from a jet-in-cloud
snapshot against the oracle's c2p of the same states."""
import numpy as np
import pytest

import knc
import oracle
from knc import mlp, train_guess
from util import gpu_state, oracle_hydro, rel_l1

pytestmark = [pytest.mark.gpu, pytest.mark.slow]


def _mean_iterations(st):
    h = st["hist"]
    return float((np.arange(h.size) * h).sum() / h.sum())


def test_harvested_pairs_match_the_oracle_c2p():
    """mlp.harvest_pairs on a 1024 x 256 jet-in-cloud snapshot after 100
    steps: every zone's (U/dV, p), with p equal to the oracle's c2p of the
    same state bit for bit."""
    print("harvest: jet_in_cloud 1024x256 after 100 steps", flush=True)
    s = knc.setups.jet_in_cloud()
    ctx = knc.context_from_setup(s, excision=True, launch_mode=knc.BATCHED)
    for _ in range(10):
        ctx.advance(10)
    h = knc.hydro_from_spec(s.hydro)
    cons, p = mlp.harvest_pairs(ctx, h)
    zones = sum(ctx.shapes[i][0] * ctx.shapes[i][1] for i in ctx.block_indexes())
    assert cons.shape == (zones, 4)
    P, code, _ = oracle.try_to_primitive(oracle_hydro(s.hydro), cons)
    assert (code == 0).all()
    assert np.array_equal(p, P[:, 3])
    # the trained network's guesses on these states: mostly within a factor of 2
    net = train_guess.load()
    _, guess = net.predict(cons, s.hydro.gamma_law_index)
    err = np.abs(np.log2(guess / p))
    print(f"harvested pairs: {cons.shape[0]}, median |log2(guess/p)| {np.median(err):.3f}", flush=True)
    assert np.median(err) < 1.0


@pytest.mark.parametrize("mode", [knc.GUESS_LEARNED_STEP, knc.GUESS_LEARNED_FOLD])
def test_committed_network_on_the_benchmark_mesh(mode):
    """The learned guess, evaluated once per RK step or once per fold, with
    the committed network against the plain batched path on bomb_jet
    4096 x 1024, 20 steps in folds of 10: the same state bit for bit (bar
    rel-L1 <= 1e-9), fewer Newton evaluations per zone, and no fallback to
    the plain guess when the guesses are the step's own (fewer than one zone
    in 10^4 when they are up to a fold old)."""
    print(f"learned (mode {mode}) vs plain: bomb_jet 4096x1024, 20 steps", flush=True)
    s = knc.setups.bomb_jet()
    net = train_guess.load()
    plain = knc.context_from_setup(s, launch_mode=knc.BATCHED, collect_stats=True)
    learned = knc.context_from_setup(s, launch_mode=knc.BATCHED, collect_stats=True, c2p_guess=mode)
    net.attach(learned)
    for _ in range(2):
        assert plain.advance(10) == learned.advance(10)
    err = rel_l1(gpu_state(learned), gpu_state(plain))
    sp, sl = plain.c2p_stats(), learned.c2p_stats()
    print(f"rel-L1 {err}; mean Newton evaluations plain {_mean_iterations(sp):.3f}, "
          f"learned {_mean_iterations(sl):.3f}; fallbacks {sl['fallbacks']}", flush=True)
    assert max(err.values()) <= 1e-9, err
    assert sl["zones"] == sp["zones"] > 0
    assert sl["fallbacks"] <= (0 if mode == knc.GUESS_LEARNED_STEP else 1e-4 * sl["zones"])
    assert _mean_iterations(sl) < _mean_iterations(sp) - 0.5
