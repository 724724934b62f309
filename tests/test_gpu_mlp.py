"""GPU tests of the learned c2p guess: the training kernels' gradient against
the FP64 oracle, the BF16-MFMA inference against the FP64 forward pass, what
a trained network gains in Newton iterations, and that the learned path
matches the plain path (and falls back to it when the network is useless)."""
import numpy as np
import pytest

import knc
import oracle
from knc import mlp
from util import gpu_state, max_rel_l1

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def trained():
    cons, p, gam = mlp.synthetic_pairs(400_000, seed=11)
    net = mlp.GuessNetwork(seed=3)
    losses = net.train(cons, p, gam, epochs=4, batch=4096, lr=3e-3, seed=5)
    return net, losses


def test_training_gradient_matches_fp64_oracle():
    rng = np.random.default_rng(0)
    net = mlp.GuessNetwork(seed=7)
    params, norm = net.get()
    x = rng.normal(0, 1, (1000, 4)).astype(np.float32)
    t = rng.normal(0, 1, 1000).astype(np.float32)
    loss, grad = net.loss_grad(x, t)
    L64, g64 = oracle.mlp_loss_grad(params.astype(np.float64), x.astype(np.float64), t.astype(np.float64))
    assert abs(loss - L64) <= 1e-5 * abs(L64)
    scale = np.abs(g64).max()
    assert np.max(np.abs(grad - g64)) <= 1e-4 * scale


def test_mfma_inference_matches_fp64_forward():
    rng = np.random.default_rng(1)
    net = mlp.GuessNetwork(seed=9)
    params, _ = net.get()
    params = params + rng.normal(0, 0.05, params.size).astype(np.float32)   # nonzero biases
    cons, p, gam = mlp.synthetic_pairs(5000, seed=2)
    f = mlp.features(cons, gam)
    norm = np.concatenate([f.mean(0), 1 / f.std(0), [0.0, 1.0]]).astype(np.float32)
    net.set(params, norm)
    y, guess = net.predict(cons, gam)
    x = (f - norm[:4]) * norm[4:8]
    ref = oracle.mlp_forward(params.astype(np.float64), x.astype(np.float64))
    err = np.abs(y - ref)
    # BF16 operands with FP32 accumulation: ~3 significant digits
    assert np.median(err) < 2e-2 * np.abs(ref).mean() + 1e-3
    assert np.corrcoef(y, ref)[0, 1] > 0.999
    assert np.allclose(guess, cons[:, 0] * 2.0 ** y.astype(np.float64), rtol=1e-5)


def test_trained_network_reduces_newton_iterations(trained):
    net, losses = trained
    assert losses[-1] < 0.05 * losses[0] or losses[-1] < 0.01
    cons, p, gam = mlp.synthetic_pairs(100_000, seed=99)
    _, guess = net.predict(cons, gam)
    err = np.abs(np.log2(guess / p))
    assert np.median(err) < 1.0            # within a factor 2 for most states
    plain_it, learned_it = [], []
    for g in (4.0 / 3.0, 5.0 / 3.0):
        m = gam == g
        h = knc.relativistic_hydro(gamma_law_index=g)
        P0, c0, i0 = knc.try_to_primitive(h, cons[m])
        P1, c1, i1 = knc.try_to_primitive(h, cons[m], guess[m])
        ok = (c0 == 0) & (c1 == 0)
        assert ok.mean() > 0.999
        # both converge to the same root: p is determined to the rounding
        # noise of the Newton function, ~1e-16 (tau + D + p)
        Q = cons[m][ok, 3] + cons[m][ok, 0] + P0[ok, 3]
        assert np.max(np.abs(P1[ok, 3] - P0[ok, 3]) / Q) < 1e-11
        plain_it.append(i0[ok])
        learned_it.append(i1[ok])
    plain_it, learned_it = np.concatenate(plain_it), np.concatenate(learned_it)
    print(f"mean Newton iterations: plain {plain_it.mean():.3f}, learned {learned_it.mean():.3f}")
    assert learned_it.mean() < plain_it.mean() - 0.3


@pytest.mark.parametrize("mode", [knc.GUESS_LEARNED, knc.GUESS_LEARNED_STEP, knc.GUESS_LEARNED_FOLD])
def test_learned_batched_path_matches_plain(trained, mode):
    net, _ = trained
    s = knc.setups.jet_in_cloud(nr=256, nq=64, nblocks=4, corr=4.0)
    plain = knc.context_from_setup(s, launch_mode=knc.PER_BLOCK)
    learned = knc.context_from_setup(s, launch_mode=knc.BATCHED, c2p_guess=mode, collect_stats=True)
    net.attach(learned)
    for _ in range(20):
        assert plain.advance(10) == learned.advance(10)
    assert max_rel_l1(gpu_state(learned), gpu_state(plain)) <= 1e-9
    st = learned.c2p_stats()
    assert st["zones"] > 0 and st["fallbacks"] <= 1e-3 * st["zones"]


def test_useless_network_falls_back_to_plain_guess():
    s = knc.setups.jet_in_cloud(nr=128, nq=64, nblocks=2, corr=4.0)
    net = mlp.GuessNetwork(seed=1)
    params, norm = net.get()
    net.set(np.full_like(params, np.nan), norm)
    plain = knc.context_from_setup(s, launch_mode=knc.BATCHED)
    learned = knc.context_from_setup(s, launch_mode=knc.BATCHED, c2p_guess=knc.GUESS_LEARNED, collect_stats=True)
    net.attach(learned)
    for _ in range(3):
        assert plain.advance(5) == learned.advance(5)
    for (x, m), (y, n) in zip(gpu_state(learned), gpu_state(plain)):
        assert np.array_equal(x, y) and np.array_equal(m, n)
    st = learned.c2p_stats()
    assert st["fallbacks"] == st["zones"] > 0


def test_learned_mode_requires_a_network():
    s = knc.setups.jet_in_cloud(nr=64, nq=64, nblocks=1, corr=4.0)
    ctx = knc.context_from_setup(s, c2p_guess=knc.GUESS_LEARNED)
    with pytest.raises(RuntimeError, match="no network"):
        ctx.advance(1)


def test_learned_two_pass_stage_equals_fused(trained):
    """The learned guess feeds the c2p pass of the two-pass stage exactly as
    it feeds the fused stage: same bits, and no fallbacks."""
    net, _ = trained
    s = knc.setups.bomb_jet(nr=512, nq=128, nblocks=8, nshell=8)
    fused = knc.context_from_setup(s, launch_mode=knc.BATCHED, c2p_guess=knc.GUESS_LEARNED)
    two = knc.context_from_setup(s, launch_mode=knc.BATCHED, c2p_guess=knc.GUESS_LEARNED, collect_stats=True,
                                 stage_passes=2)
    net.attach(fused)
    net.attach(two)
    for _ in range(3):
        assert fused.advance(5) == two.advance(5)
    for (x, m), (y, n) in zip(gpu_state(fused), gpu_state(two)):
        assert np.array_equal(x, y) and np.array_equal(m, n)
    st = two.c2p_stats()
    # the c2p pass recovers each zone once per stage: 3 x 5 steps x 2 stages
    assert int(st["hist"].sum()) == 30 * s.total_zones
    assert st["fallbacks"] <= 1e-3 * st["zones"]
