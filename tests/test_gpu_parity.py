"""GPU parity: the HIP path through the C ABI against the CPU oracle.

Bars (rel-L1 per field, sum|gpu-cpu|/sum|cpu|): <= 1e-12 after one step and
<= 1e-9 after 1000 steps.  Per-zone methods are compared element-wise.
"""
import numpy as np
import pytest

import knc
import oracle
from util import gpu_state, max_rel_l1, oracle_from_setup, oracle_hydro, oracle_state, rel_l1

pytestmark = pytest.mark.gpu

ONE_STEP = 1e-12
MANY_STEPS = 1e-9


def _near_midpoint_squares(rng, k):
    """Doubles x whose square root lies within 2^-40 ulp of a rounding
    midpoint -- the hardest inputs to round -- and their neighbours.  A
    midpoint between consecutive doubles is m = t 2^E with t odd and of 54
    bits; RN(m^2) is that close to m^2 when t^2 is congruent to a small c
    modulo 2^(bits of t^2 - 53), so t is a square root of c there, found by
    Hensel lifting (c = 1 mod 8; c < 0 too)."""
    def roots(c, bits):
        ts = {1, 3, 5, 7}
        for j in range(3, bits):
            mod = 1 << (j + 1)
            ts = {t if (t * t - c) % mod == 0 else t + (1 << (j - 1)) for t in ts}
        mod = 1 << bits
        ts = {t % mod for t in ts} | {(-t) % mod for t in ts}
        # the 54-bit representatives t + j 2^bits in [2^53, 2^54)
        return {t + j * mod for t in ts for j in range(4) if (t + j * mod) >> 53 == 1}
    out = []
    for c in list(range(1, 8 * k, 8)) + list(range(-7, -8 * k, -8)):
        for bits in (53, 55):
            for t in roots(c, bits):
                if (t * t).bit_length() - 53 != bits:
                    continue
                e = int(rng.integers(-436, 458))
                x = float(t * t) * 2.0 ** (2 * e) if e >= 0 else float(t * t) / 2.0 ** (-2 * e)
                out += [np.nextafter(x, 0), x, np.nextafter(x, np.inf)]
    return np.array(out)


def test_fp64_arithmetic_rounds_like_host():
    """Device division, the stage's square root, unfused multiply-add and the
    stage's U/dV division (from the correctly rounded 1/dV with one
    correction) round like the host's: over the whole exponent range, at the
    edges of the square root's fast range (2^-767 and DBL_MAX) and of the
    compiler's range scaling (subnormals, 0, inf, NaN), on 2^20 random normal
    inputs, on inputs whose root is next to a rounding midpoint, and on
    divisors and dividends with significands near 1 and near 2."""
    rng = np.random.default_rng(0)
    n = 1 << 20
    a = rng.standard_normal(n) * np.exp(rng.uniform(-300, 300, n) / 10)
    a[: n // 4] = np.ldexp(rng.uniform(1.0, 2.0, n // 4), rng.integers(-1074, 1024, n // 4))
    a[n // 4: n // 2] = np.ldexp(rng.uniform(1.0, 2.0, n // 4), rng.integers(-767, 1024, n // 4))
    edges = np.array([0.0, -0.0, 5e-324, 2.0 ** -1022, 2.0 ** -767, np.nextafter(2.0 ** -767, 0),
                      np.nextafter(2.0 ** -767, 1), 1.0, np.nextafter(1.0, 2), 1.7976931348623157e308,
                      np.nextafter(1.7976931348623157e308, 0), np.inf, -np.inf, np.nan, -1.0, -5e-324])
    a[n // 2: n // 2 + edges.size] = edges
    hard = _near_midpoint_squares(rng, 1000)
    a[n // 2 + 64: n // 2 + 64 + hard.size] = hard
    b = rng.standard_normal(n) * np.exp(rng.uniform(-300, 300, n) / 10)
    # quotients at and next to the edges of the three-operation division's
    # range, |a/b| = 2^-500 and 2^500, for divisors with and without a
    # significand
    qe = np.array([2.0 ** -500, 2.0 ** 500])
    qe = np.concatenate([qe, np.nextafter(qe, 0), np.nextafter(qe, np.inf)])
    qb = np.array([1.0, 3.0, 0.75, 2.0 ** -300, 5.0 * 2.0 ** 299])
    qa = (qe[:, None] * qb[None, :]).ravel()
    qb = np.broadcast_to(qb, (qe.size, qb.size)).ravel()
    k1 = n // 2 + 4096
    a[k1: k1 + 2 * qa.size] = np.concatenate([qa, -qa])
    b[k1: k1 + 2 * qa.size] = np.concatenate([qb, qb])
    # significands of the divisor next to 2 and to 1, and of the dividend next to 2
    m, k0 = 1 << 16, 3 * n // 4
    e = rng.integers(-60, 60, (3, m))
    b[k0: k0 + m] = np.ldexp(2.0 - rng.integers(1, 256, m) * 2.0 ** -52, e[0])
    b[k0 + m: k0 + 2 * m] = np.ldexp(1.0 + rng.integers(0, 256, m) * 2.0 ** -52, e[1])
    a[k0 + 2 * m: k0 + 3 * m] = np.ldexp(2.0 - rng.integers(1, 256, m) * 2.0 ** -52, e[2])
    c = rng.standard_normal(n)
    with np.errstate(all="ignore"):
        div, sq, mad, quot = knc.fp64_probe(a, b, c)
        assert np.array_equal(div, a / b, equal_nan=True)
        ref = a / np.abs(b)
        # every dividend, the divisors in the range of cell volumes, signed zeros included
        q = (np.abs(b) >= 2.0 ** -400) & (np.abs(b) <= 2.0 ** 400)
        assert q.sum() > 3 * n // 4
        assert np.array_equal(quot[q], ref[q], equal_nan=True)
        assert np.array_equal(np.signbit(quot[q]), np.signbit(ref[q]))
        assert np.array_equal(sq, np.sqrt(np.abs(a)), equal_nan=True)
        m = np.isfinite(a * b + c)
        assert np.array_equal(mad[m], (a * b + c)[m])


def test_fp64_fma_probe_rates():
    """The FMA probe behind the stage's VALU roofline: eight independent
    chains at eight waves per SIMD issue faster than one dependent chain at
    two, and neither exceeds the MI355X issue peak (614 G wave64
    instructions/s) by more than clock tolerance."""
    peak, ms = knc.fp64_fma_rate(8, 8, 1024)
    dep, _ = knc.fp64_fma_rate(1, 2, 1024)
    print(f"FP64 FMA issue: {peak / 64 / 1e9:.1f} G/s (8 chains, 8 waves), {dep / 64 / 1e9:.1f} G/s (1 chain, 2 waves)",
          flush=True)
    assert ms > 0 and peak > dep > 0
    assert peak / 64 < 1.2 * 614.4e9


def _random_prims(system, n, rng):
    if system == knc.SRHD:
        G = np.exp(rng.uniform(0, np.log(100), n))
        rho = np.exp(rng.uniform(np.log(1e-8), 0, n))
        T = np.exp(rng.uniform(np.log(1e-6), np.log(10), n))
        u = np.sqrt(G ** 2 - 1)
    else:
        rho = np.exp(rng.uniform(np.log(1e-3), np.log(10), n))
        T = np.exp(rng.uniform(np.log(1e-4), np.log(10), n))
        u = rng.uniform(0, 3, n)
    phi = rng.uniform(0, 2 * np.pi, n)
    return np.stack([rho, u * np.cos(phi), u * np.sin(phi), rho * T], 1)


@pytest.mark.parametrize("system", [knc.EULER, knc.SRHD])
@pytest.mark.parametrize("solver", [knc.HLLE, knc.HLLC])
def test_zone_methods_match_oracle(system, solver):
    rng = np.random.default_rng(7 + system + 2 * solver)
    gamma = 4.0 / 3.0 if system == knc.SRHD else 5.0 / 3.0
    h = knc.Hydro(system, solver, 2, 1, gamma, 1.5, 0.4, 3e10 if system == knc.SRHD else 1.0)
    ho = oracle.KoHydro(system, solver, 2, 1, gamma, 1.5, 0.4, h.light_speed)
    n = 50000
    pl = _random_prims(system, n, rng)
    pr = _random_prims(system, n, rng)
    sl = rng.uniform(0, 1, n) * pl[:, 0]
    sr = rng.uniform(0, 1, n) * pr[:, 0]
    u = knc.to_conserved(h, pl)
    assert np.array_equal(u, oracle.to_conserved(ho, pl))
    p1, c1, i1 = knc.try_to_primitive(h, u)
    p2, c2, i2 = oracle.try_to_primitive(ho, u)
    assert np.array_equal(c1, c2) and np.array_equal(i1, i2)
    assert np.array_equal(p1, p2)
    for axis in (knc.RADIAL, knc.POLAR):
        f1, g1 = knc.intercell_flux(h, pl, pr, sl, sr, axis)
        f2, g2 = oracle.intercell_flux(ho, pl, pr, sl, sr, axis)
        assert np.array_equal(f1, f2) and np.array_equal(g1, g2)
    r = np.exp(rng.uniform(0, np.log(100), n))
    q = rng.uniform(0.01, np.pi - 0.01, n)
    assert np.array_equal(knc.geometrical_source_terms(h, pl, r, q), oracle.geometrical_source_terms(ho, pl, r, q))
    assert np.array_equal(knc.max_signal_speed(h, pl), oracle.max_signal_speed(ho, pl))
    a, b, c = rng.standard_normal((3, n))
    assert np.array_equal(knc.plm_gradient(1.5, a, b, c), oracle.plm_gradient(1.5, a, b, c))


def _run_pair(setup, nsteps, fold, **options):
    ctx = knc.context_from_setup(setup, **options)
    sim = oracle_from_setup(setup)
    done = 0
    while done < nsteps:
        f = min(fold, nsteps - done)
        dt_o = sim.time_step()
        sim.advance(dt_o, f)
        dt_g = ctx.advance(f)
        assert dt_g == dt_o
        done += f
    assert ctx.time == sim.time
    return ctx, sim


def test_plm_gradient_bits_match_oracle_at_zeros_and_edges():
    """The slope limiter bit for bit, zero signs included: every triple of
    signed zeros, tiny and subnormal values, equal neighbours and
    extrema, against the oracle's select form."""
    vals = np.array([0.0, -0.0, 1.0, -1.0, 5e-324, -5e-324, 2.2e-308, -2.2e-308, 1e-300, -1e-300, 3.0, 1e300,
                     -1e300, 0.5])
    a, b, c = (x.ravel() for x in np.meshgrid(vals, vals, vals, indexing="ij"))
    for theta in (1.0, 1.5, 2.0):
        g = knc.plm_gradient(theta, a, b, c)
        go = oracle.plm_gradient(theta, a, b, c)
        assert np.array_equal(g.view(np.uint64), go.view(np.uint64)), theta


def test_euler_blast_one_step():
    s = knc.setups.euler_blast()
    ctx, sim = _run_pair(s, 1, 1)
    err = rel_l1(gpu_state(ctx), oracle_state(sim, s.blocks))
    assert max(err.values()) <= ONE_STEP, err


@pytest.mark.parametrize("mode,passes", [(knc.PER_BLOCK, 1), (knc.BATCHED, 1), (knc.BATCHED, 2)])
def test_euler_blast_1000_steps(mode, passes):
    s = knc.setups.euler_blast()
    ctx, sim = _run_pair(s, 1000, 10, launch_mode=mode, stage_passes=passes)
    err = rel_l1(gpu_state(ctx), oracle_state(sim, s.blocks))
    assert max(err.values()) <= MANY_STEPS, err


def test_jet_in_cloud_one_step_full_size():
    s = knc.setups.jet_in_cloud()
    ctx, sim = _run_pair(s, 1, 1)
    err = rel_l1(gpu_state(ctx), oracle_state(sim, s.blocks))
    assert max(err.values()) <= ONE_STEP, err


@pytest.mark.parametrize("solver,passes", [(knc.HLLE, 1), (knc.HLLC, 1), (knc.HLLC, 2)])
def test_jet_in_cloud_1000_steps_reduced(solver, passes):
    s = knc.setups.jet_in_cloud(nr=256, nq=64, nblocks=4, riemann=solver, corr=4.0)
    ctx, sim = _run_pair(s, 1000, 10, launch_mode=knc.BATCHED, stage_passes=passes)
    err = rel_l1(gpu_state(ctx), oracle_state(sim, s.blocks))
    assert max(err.values()) <= MANY_STEPS, err


def test_bomb_jet_uneven_blocks_per_block_equals_batched():
    """Per-block, batched (graph / no graph) and two-pass stages give the
    same bits."""
    s = knc.setups.bomb_jet(nr=512, nq=128, nblocks=8, nshell=8)
    ctxs = [knc.context_from_setup(s, launch_mode=knc.PER_BLOCK),
            knc.context_from_setup(s, launch_mode=knc.BATCHED, use_graph=True),
            knc.context_from_setup(s, launch_mode=knc.BATCHED, use_graph=False, rows_per_item=5),
            knc.context_from_setup(s, launch_mode=knc.BATCHED, stage_passes=2),
            knc.context_from_setup(s, launch_mode=knc.PER_BLOCK, stage_passes=2, rows_per_item=7)]
    for _ in range(3):
        dts = [c.advance(10) for c in ctxs]
        assert all(d == dts[0] for d in dts)
    ref = gpu_state(ctxs[0])
    for c in ctxs[1:]:
        for (x, m), (y, n) in zip(gpu_state(c), ref):
            assert np.array_equal(x, y) and np.array_equal(m, n)


def test_bomb_jet_one_step_vs_oracle():
    s = knc.setups.bomb_jet(nr=512, nq=256, nblocks=8, nshell=8)
    ctx, sim = _run_pair(s, 2, 1, launch_mode=knc.BATCHED)
    err = rel_l1(gpu_state(ctx), oracle_state(sim, s.blocks))
    assert max(err.values()) <= ONE_STEP, err


@pytest.mark.parametrize("mode", [knc.PER_BLOCK, knc.BATCHED])
def test_negative_density_is_reported_with_position(mode):
    """A planted zone fails the time step's c2p (State::time_step meets it
    first): the dt pass per block, or the first stage's c2p pass, which
    computes dt in the batched runs."""
    s = knc.setups.jet_in_cloud(nr=128, nq=32, nblocks=2, corr=2.0)
    s.blocks[1].conserved[5, 7, 0] = -1.0
    ctx = knc.context_from_setup(s, launch_mode=mode)
    with pytest.raises(knc.HydroError) as e:
        ctx.advance(1)
    assert e.value.kind == "NegativeDensity"
    assert (e.value.block, e.value.i, e.value.j) == (1, 5, 7)
    sim = oracle_from_setup(s)
    with pytest.raises(oracle.OracleError) as eo:
        sim.advance(sim.time_step(), 1)
    assert (eo.value.block, eo.value.i, eo.value.j) == (1, 5, 7)
    assert e.value.r == eo.value.r and e.value.theta == eo.value.theta


@pytest.mark.parametrize("name,rk,guess", [("bomb_jet", 2, knc.GUESS_PLAIN), ("bomb_jet", 2, knc.GUESS_PREVIOUS),
                                            ("bomb_jet", 2, knc.GUESS_LEARNED_FOLD), ("bomb_jet", 3, knc.GUESS_PLAIN),
                                            ("euler_blast", 1, knc.GUESS_PLAIN), ("euler_blast", 3, knc.GUESS_PLAIN)])
def test_dt_in_the_first_stage_equals_the_dt_pass(name, rk, guess):
    """Batched runs compute each fold's dt in the first stage's c2p pass;
    per-block runs in a separate pass over the state (k_dt).  Same dt at
    every fold, same state bit for bit, over five folds of 4 steps of the
    kinetic bomb + jet in 8 uneven blocks and of the Newtonian blast, for
    the guesses and RK orders listed."""
    from knc import train_guess
    s = (knc.setups.bomb_jet(nr=512, nq=128, nblocks=8, nshell=8) if name == "bomb_jet"
         else knc.setups.euler_blast())
    s.hydro.runge_kutta_order = rk
    ctxs = [knc.context_from_setup(s, launch_mode=m, c2p_guess=guess) for m in (knc.PER_BLOCK, knc.BATCHED)]
    if guess == knc.GUESS_LEARNED_FOLD:
        net = train_guess.load()
        for c in ctxs:
            net.attach(c)
    for _ in range(5):
        dts = [c.advance(4) for c in ctxs]
        assert dts[0] == dts[1]
    err = rel_l1(gpu_state(ctxs[1]), gpu_state(ctxs[0]))
    assert max(err.values()) == 0.0, err


@pytest.mark.parametrize("guess", [knc.GUESS_LEARNED_FOLD, knc.GUESS_LEARNED_STEP])
def test_learned_guess_with_dt_in_the_first_stage(guess):
    """A batched run's first stage with the learned guess: the guess pass,
    then the c2p pass that also reduces the CFL minimum; per-block runs use
    the guess pass, the c2p pass and a separate dt pass.  In the two-pass
    form every zone is recovered once per stage either way, so the Newton
    histograms agree bin for bin (the same guesses), and dt and state bit for
    bit."""
    from knc import train_guess
    s = knc.setups.bomb_jet(nr=512, nq=128, nblocks=8, nshell=8)
    net = train_guess.load()
    ctxs = [knc.context_from_setup(s, launch_mode=m, c2p_guess=guess, collect_stats=True, stage_passes=2)
            for m in (knc.PER_BLOCK, knc.BATCHED)]
    for c in ctxs:
        net.attach(c)
    for _ in range(3):
        dts = [c.advance(5) for c in ctxs]
        assert dts[0] == dts[1]
    err = rel_l1(gpu_state(ctxs[1]), gpu_state(ctxs[0]))
    assert max(err.values()) == 0.0, err
    a, b = (c.c2p_stats() for c in ctxs)
    assert np.array_equal(a["hist"], b["hist"]) and a["hist"].sum() > 0
    assert a["fallbacks"] == b["fallbacks"]


@pytest.mark.parametrize("mode,passes", [("per_block", 1), ("per_block", 2), ("batched", 1), ("batched", 2),
                                         ("batched", 3)])
def test_first_failing_stage_is_reported(mode, passes):
    """dt = 1e-12 is beyond the CFL limit next to the inner jet (c = 3e10):
    the second RK stage fails at block 0, but the planted zone fails in the
    first stage, which the reference returns from (scheme.rs:162-166)."""
    s = knc.setups.bomb_jet(nr=256, nq=128, nblocks=2, nshell=4)
    s.blocks[1].conserved[17, 40, 0] = -1.0
    launch = knc.BATCHED if mode == "batched" else knc.PER_BLOCK
    ctx = knc.context_from_setup(s, launch_mode=launch, stage_passes=passes)
    with pytest.raises(knc.HydroError) as e:
        ctx.advance_steps(1e-12, 1)
    sim = oracle_from_setup(s)
    with pytest.raises(oracle.OracleError) as eo:
        sim.advance(1e-12, 1)
    got = (e.value.code, e.value.block, e.value.i, e.value.j, e.value.value, e.value.r, e.value.theta)
    want = (1, eo.value.block, eo.value.i, eo.value.j, eo.value.value, eo.value.r, eo.value.theta)
    assert got == want and want[1:4] == (1, 17, 40)


def test_later_stage_failure_is_reported_at_its_zone():
    """Without the planted zone the same step fails in its second stage, at
    the zone the oracle names."""
    s = knc.setups.bomb_jet(nr=256, nq=128, nblocks=2, nshell=4)
    ctx = knc.context_from_setup(s, launch_mode=knc.BATCHED)
    with pytest.raises(knc.HydroError) as e:
        ctx.advance_steps(1e-12, 1)
    sim = oracle_from_setup(s)
    with pytest.raises(oracle.OracleError) as eo:
        sim.advance(1e-12, 1)
    assert (e.value.code, e.value.block, e.value.i, e.value.j, e.value.value) == \
        (eo.value.code, eo.value.block, eo.value.i, eo.value.j, eo.value.value)


@pytest.mark.parametrize("nq,threads,passes", [(256, 64, 1), (100, 64, 1), (65, 64, 1), (130, 128, 1),
                                                (100, 64, 2), (130, 128, 2), (300, 256, 1), (300, 256, 2),
                                                (700, 256, 1), (700, 512, 1), (1000, 512, 2)])
def test_polar_tiles_match_untiled_and_oracle(nq, threads, passes):
    """Work items narrower than the mesh (halo columns recovered per item,
    a partial last tile) give the same bits as one item spanning every
    column (or 512-column items), and as the oracle."""
    s = knc.setups.bomb_jet(nr=96, nq=nq, nblocks=3, nshell=4)
    tiled = knc.context_from_setup(s, launch_mode=knc.BATCHED, stage_threads=threads, rows_per_item=7,
                                   stage_passes=passes)
    whole = knc.context_from_setup(s, launch_mode=knc.PER_BLOCK, stage_threads=512)
    sim = oracle_from_setup(s)
    for _ in range(2):
        dt = sim.time_step()
        sim.advance(dt, 3)
        assert tiled.advance(3) == dt
        assert whole.advance(3) == dt
    st, sw = gpu_state(tiled), gpu_state(whole)
    for (x, m), (y, n) in zip(st, sw):
        assert np.array_equal(x, y) and np.array_equal(m, n)
    assert max_rel_l1(st, oracle_state(sim, s.blocks)) <= ONE_STEP


@pytest.mark.parametrize("maker", ["euler", "srhd"])
def test_one_dimensional_mesh_matches_oracle(maker):
    """nq = 1: the 1-D update of src/scheme.rs:99-105 (no polar fluxes)."""
    if maker == "euler":
        s = knc.setups.euler_blast(nr=256, nq=1)
    else:
        s = knc.setups.jet_in_cloud(nr=256, nq=1, nblocks=4, corr=2.0)
    ctx, sim = _run_pair(s, 1, 1, launch_mode=knc.BATCHED)
    assert max_rel_l1(gpu_state(ctx), oracle_state(sim, s.blocks)) <= ONE_STEP
    ctx, sim = _run_pair(s, 200, 10, launch_mode=knc.BATCHED)
    assert max_rel_l1(gpu_state(ctx), oracle_state(sim, s.blocks)) <= MANY_STEPS


@pytest.mark.parametrize("order", [1, 3])
def test_runge_kutta_orders_match_oracle(order):
    """RK1 and RK3 (SSP weights 3/4 and 1/3 fused into the last stage)."""
    s = knc.setups.jet_in_cloud(nr=256, nq=64, nblocks=4, riemann=knc.HLLC, corr=4.0)
    s.hydro.runge_kutta_order = order
    ctx, sim = _run_pair(s, 100, 10, launch_mode=knc.BATCHED)
    assert max_rel_l1(gpu_state(ctx), oracle_state(sim, s.blocks)) <= MANY_STEPS


def test_euler_hlle_and_outflow_both_sides_match_oracle():
    """The reference's Euler solver (HLLE) with zero-gradient rows at both
    ends of the mesh."""
    s = knc.setups.euler_blast()
    s.hydro.riemann_solver = knc.HLLE
    s.inner = knc.setups.Boundary(knc.OUTFLOW)
    ctx, sim = _run_pair(s, 300, 10, launch_mode=knc.BATCHED)
    assert max_rel_l1(gpu_state(ctx), oracle_state(sim, s.blocks)) <= MANY_STEPS


def test_c2p_error_codes_and_floor_match_oracle():
    """The error mapping of RelativisticHydro::try_to_primitive
    (src/physics/relativistic_hydro.rs:68-88) on crafted states: D < 0,
    tau < 0, no root (|S| > tau + D), and a recovered pressure below zero
    (floored to 1e-3 rho, reported as NegativePressure)."""
    h = knc.relativistic_hydro()
    ho = oracle.hydro(system=oracle.SRHD, gamma=4.0 / 3.0)
    cold = knc.to_conserved(h, np.array([[1.0, 10.0, 0.0, 1e-12], [1e-3, 3.0, 4.0, 1e-14]]))
    cold[:, 3] *= 1.0 - 1e-9                      # root moves below p = 0
    cons = np.vstack([[[-1.0, 0.0, 0.0, 1.0]],      # NegativeDensity
                      [[1.0, 0.0, 0.0, -1.0]],      # NegativeEnergyDensity
                      [[1.0, 10.0, 0.0, 0.1]],      # no root: S^2 > (tau + D)^2
                      cold])
    p1, c1, i1 = knc.try_to_primitive(h, cons)
    p2, c2, i2 = oracle.try_to_primitive(ho, cons)
    assert list(c1) == [1, 3, 4, 2, 2]
    assert np.array_equal(c1, c2) and np.array_equal(i1, i2)
    floored = c1 == 2
    assert np.array_equal(p1[floored], p2[floored])
    assert np.array_equal(p1[floored, 3], 1e-3 * p1[floored, 0])
