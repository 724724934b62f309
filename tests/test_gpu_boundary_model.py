"""GPU tests of model boundaries: ghost rows evaluated at every stage's
input time (BlockState::from_model on the boundary blocks at state.time,
src/scheme.rs:46-52), against the oracle driven by the same source."""
import numpy as np
import pytest

import knc
from knc import setups as S
from util import gpu_state, max_rel_l1, oracle_from_setup

pytestmark = pytest.mark.gpu

ONE_STEP = 1e-12
MANY_STEPS = 1e-9


def _inlet(setup, ramp_time):
    """A jet inlet whose Lorentz factor ramps from 2 to 10 and whose density
    doubles over ramp_time, in the ghost rows below the inner block; the
    outer ghost rows are the ambient r^-2 medium beyond the outer block."""
    nq = setup.blocks[0].nq
    qv = setup.blocks[0].qv
    qc = 0.5 * (qv[:-1] + qv[1:])
    jet = qc < 0.1
    rv_in = setup.blocks[0].rv
    r_in = rv_in[0] * (rv_in[0] / rv_in[1]) ** np.array([1.5, 0.5])      # centres of the two rows below
    rv_out = setup.blocks[-1].rv
    r_out = rv_out[-1] * (rv_out[-1] / rv_out[-2]) ** np.array([0.5, 1.5])
    seen = []

    def source(side, t, inner, outer):
        seen.append((side, t, inner, outer))
        f = min(1.0, t / ramp_time)
        prim = np.zeros((2, nq, 4))
        scal = np.zeros((2, nq))
        if side == 0:
            rho = r_in[:, None] ** -2.0 * np.ones((1, nq))
            prim[..., 0] = np.where(jet, 1e-3 * (1.0 + f), rho)
            prim[..., 1] = np.where(jet, np.sqrt((2.0 + 8.0 * f) ** 2 - 1.0), 0.0)
            prim[..., 3] = np.where(jet, 1e-2, 1e-4 * rho)
            scal[:] = np.where(jet, prim[..., 0], 0.0)
        else:
            rho = r_out[:, None] ** -2.0 * np.ones((1, nq))
            prim[..., 0] = rho
            prim[..., 3] = 1e-4 * rho
        return prim, scal

    return source, seen


@pytest.mark.parametrize("mode", [knc.PER_BLOCK, knc.BATCHED])
def test_time_dependent_inlet_matches_oracle(mode):
    s = S.jet_in_cloud(nr=256, nq=64, nblocks=4, corr=4.0)
    sim = oracle_from_setup(s)
    ramp = 50 * sim.time_step()
    src_g, seen_g = _inlet(s, ramp)
    src_o, seen_o = _inlet(s, ramp)
    ctx = knc.context_from_setup(s, launch_mode=mode)
    for side in (0, 1):
        ctx.set_boundary_source(side, src_g)
        sim.set_boundary_source(side, lambda sd, t: src_o(sd, t, 0, 3))
    dt = sim.time_step()
    sim.advance(dt, 1)
    assert ctx.advance(1) == dt
    assert max_rel_l1(gpu_state(ctx), sim.blocks()) <= ONE_STEP
    for _ in range(10):
        dt = sim.time_step()
        sim.advance(dt, 10)
        assert ctx.advance(10) == dt
    assert ctx.time == sim.time
    # every stage of every step asked for its rows at the oracle's stage times
    assert [t for _, t, _, _ in seen_g] == [t for _, t, _, _ in seen_o]
    assert len(seen_g) == 2 * 2 * 101
    assert max_rel_l1(gpu_state(ctx), sim.blocks()) <= MANY_STEPS


def test_model_boundary_follows_excision():
    """With the inner excision surface moving, the source is told the
    current inner block index; results match the oracle's scheme_advance."""
    s = S.jet_in_cloud(nr=256, nq=64, nblocks=4, corr=4.0)
    sim = oracle_from_setup(s)
    ramp = 200 * sim.time_step()
    src_g, seen_g = _inlet(s, ramp)
    src_o, _ = _inlet(s, ramp)
    ctx = knc.context_from_setup(s, excision=True, launch_mode=knc.BATCHED)
    ctx.set_boundary_source(0, src_g)
    sim.set_boundary_source(0, lambda sd, t: src_o(sd, t, 0, 3))
    first = 0
    for _ in range(70):
        dt, first = sim.scheme_advance(10, s.excision, None, first)
        assert ctx.advance(10) == dt
    assert first >= 1
    assert ctx.block_indexes()[0] == first
    assert {inner for _, _, inner, _ in seen_g} == set(range(first + 1))
    assert max_rel_l1(gpu_state(ctx), sim.blocks()) <= MANY_STEPS


def test_clearing_the_source_restores_fixed_rows():
    s = S.jet_in_cloud(nr=128, nq=32, nblocks=2, corr=2.0)
    a = knc.context_from_setup(s, launch_mode=knc.BATCHED)
    b = knc.context_from_setup(s, launch_mode=knc.BATCHED)
    src, seen = _inlet(s, 1.0)
    b.set_boundary_source(0, src)
    b.set_boundary_source(0, None)
    for _ in range(3):
        assert a.advance(5) == b.advance(5)
    assert not seen
    for (x, m), (y, n) in zip(gpu_state(a), gpu_state(b)):
        assert np.array_equal(x, y) and np.array_equal(m, n)
