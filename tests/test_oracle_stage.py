"""Stage-level invariants of the oracle (try_advance_rk / advance / time_step).

Mass conservation, equatorial mirror symmetry, the RK weighted averages and
the CFL time step are checked against what the mathematics fixes, using the
oracle's whole-mesh simulation.
"""
import numpy as np
import pytest

import knc.setups as S
import oracle
from util import oracle_from_setup, oracle_hydro, oracle_state


def blast_setup(system=0, nr=64, nq=32, nblocks=2, solver=1, rk=2, ur=0.0):
    """A blast in a uniform medium, symmetric about the equator, with an
    optional uniform radial velocity."""
    gamma = 5.0 / 3.0 if system == 0 else 4.0 / 3.0
    hydro = S.HydroSpec(system, solver, rk, True, gamma, 1.5, 0.4)
    qv = S.polar_vertices(nq)
    rv = S.log_vertices(1.0, 10.0, nr)
    rc, qc = S.cell_centers(rv, qv)
    rho = np.ones((nr, nq))
    p = np.where(rc < 2.0, 1.0, 1e-3) * (1.0 + 0.1 * np.cos(qc - np.pi / 2) ** 2)
    prim = np.stack([rho, ur + 0 * rho, 0 * rho, p], -1)
    X = (rc < 2.0).astype(float) + 0.5
    blocks = S._blocks_from_prim(hydro, rv, qv, [nr // nblocks] * nblocks, prim, X)
    return S.Setup("blast", hydro, blocks, S.Boundary(1), S.Boundary(1))


def totals(state):
    """Sums of D dV, tau dV and the scalar mass over the mesh."""
    return np.array([sum(c[..., 0].sum() for c, m in state), sum(c[..., 3].sum() for c, m in state),
                     sum(m.sum() for c, m in state)])


def edge_flux(h, block, row, c, m):
    """Mass, energy and scalar flux through the boundary face next to `row`:
    with zero-gradient ghosts both PLM face states equal the edge row."""
    g = oracle.geometry(block.rv, block.qv)
    dv = g["dV"][row]
    prim, code, _ = oracle.try_to_primitive(h, c[row] / dv[:, None])
    assert (code == 0).all()
    W = np.sqrt(1 + prim[:, 1] ** 2 + prim[:, 2] ** 2) if h.system == 1 else 1.0
    sc = m[row] / dv / W
    F, gs = oracle.intercell_flux(h, prim, prim, sc, sc, 0)
    area = g["Ar"][0 if row == 0 else -1]
    return np.array([(F[:, 0] * area).sum(), (F[:, 3] * area).sum(), (gs * area).sum()])


@pytest.mark.parametrize("system", [0, 1])
def test_mass_energy_and_scalar_conserved_up_to_boundary_flux(system):
    """Sums of D dV, tau dV (E dV for Euler) and M over the mesh change only
    by dt times the net flux through the two radial boundary faces (mass,
    energy and scalar sources are zero; interior and pole fluxes telescope).
    A wrong energy flux or a nonzero energy source breaks the second."""
    s = blast_setup(system, rk=1, ur=0.1)
    h = oracle_hydro(s.hydro)
    sim = oracle_from_setup(s, 4)
    for _ in range(3):
        st = oracle_state(sim, s.blocks)
        dt = sim.time_step()
        inflow = edge_flux(h, s.blocks[0], 0, *st[0]) - edge_flux(h, s.blocks[-1], s.blocks[-1].nr - 1, *st[-1])
        before = totals(st)
        sim.advance(dt, 1)
        after = totals(oracle_state(sim, s.blocks))
        assert (np.abs(dt * inflow) > 1e-6 * before).all()
        assert np.allclose(after - before, dt * inflow, rtol=1e-9, atol=1e-14 * before)


@pytest.mark.parametrize("system", [0, 1])
def test_equatorial_mirror_symmetry(system):
    s = blast_setup(system)
    sim = oracle_from_setup(s, 4)
    for _ in range(3):
        sim.advance(sim.time_step(), 5)
    for c, m in oracle_state(sim, s.blocks):
        mir = c[:, ::-1, :].copy()
        mir[..., 2] *= -1.0
        # momenta are compared on the momentum scale sqrt(D tau): the mirrored
        # polar grid and cot(theta) agree only to rounding
        scale = np.abs(c).max(axis=(0, 1))
        scale[1:3] = np.sqrt(scale[0] * scale[3])
        assert np.all(np.abs(c - mir) <= 1e-12 * scale)
        assert np.allclose(m, m[:, ::-1], rtol=1e-12, atol=1e-12 * np.abs(m).max())


def test_rk2_is_the_half_average_of_two_euler_stages():
    """RK2 = U0/2 + L(L(U0))/2 exactly (WeightedAverage with b = 1/2)."""
    s1 = blast_setup(rk=1)
    s2 = blast_setup(rk=2)
    a = oracle_from_setup(s1, 2)
    b = oracle_from_setup(s2, 2)
    dt = b.time_step()
    u0 = oracle_state(a, s1.blocks)
    a.advance(dt, 2)
    b.advance(dt, 1)
    ua = oracle_state(a, s1.blocks)
    ub = oracle_state(b, s2.blocks)
    for (c0, m0), (ca, ma), (cb, mb) in zip(u0, ua, ub):
        assert np.array_equal(cb, ca * (-0.5 + 1.0) + c0 * 0.5)
        assert np.array_equal(mb, ma * (-0.5 + 1.0) + m0 * 0.5)
    assert b.time == pytest.approx(dt, rel=1e-15) and b.iteration == 1


def sim_with_state(setup, state, nthreads=2):
    sim = oracle.Sim(oracle_hydro(setup.hydro), nthreads)
    for b, (c, m) in zip(setup.blocks, state):
        sim.push_block(b.rv, b.qv, c, m)
    for side, bd in ((0, setup.inner), (1, setup.outer)):
        sim.set_boundary(side, bd.mode, bd.primitive, bd.scalar)
    return sim


def test_rk3_is_the_shu_osher_combination_of_euler_stages():
    """RK3: s1 = L(U0); s2 = L(s1)(1-3/4) + U0 3/4; U = L(s2)(1-1/3) + U0/3,
    each L one forward-Euler stage with the same dt."""
    s1, s3 = blast_setup(rk=1), blast_setup(rk=3)
    c = oracle_from_setup(s3, 2)
    dt = c.time_step()
    u0 = oracle_state(c, s3.blocks)

    def euler(state):
        sim = sim_with_state(s1, state)
        sim.advance(dt, 1)
        return oracle_state(sim, s1.blocks)

    def avg(a, b0, w):
        return [(x * (-w + 1.0) + y * w, m * (-w + 1.0) + n * w) for (x, m), (y, n) in zip(a, b0)]

    st1 = euler(u0)
    st2 = avg(euler(st1), u0, 0.75)
    st3 = avg(euler(st2), u0, 1.0 / 3.0)
    c.advance(dt, 1)
    for (x, m), (y, n) in zip(oracle_state(c, s3.blocks), st3):
        assert np.array_equal(x, y) and np.array_equal(m, n)


def test_time_step_closed_form_on_uniform_state():
    """Uniform Euler gas at rest: dt = cfl * min(dl) / c_s."""
    nr, nq = 32, 16
    hydro = S.HydroSpec(0, 0, 2, True, 5.0 / 3.0, 1.5, 0.3)
    qv = S.polar_vertices(nq)
    rv = S.log_vertices(1.0, 3.0, nr)
    prim = np.stack([np.full((nr, nq), 2.0), np.zeros((nr, nq)), np.zeros((nr, nq)), np.full((nr, nq), 0.5)], -1)
    blocks = S._blocks_from_prim(hydro, rv, qv, [nr], prim, np.zeros((nr, nq)))
    s = S.Setup("uniform", hydro, blocks, S.Boundary(1), S.Boundary(1))
    sim = oracle_from_setup(s, 1)
    dl = np.minimum(np.diff(rv)[:, None], np.diff(qv)[None, :] * rv[:-1, None])
    cs = np.sqrt(5.0 / 3.0 * 0.5 / 2.0)
    assert sim.time_step() == pytest.approx(0.3 * dl.min() / cs, rel=1e-13)


def test_srhd_fixed_time_step_is_cfl_dr0_over_c():
    s = S.jet_in_cloud(nr=64, nq=16, nblocks=2, corr=2.0)
    sim = oracle_from_setup(s, 1)
    rv = s.blocks[0].rv
    assert sim.time_step() == 0.4 * (rv[1] - rv[0]) / S.LIGHT_SPEED


def test_block_split_does_not_change_the_result():
    """The stage is the same whether the radial range is one block or many
    (halo rows come from the neighbour block, src/scheme.rs:73-77)."""
    one = S.jet_in_cloud(nr=96, nq=24, nblocks=1, corr=3.0)
    three = S.jet_in_cloud(nr=96, nq=24, nblocks=3, corr=3.0)
    a = oracle_from_setup(one, 2)
    b = oracle_from_setup(three, 2)
    for _ in range(3):
        dt = a.time_step()
        assert dt == b.time_step()
        a.advance(dt, 4)
        b.advance(dt, 4)
    ca = oracle_state(a, one.blocks)[0]
    cb = oracle_state(b, three.blocks)
    assert np.array_equal(ca[0], np.concatenate([c for c, _ in cb]))
    assert np.array_equal(ca[1], np.concatenate([m for _, m in cb]))


def test_excision_removes_inner_block():
    s = S.jet_in_cloud(nr=64, nq=16, nblocks=4, corr=2.0)
    sim = oracle_from_setup(s, 1)
    assert sim.num_blocks() == 4
    sim.pop_inner_block()
    assert sim.num_blocks() == 3
    sim.advance(sim.time_step(), 2)


def test_inner_excision_surface_removes_block_when_it_passes():
    """IES(t) = r0 + t v (src/mesh.rs:318-321) passes the inner block's outer
    radius between the 4th and 5th step: the block goes before step 5
    (the check precedes each step, src/scheme.rs:228-232)."""
    s = S.jet_in_cloud(nr=64, nq=16, nblocks=4, corr=2.0)
    sim = oracle_from_setup(s, 1)
    dt = sim.time_step()
    r_out = s.blocks[0].rv[-1]
    ex = S.Excision(inner_radius=1.0, outer_radius=100.0, inner_speed=(r_out - 1.0) / (3.5 * dt))
    dt1, first = sim.scheme_advance(4, ex)
    assert dt1 == dt and first == 0 and sim.num_blocks() == 4
    _, first = sim.scheme_advance(1, ex, first_index=first)
    assert first == 1 and sim.num_blocks() == 3


def test_outer_excision_surface_adds_model_blocks():
    """Each check removes the inner block and adds block outer+1 from the
    model (src/scheme.rs:196-205); the new block continues the log spacing."""
    s = S.wind_shock(nq=16, nblocks=6, choices=(8, 16), check_every=3)
    sim = oracle_from_setup(s, 2)
    first = 0
    for _ in range(3):
        _, first = sim.scheme_advance(3, s.excision, s.block_source, first)
    # checks at iterations 3 and 6 (iteration 0 is at t = 0: nothing to do)
    assert first == 2 and sim.num_blocks() == 6
    rvs = sim.radial_vertices
    for a, b in zip(rvs[:-1], rvs[1:]):
        assert a[-1] == b[0]
    ratio = rvs[-1][1:] / rvs[-1][:-1]
    assert np.allclose(ratio, ratio[0], rtol=1e-12)
    assert np.isclose(ratio[0], s.blocks[0].rv[1] / s.blocks[0].rv[0], rtol=1e-12)
