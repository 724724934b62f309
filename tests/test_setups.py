"""Host-side workload builders: block sizes, vertex continuity, and the
stated initial conditions of each workload."""
import numpy as np

import oracle
from knc import setups as S


def _check_chain(blocks):
    for a, b in zip(blocks[:-1], blocks[1:]):
        assert b.index == a.index + 1
        assert a.rv[-1] == b.rv[0]
        assert np.array_equal(a.qv, b.qv)
    for b in blocks:
        assert np.all(np.diff(b.rv) > 0)
        assert b.conserved.shape == (b.nr, b.nq, 4) and b.scalar_mass.shape == (b.nr, b.nq)


def test_bomb_jet_uneven_blocks_cover_the_mesh():
    s = S.bomb_jet(nr=512, nq=32, nblocks=8)
    _check_chain(s.blocks)
    sizes = [b.nr for b in s.blocks]
    assert sum(sizes) == 512 and set(sizes) <= {32, 64, 128} and len(set(sizes)) > 1
    assert s.blocks[0].rv[0] == 1.0 and np.isclose(s.blocks[-1].rv[-1], 100.0, rtol=1e-12)
    # shells: Lorentz factors in [2, 100], p = 1e-2 rho (recovered by the oracle's c2p)
    h = oracle.hydro(system=oracle.SRHD, gamma=4.0 / 3.0)
    for b in s.blocks:
        dv = S.cell_volumes(b.rv, b.qv)
        prim, code, _ = oracle.try_to_primitive(h, (b.conserved / dv[..., None]).reshape(-1, 4))
        assert np.all(code == 0)
        G = np.sqrt(1.0 + prim[:, 1] ** 2 + prim[:, 2] ** 2)
        assert G.min() > 2.0 * (1 - 1e-9) and G.max() < 100.0 * (1 + 1e-9)
        assert np.allclose(prim[:, 3], 1e-2 * prim[:, 0], rtol=1e-9)


def test_wind_shock_chain_continues_past_the_initial_mesh():
    s = S.wind_shock(nq=16, nblocks=10, choices=(8, 16, 32), check_every=7)
    _check_chain(s.blocks)
    assert set(b.nr for b in s.blocks) <= {8, 16, 32}
    assert np.isclose(s.blocks[-1].rv[-1], 1e4, rtol=1e-12)
    rv, cons, sm = s.block_source(10, 0.0)
    assert rv[0] == s.blocks[-1].rv[-1] and rv[-1] > 1e4
    assert cons.shape == (rv.size - 1, 16, 4) and np.all(sm == 0.0)
    again = S.wind_shock(nq=16, nblocks=10, choices=(8, 16, 32), check_every=7).block_source(10, 0.0)
    assert np.array_equal(again[0], rv) and np.array_equal(again[1], cons)
    assert s.excision.check_every == 7 and s.excision.moving


def test_euler_blast_deposits_unit_thermal_energy():
    s = S.euler_blast()
    b = s.blocks[0]
    dv = S.cell_volumes(b.rv, b.qv)
    rc, _ = S.cell_centers(b.rv, b.qv)
    rho = b.conserved[..., 0] / dv
    e_th = b.conserved[..., 3] - 0.5 * (b.conserved[..., 1] ** 2 + b.conserved[..., 2] ** 2) / b.conserved[..., 0]
    hot = rc < 1.5
    background = 1e-6 * rho / (5.0 / 3.0 - 1.0) * dv
    assert np.isclose((e_th - background)[hot].sum(), 1.0, rtol=1e-12)
    assert np.allclose((e_th - background)[~hot], 0.0, atol=1e-18)


def test_jet_in_cloud_inlet_rows():
    s = S.jet_in_cloud(nr=64, nq=32, nblocks=2, corr=2.0)
    qc = 0.5 * (s.blocks[0].qv[:-1] + s.blocks[0].qv[1:])
    prim = s.inner.primitive
    jet = qc < 0.1
    assert np.allclose(prim[:, jet, 1], np.sqrt(99.0)) and np.allclose(prim[:, jet, 0], 1e-3)
    assert np.allclose(prim[:, jet, 3], 1e-2) and np.allclose(prim[:, ~jet, 1], 0.0)
    assert np.allclose(s.inner.scalar[:, jet], 1e-3) and np.allclose(s.inner.scalar[:, ~jet], 0.0)
    assert s.excision.inner_speed == 0.5 * S.LIGHT_SPEED and not s.hydro.adaptive_time_step
