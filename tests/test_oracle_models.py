"""CPU checks that pin the oracle's restatement of the reference's initial
models (oracle/knc_model_oracle.cpp) and block mesh: the reference's own
unit tests of the lookup table (src/lookup_table_v2.rs:138-180), the
defining relations of each model (mass and energy budgets, normalisations,
Poisson's equation for the galactic density) and the mesh's geometric
progression.  Each check is independent of the formula it tests."""
import math

import numpy as np
import pytest
from scipy import integrate

import oracle

C = 3e10
KB = 1
JIC = 2
JIS = 3
HALO = 4
WIND = 5

BOMB = [1e-23, 1e18, 1e18, 1e51, 2e32]                     # kinetic_bomb.yaml
CLOUD = [1e32, 1.0, 1.0, 1.0, 0.1, 30.0, 2e27, 0.1, 0.995, 0.25]   # jet_in_cloud.yaml
STAR = [1e34, 10.0, 1e51, 0.1, 50.0, 1e11, 1e31, 1.0]
HALO_P = [1e20, 1e18, 1e18, 1e51, 2e32, 1e22]              # halo_kilonova.yaml
WIND_P = [1e20, 1e2, 1e-17, 1e-9, 5e8, 1e-2, 1e33, 1e2, 1.0, 0.01]  # wind_shock.yaml


def _eval(kind, params, r, q, t, table=None):
    return oracle.model_eval(kind, params, np.atleast_1d(np.asarray(r, float)),
                             np.broadcast_to(np.asarray(q, float), np.shape(np.atleast_1d(r))), t, table)


def test_lookup_table_known_answers():
    """lookup_table_v2.rs tests: rows (0, .1), (1, .2), (2, .3); sampled at
    0.5, 1.0, 1.5 -> .15, .20, .25; panics at the lower bound, not at the
    upper bound.  Column 1 of the table is the wind's gamma-beta."""
    table = np.array([[0.0, 0.1, 1.0, 2 * C * C], [1.0, 0.2, 1.0, 2 * C * C], [2.0, 0.3, 1.0, 2 * C * C]])
    quiet = WIND_P[:8] + [100.0, 0.01]            # flares long after t = 0
    P, _, bad = _eval(WIND, quiet, [0.5, 1.0, 1.5, 2.0], 1.0, 0.0, table)
    assert bad == -1
    assert np.all(np.abs(P[:3, 1] - [0.15, 0.20, 0.25]) < 1e-10)
    assert abs(P[3, 1] - 0.3) < 1e-10
    for x in (0.0, -1.0, 2.5):
        assert _eval(WIND, quiet, [x], 1.0, 0.0, table)[2] == 0
    # h = 2 c^2: mu = 1, e = 3/4, p = d e (gamma - 1) with gamma = 4/3
    assert abs(P[0, 3] - 0.25) < 1e-15


def test_wind_shock_flares_and_panics():
    """Three flares at flare_time + {0, 3, 6} (wind_shock.rs:81-126): mass
    outflow 4 pi r^2 rho u c = flare_outflow_rate tapered linearly to zero
    over flare_duration, p = 1e-6 rho; between flares without a table the
    reference panics."""
    r = np.array([1e9, 3e9])
    for start in (1.0, 4.0, 7.0):
        t = start + 0.004
        P, X, bad = _eval(WIND, WIND_P, r, 1.0, t)
        assert bad == -1 and np.all(X == 0.0)
        mdot = 4 * math.pi * r * r * P[:, 0] * P[:, 1] * C
        assert np.allclose(mdot, 1e33 * (0.01 - 0.004) / 0.01, rtol=1e-13)
        assert np.allclose(P[:, 3], 1e-6 * P[:, 0], rtol=1e-15)
    assert _eval(WIND, WIND_P, r, 1.0, 2.0)[2] == 0


def test_kinetic_bomb_shell_budget():
    """The shell carries shell_mass at velocity sqrt(2E/M): its density
    integrated over the shell volume gives the mass, and its momentum the
    kinetic energy; outside, rho = rho0 (r/r_launch)^2 and p = 1e-3 rho."""
    M, E = BOMB[4], BOMB[3]
    v = math.sqrt(2 * E / M)
    t = 2e8
    r_out = 1e18 + v * t
    r_in = 1e18 + v * (t - 1e18 / v)
    rs = np.linspace(r_in, r_out, 20001)[:-1]
    P, X, _ = _eval(KB, BOMB, rs, 1.0, t)
    assert np.all(X == 1.0)
    mass = np.trapezoid(P[:, 0] * 4 * math.pi * rs ** 2, rs)
    assert abs(mass / M - 1) < 1e-3
    assert np.allclose(0.5 * mass * (P[:, 1] * C) ** 2, E, rtol=2e-3)
    P2, X2, _ = _eval(KB, BOMB, [3e18], 1.0, t)
    assert X2[0] == 0.0 and P2[0, 1] == 0.0
    assert P2[0, 0] == pytest.approx(1e-23 * 9.0, rel=1e-15)
    assert P2[0, 3] == pytest.approx(1e-3 * P2[0, 0], rel=1e-15)


def test_jet_in_cloud_budgets():
    """Cloud: the mass rate per steradian times 4 pi times the engine delay
    is the cloud mass; jet: Gamma times the mass rate, times 4 pi times the
    engine duration, is E = engine_strength * cloud_mass (jet_in_cloud.rs);
    the envelope beyond the cloud moves with u = b/sqrt(1-b^2), b = r/(ct)."""
    t = 1.5
    r_cloud = 0.05 * C * t
    P, X, _ = _eval(JIC, CLOUD, [r_cloud], 1.0, t)
    f = P[0, 0] * r_cloud ** 2 * P[0, 1] * C
    assert X[0] == 1.0 and abs(4 * math.pi * CLOUD[1] * f / CLOUD[0] - 1) < 1e-14
    v_jet = 30 / math.sqrt(1 + 900) * C
    r_jet = 0.5 * v_jet * (t - 1.0)
    P, X, _ = _eval(JIC, CLOUD, [r_jet], 0.05, t)
    f = P[0, 0] * r_jet ** 2 * P[0, 1] * C
    assert X[0] == 100.0 and P[0, 1] == 30.0
    assert abs(4 * math.pi * CLOUD[2] * f * math.sqrt(1 + 900) / (CLOUD[3] * CLOUD[0]) - 1) < 1e-14
    r_env = 0.5 * C * t
    P, X, _ = _eval(JIC, CLOUD, [r_env], 1.0, t)
    assert X[0] == 0.01 and abs(P[0, 1] - 0.5 / math.sqrt(0.75)) < 1e-15
    assert np.all(P[:, 3] == 1e-3 * P[:, 0])


def test_jet_in_star_nozzle_normalisation():
    """The nozzle function of Duffell & MacFadyen (2015) is normalised to a
    unit volume integral (in units of R0); the jet zone keeps the part within
    engine_theta of either pole, 1 - 1/e of it for a Gaussian in theta."""
    th, u, T, E = STAR[3], STAR[4], STAR[1], STAR[2]
    R0 = 7e10

    def nozzle(x, q):
        r = x * R0
        P, X, _ = _eval(JIS, STAR, [r], q, 1e3)
        rate = P[0, 0] * r * r * u * C
        return rate * math.sqrt(1 + u * u) * C * C * 4 * math.pi * T / E

    cone, _ = integrate.dblquad(lambda q, x: nozzle(x, q) * 2 * math.pi * x * x * math.sin(q), 1e-6, 0.1, 0.0, th,
                                epsabs=1e-12, epsrel=1e-8)
    assert abs(2 * cone / (1 - math.exp(-1)) - 1) < 0.02


def test_halo_density_solves_poisson():
    """Outside the shell the gas density is the thin-disk Miyamoto-Nagai
    density, i.e. the Laplacian of its potential -G M / sqrt(R^2 + (a +
    sqrt(z^2 + b^2))^2) over 4 pi G (galmod.rs), evaluated at
    (radial_distance, z); the hydrostatic pressure falls from 1e-3 with
    altitude; below the midplane the reference panics."""
    M, a, b = 1.538e44, 1.461e22, 1.790e21
    R = HALO_P[5]

    def phi(Rc, z):
        return -M / np.sqrt(Rc * Rc + (a + np.sqrt(z * z + b * b)) ** 2)

    for z_target in (5e20, 2e21):
        r = z_target - HALO_P[0]            # q = 0: z = r + altitude
        assert r > 0
        P, X, bad = _eval(HALO, HALO_P, [r], 0.0, 0.0)
        assert bad == -1 and X[0] == 0.0
        h = 1e-3 * b
        lap = ((phi(R + h, z_target) - 2 * phi(R, z_target) + phi(R - h, z_target)) / h ** 2 +
               (phi(R + h, z_target) - phi(R - h, z_target)) / (2 * h) / R +
               (phi(R, z_target + h) - 2 * phi(R, z_target) + phi(R, z_target - h)) / h ** 2)
        assert abs(P[0, 0] / (lap / (4 * math.pi)) - 1) < 1e-4
        assert 0.0 < P[0, 3] < 1e-3
    P1, _, _ = _eval(HALO, HALO_P, [1e19], 0.0, 0.0)
    P2, _, _ = _eval(HALO, HALO_P, [5e19], 0.0, 0.0)
    assert P2[0, 3] < P1[0, 3]
    assert _eval(HALO, HALO_P, [2e20], math.pi, 0.0)[2] == 0


def test_mesh_blocks_form_a_geometric_progression():
    """Mesh::subgrid_extent: block i spans R (1+b)^i .. R (1+b)^(i+1), b =
    block_size * zone_dlogr; zones are log-spaced inside a block; the 1-D
    polar band is pi/2 -+ zone_dlogr."""
    m = oracle.omesh(3e8, 0, 512, 4)
    b = 4 * math.pi / 512
    rv5, qv = oracle.mesh_block_vertices(m, 5)
    rv6, _ = oracle.mesh_block_vertices(m, 6)
    assert abs(rv5[0] / (3e8 * (1 + b) ** 5) - 1) < 1e-14
    assert abs(rv5[-1] / rv6[0] - 1) < 1e-14
    ratios = rv5[1:] / rv5[:-1]
    assert np.allclose(ratios, (1 + b) ** 0.25, rtol=1e-14)
    assert qv[0] == 0.0 and abs(qv[-1] - math.pi) < 1e-15 and qv.size == 513
    m1 = oracle.omesh(5e8, 32, 1, 32)
    rv, qv = oracle.mesh_block_vertices(m1, 0)
    assert abs(rv[-1] / rv[0] - (1 + 32 / 32) ) < 1e-14
    assert np.allclose(qv, [math.pi / 2 - 1 / 32, math.pi / 2 + 1 / 32], rtol=0, atol=1e-16)


def test_block_from_model_is_volume_integrated():
    """from_model: conserved = to_conserved(P) dV and scalar_mass = D dV X;
    summed over a block, D dV is the model's mass in it."""
    h = oracle.hydro(oracle.SRHD, gamma=4.0 / 3.0)
    m = oracle.omesh(1e18, 512, 1, 512)
    t = 1e8
    rv, qv, cons, sm, bad = oracle.block_from_model(h, m, KB, BOMB, 0, t)
    assert bad == -1
    rc = np.sqrt(rv[:-1] * rv[1:])
    P, X, _ = _eval(KB, BOMB, rc, 0.5 * (qv[0] + qv[1]), t)
    dV = 2 * math.pi * (rv[1:] ** 3 - rv[:-1] ** 3) / 3 * (math.cos(qv[0]) - math.cos(qv[1]))
    W = np.sqrt(1 + P[:, 1] ** 2)
    assert np.allclose(cons[:, 0, 0], P[:, 0] * W * dV, rtol=1e-11)
    assert np.allclose(sm[:, 0], cons[:, 0, 0] * X, rtol=0, atol=0)
