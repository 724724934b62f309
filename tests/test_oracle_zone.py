"""The oracle's per-zone arithmetic against closed forms and invariants.

These pin the restated crate arithmetic (c2p, p2c, wave speeds, sources,
PLM, geometry) independently of the oracle's own code: every check below is
a mathematical identity or special case, not a re-typing of its formula.
"""
import math

import numpy as np
import pytest

import oracle

SRHD = oracle.hydro(oracle.SRHD, gamma=4.0 / 3.0)
EULER = oracle.hydro(oracle.EULER, gamma=5.0 / 3.0)


def random_srhd_prims(n, rng, gmax=100.0, tmin=1e-6, tmax=10.0):
    G = np.exp(rng.uniform(0, np.log(gmax), n))
    rho = np.exp(rng.uniform(np.log(1e-8), 0, n))
    T = np.exp(rng.uniform(np.log(tmin), np.log(tmax), n))
    u = np.sqrt(G ** 2 - 1)
    phi = rng.uniform(0, 2 * np.pi, n)
    return np.stack([rho, u * np.cos(phi), u * np.sin(phi), rho * T], 1)


def srhd_p2c_reference(prim, gamma):
    """Closed form D = rho W, S = rho h W u, tau = rho h W^2 - p - D, written
    with the cancellation-free tau = D u^2/(W+1) + p (gamma/(gamma-1) W^2 - 1)."""
    rho, u1, u2, p = prim.T
    uu = u1 * u1 + u2 * u2
    W = np.sqrt(1 + uu)
    h = 1 + gamma / (gamma - 1) * p / rho
    D = rho * W
    tau = D * uu / (W + 1) + p * (gamma / (gamma - 1) * W * W - 1)
    return np.stack([D, rho * h * W * u1, rho * h * W * u2, tau], 1)


def test_srhd_p2c_matches_closed_form():
    rng = np.random.default_rng(1)
    prim = random_srhd_prims(20000, rng, tmin=1e-3)
    u = oracle.to_conserved(SRHD, prim)
    ref = srhd_p2c_reference(prim, 4.0 / 3.0)
    assert np.allclose(u, ref, rtol=1e-11, atol=0)


def test_srhd_c2p_is_backward_stable():
    """c2p recovers a primitive whose conserved state reproduces the input to
    ~1e-12 relative, over rho in [1e-8,1], p/rho in [1e-6,10], Gamma in [1,100]."""
    rng = np.random.default_rng(2)
    prim = random_srhd_prims(100000, rng)
    u = oracle.to_conserved(SRHD, prim)
    p2, code, iters = oracle.try_to_primitive(SRHD, u)
    assert (code == 0).all()
    assert iters.max() <= 10
    u2 = oracle.to_conserved(SRHD, p2)
    scale = np.abs(u).max(axis=1, keepdims=True)
    assert np.max(np.abs(u2 - u) / scale) < 1e-11


def test_srhd_c2p_recovers_well_conditioned_primitives():
    rng = np.random.default_rng(3)
    prim = random_srhd_prims(50000, rng, gmax=20.0, tmin=1e-2)
    p2, code, _ = oracle.try_to_primitive(SRHD, oracle.to_conserved(SRHD, prim))
    assert (code == 0).all()
    assert np.allclose(p2, prim, rtol=1e-9, atol=0)


def test_srhd_c2p_at_rest_is_gamma_minus_one_tau():
    rng = np.random.default_rng(4)
    D = np.exp(rng.uniform(-10, 0, 1000))
    tau = np.exp(rng.uniform(-10, 2, 1000))
    u = np.stack([D, 0 * D, 0 * D, tau], 1)
    p, code, iters = oracle.try_to_primitive(SRHD, u)
    assert (code == 0).all()
    # exact up to the rounding of tau + D + p - D - p (a few ulp of D + tau)
    assert (np.abs(p[:, 3] - (4.0 / 3.0 - 1.0) * tau) <= 1e-15 * (D + tau)).all()
    assert np.allclose(p[:, 0], D, rtol=5e-16, atol=0)  # D sqrt(Q^2)/Q
    assert (p[:, 1] == 0).all() and (p[:, 2] == 0).all()


def test_euler_c2p_closed_form():
    rng = np.random.default_rng(5)
    rho = np.exp(rng.uniform(-5, 5, 1000))
    v = rng.standard_normal((1000, 2))
    p = np.exp(rng.uniform(-5, 5, 1000))
    prim = np.stack([rho, v[:, 0], v[:, 1], p], 1)
    u = oracle.to_conserved(EULER, prim)
    assert np.allclose(u[:, 3], p / (2.0 / 3.0) + 0.5 * rho * (v ** 2).sum(1), rtol=1e-15)
    p2, code, _ = oracle.try_to_primitive(EULER, u)
    assert (code == 0).all()
    assert np.allclose(p2[:, :3], prim[:, :3], rtol=1e-14)
    assert np.allclose(p2[:, 3], p, rtol=1e-10)


def test_c2p_error_kinds():
    """Error mapping of relativistic_hydro.rs:68-88 / newtonian_hydro.rs:60-65."""
    u = np.array([[-1.0, 0, 0, 1.0],      # D < 0
                  [1.0, 0, 0, -1e-3],     # tau < 0
                  [1.0, 1e6, 0, 1.0],     # |S| >> tau + D: no root
                  ])
    _, code, _ = oracle.try_to_primitive(SRHD, u)
    assert list(code) == [1, 3, 4]
    _, code, _ = oracle.try_to_primitive(EULER, u[:1])
    assert list(code) == [1]


def test_srhd_negative_pressure_is_floored():
    """A state whose pressure root is negative (tau below the cold value for
    its momentum) returns NegativePressure with p = 1e-3 rho."""
    prim = np.array([[1.0, 3.0, 0.0, 1e-12]])
    u = oracle.to_conserved(SRHD, prim)
    u[0, 3] *= 0.999
    p, code, _ = oracle.try_to_primitive(SRHD, u)
    assert code[0] == 2
    assert p[0, 3] == 1e-3 * p[0, 0]


def test_srhd_wave_speeds_reduce_to_velocity_addition():
    """For motion along the normal, Marti-Mueller's outer speeds are the
    relativistic sum (v +- c_s)/(1 +- v c_s); at rest they are +-c_s."""
    g = 4.0 / 3.0
    rng = np.random.default_rng(6)
    n = 2000
    rho = np.exp(rng.uniform(-5, 0, n))
    p = rho * np.exp(rng.uniform(-8, 2, n))
    u = rng.uniform(-50, 50, n)
    for axis in (0, 1):
        prim = np.stack([rho, u if axis == 0 else 0 * u, u if axis == 1 else 0 * u, p], 1)
        lm, lp = oracle.wavespeeds(SRHD, prim, axis)
        cs = np.sqrt(g * p / (rho + g / (g - 1) * p))
        v = u / np.sqrt(1 + u * u)
        assert np.allclose(lm, (v - cs) / (1 - v * cs), rtol=1e-9, atol=1e-12)
        assert np.allclose(lp, (v + cs) / (1 + v * cs), rtol=1e-9, atol=1e-12)
    assert (np.abs(lm) < 1).all() and (np.abs(lp) < 1).all()


def test_max_signal_speed_bounds_wave_speeds():
    rng = np.random.default_rng(7)
    prim = random_srhd_prims(20000, rng)
    m = oracle.max_signal_speed(SRHD, prim) / oracle.LIGHT_SPEED
    for axis in (0, 1):
        lm, lp = oracle.wavespeeds(SRHD, prim, axis)
        assert (np.maximum(np.abs(lm), np.abs(lp)) <= m * (1 + 1e-12)).all()
    assert (m <= 1.0).all()
    pe = np.array([[2.0, 3.0, 4.0, 5.0]])
    assert np.isclose(oracle.max_signal_speed(EULER, pe)[0], 5.0 + np.sqrt(5.0 / 3.0 * 5.0 / 2.0))


def test_geometric_sources_textbook_form():
    """S_r = (2p + rho h u_q^2)/r, S_q = (p cot q - rho h u_r u_q)/r, times c."""
    rng = np.random.default_rng(8)
    prim = random_srhd_prims(1000, rng, tmin=1e-3)
    r = np.exp(rng.uniform(0, 4, 1000))
    q = rng.uniform(0.1, 3.0, 1000)
    s = oracle.geometrical_source_terms(SRHD, prim, r, q) / oracle.LIGHT_SPEED
    rho, ur, uq, p = prim.T
    rhoh = rho + 4.0 * p
    assert (s[:, 0] == 0).all() and (s[:, 3] == 0).all()
    assert np.allclose(s[:, 1], (2 * p + rhoh * uq * uq) / r, rtol=1e-13)
    assert np.allclose(s[:, 2], (p / np.tan(q) - rhoh * ur * uq) / r, rtol=1e-12, atol=1e-14 * np.abs(rhoh * ur * uq / r))


@pytest.mark.parametrize("theta", [1.0, 1.5, 2.0])
def test_plm_brute_force(theta):
    rng = np.random.default_rng(9)
    a, b, c = rng.integers(-3, 4, (3, 5000)).astype(float) * rng.uniform(0.5, 2, (3, 5000))
    g = oracle.plm_gradient(theta, a, b, c)
    cand = np.stack([(b - a) * theta, (c - a) * 0.5, (c - b) * theta])
    ref = np.where((cand > 0).all(0), cand.min(0), np.where((cand < 0).all(0), cand.max(0), 0.0))
    assert np.array_equal(g, ref)
    # monotone data keeps the face values inside the neighbours' range
    mono = (a <= b) & (b <= c)
    assert ((b[mono] + 0.5 * g[mono] <= c[mono] + 1e-12) & (b[mono] - 0.5 * g[mono] >= a[mono] - 1e-12)).all()


def _mesh_rs_face_area(r0, q0, r1, q1):
    """face_area of src/mesh.rs:125-134, operation for operation."""
    s0, s1 = r0 * math.sin(q0), r1 * math.sin(q1)
    z0, z1 = r0 * math.cos(q0), r1 * math.cos(q1)
    ds, dz = s1 - s0, z1 - z0
    return math.pi * (s0 + s1) * math.sqrt(ds * ds + dz * dz)


def test_geometry_is_the_reference_expression():
    """Face areas and volumes are src/mesh.rs:119-134 evaluated zone by zone
    as written there, bit for bit."""
    nr, nq = 12, 9
    rv = 10.0 ** np.linspace(0.3, 1.1, nr + 1)
    qv = 0.1 + (2.9 / nq) * np.arange(nq + 1)
    g = oracle.geometry(rv, qv)
    for i in range(nr + 1):
        for j in range(nq):
            assert g["Ar"][i, j] == _mesh_rs_face_area(rv[i], qv[j], rv[i], qv[j + 1])
    for i in range(nr):
        for j in range(nq + 1):
            assert g["Aq"][i, j] == _mesh_rs_face_area(rv[i], qv[j], rv[i + 1], qv[j])
        for j in range(nq):
            dcost = -(math.cos(qv[j + 1]) - math.cos(qv[j]))
            # powi(3) is r * r * r (repeated multiplication)
            r0, r1 = rv[i], rv[i + 1]
            assert g["dV"][i, j] == 2.0 * math.pi * (r1 * r1 * r1 - r0 * r0 * r0) / 3.0 * dcost


def test_geometry_telescopes_and_closed_forms():
    nr, nq = 40, 24
    rv = 10.0 ** np.linspace(0.0, 1.0, nr + 1)
    qv = 0.0 + (np.pi / nq) * np.arange(nq + 1)
    g = oracle.geometry(rv, qv)
    total = 2 * np.pi / 3 * (rv[-1] ** 3 - rv[0] ** 3) * (np.cos(qv[0]) - np.cos(qv[-1]))
    assert np.isclose(g["dV"].sum(), total, rtol=1e-13)
    # every radial face is the frustum of its chord: pi r^2 (sin q0 + sin q1) 2 sin(dq/2)
    dq = np.diff(qv)
    ref_r = np.pi * rv[:, None] ** 2 * (np.sin(qv[:-1]) + np.sin(qv[1:]))[None, :] * (2 * np.sin(dq / 2))[None, :]
    assert np.max(np.abs(g["Ar"] / ref_r - 1)) <= 1e-13
    # every polar face: pi sin(q) (r1^2 - r0^2) (zero at the pole vertex q = 0)
    ref = np.pi * np.sin(qv)[None, :] * (rv[1:, None] ** 2 - rv[:-1, None] ** 2)
    assert g["Aq"][:, 0].max() == 0.0
    assert np.max(np.abs(g["Aq"][:, 1:] / ref[:, 1:] - 1)) <= 1e-13
    # radial faces over the full polar range sum to 4 pi r^2 (chord areas: O(dq^2))
    assert np.allclose(g["Ar"].sum(1), 4 * np.pi * rv ** 2, rtol=5e-3)
    assert np.allclose(g["rc"], np.sqrt(rv[:-1] * rv[1:])[:, None])
    dq = np.diff(qv)
    assert np.allclose(g["dl"], np.minimum(np.diff(rv)[:, None], dq[None, :] * rv[:-1, None]))


def test_oracle_products_recover_the_model_primitive():
    """BlockProducts (src/products.rs:44-66): a block built as from_model does
    (conserved = to_conserved(P) dV, scalar_mass = D dV X) gives back P in
    AnyPrimitive's array order [velocity_r, velocity_q, mass_density,
    gas_pressure] (src/physics/mod.rs:59-63) and the concentration X; a
    negative density is reported at its zone."""
    rng = np.random.default_rng(11)
    rv = np.geomspace(1.0, 3.0, 6)
    qv = np.linspace(0.0, np.pi, 5)
    dV = oracle.geometry(rv, qv)["dV"]
    for system, gamma in ((oracle.SRHD, 4.0 / 3.0), (oracle.EULER, 5.0 / 3.0)):
        h = oracle.hydro(system, gamma=gamma)
        P = np.stack([10 ** rng.uniform(-3, 0, (5, 4)), rng.uniform(-3, 3, (5, 4)), rng.uniform(-1, 1, (5, 4)),
                      10 ** rng.uniform(-4, -1, (5, 4))], -1)
        X = rng.uniform(0, 1, (5, 4))
        cons = oracle.to_conserved(h, P).reshape(5, 4, 4) * dV[..., None]
        sm = cons[..., 0] * X
        prim, scal = oracle.products(h, rv, qv, cons, sm)
        assert np.allclose(prim, P[..., [1, 2, 0, 3]], rtol=1e-11, atol=0)
        assert np.allclose(scal, X, rtol=1e-15, atol=1e-300)
        cons[3, 2, 0] = -abs(cons[3, 2, 0])
        with pytest.raises(oracle.OracleError) as e:
            oracle.products(h, rv, qv, cons, sm)
        assert (e.value.code, e.value.i, e.value.j) == (1, 3, 2)
