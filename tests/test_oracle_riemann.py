"""The oracle's HLLE / HLLC fluxes against invariants and exact Riemann
solutions (Sod: Toro, Riemann Solvers 3rd ed., sec. 4.3 test 1; SRHD blast
wave 1: Marti & Mueller, Living Rev. Relativ. 6 (2003) 7, sec. 6.2)."""
import json
import os

import numpy as np
import pytest

import oracle

GOLDEN = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "known_answers.json")))


def physical_flux(system, gamma, prim, s, axis, c):
    """Independent numpy flux: F_D = D v_n, F_S = S v_n + p n, F_tau = (tau+p) v_n."""
    rho, u1, u2, p = prim.T
    un = u1 if axis == 0 else u2
    if system == oracle.SRHD:
        W = np.sqrt(1 + u1 * u1 + u2 * u2)
        h = 1 + gamma / (gamma - 1) * p / rho
        D, S1, S2 = rho * W, rho * h * W * u1, rho * h * W * u2
        tau = rho * h * W * W - p - D
        vn = un / W
        us = s * W
    else:
        D, S1, S2 = rho, rho * u1, rho * u2
        tau = p / (gamma - 1) + 0.5 * rho * (u1 * u1 + u2 * u2)
        vn = un
        us = s
    F = np.stack([D * vn, S1 * vn + (p if axis == 0 else 0), S2 * vn + (p if axis == 1 else 0), (tau + p) * vn], 1)
    return F * c, us * vn * c


def cases():
    out = []
    for system, gamma in ((oracle.EULER, 5.0 / 3.0), (oracle.SRHD, 4.0 / 3.0)):
        for solver in (oracle.HLLE, oracle.HLLC):
            out.append((system, gamma, solver))
    return out


def rand_prims(system, n, rng):
    rho = np.exp(rng.uniform(-3, 1, n))
    p = rho * np.exp(rng.uniform(-6, 1, n))
    u = rng.uniform(-3, 3, (n, 2)) if system == oracle.SRHD else rng.uniform(-2, 2, (n, 2))
    return np.stack([rho, u[:, 0], u[:, 1], p], 1)


@pytest.mark.parametrize("system,gamma,solver", cases())
def test_consistency(system, gamma, solver):
    rng = np.random.default_rng(10)
    h = oracle.hydro(system, gamma=gamma, riemann=solver)
    P = rand_prims(system, 5000, rng)
    s = rng.uniform(0, 1, 5000)
    for axis in (0, 1):
        F, g = oracle.intercell_flux(h, P, P, s, s, axis)
        Fr, gr = physical_flux(system, gamma, P, s, axis, h.light_speed)
        scale = np.abs(Fr).max(1, keepdims=True)
        assert np.max(np.abs(F - Fr) / scale) < 1e-12
        assert np.allclose(g, gr, rtol=1e-12, atol=1e-12 * np.abs(gr).max())


@pytest.mark.parametrize("system,gamma,solver", cases())
def test_mirror_symmetry(system, gamma, solver):
    rng = np.random.default_rng(11)
    h = oracle.hydro(system, gamma=gamma, riemann=solver)
    L, R = rand_prims(system, 5000, rng), rand_prims(system, 5000, rng)
    sl, sr = rng.uniform(0, 1, (2, 5000))
    for axis in (0, 1):
        n = 1 + axis
        Lm, Rm = R.copy(), L.copy()
        Lm[:, n] *= -1
        Rm[:, n] *= -1
        F, g = oracle.intercell_flux(h, L, R, sl, sr, axis)
        Fm, gm = oracle.intercell_flux(h, Lm, Rm, sr, sl, axis)
        sign = np.array([-1.0, -1.0, -1.0, -1.0])
        sign[n] = 1.0
        scale = np.abs(F).max(1, keepdims=True) + 1e-300
        assert np.max(np.abs(Fm - sign * F) / scale) < 1e-10
        assert np.allclose(gm, -g, rtol=1e-10, atol=1e-10 * np.abs(g).max())


@pytest.mark.parametrize("system,gamma,solver", cases())
def test_supersonic_upwinding(system, gamma, solver):
    h = oracle.hydro(system, gamma=gamma, riemann=solver)
    L = np.array([[1.0, 20.0, 0.3, 0.01]])
    R = np.array([[0.5, 15.0, -0.2, 0.02]])
    F, g = oracle.intercell_flux(h, L, R, 0.3, 0.7, 0)
    FL, gL = physical_flux(system, gamma, L, np.array([0.3]), 0, h.light_speed)
    assert np.allclose(F, FL, rtol=1e-13)
    assert np.allclose(g, gL, rtol=1e-13)


@pytest.mark.parametrize("system,gamma", [(oracle.EULER, 1.4), (oracle.SRHD, 4.0 / 3.0)])
def test_hllc_keeps_stationary_contact(system, gamma):
    h = oracle.hydro(system, gamma=gamma, riemann=oracle.HLLC)
    L = np.array([[1.0, 0.0, 0.0, 0.5]])
    R = np.array([[0.1, 0.0, 0.0, 0.5]])
    F, g = oracle.intercell_flux(h, L, R, 1.0, 0.0, 0)
    assert F[0, 0] == 0.0 and F[0, 3] == 0.0 and g[0] == 0.0
    assert np.isclose(F[0, 1], 0.5 * h.light_speed, rtol=1e-15)
    he = oracle.hydro(system, gamma=gamma, riemann=oracle.HLLE)
    Fe, _ = oracle.intercell_flux(he, L, R, 1.0, 0.0, 0)
    assert Fe[0, 0] != 0.0  # HLLE diffuses the contact


def godunov_1d(h, left, right, n, t_end, cfl=0.4):
    """First-order Godunov on x in [0,1] with the oracle's flux; returns
    cell centres and primitives at t_end."""
    x = (np.arange(n) + 0.5) / n
    prim = np.where((x < 0.5)[:, None], np.array(left), np.array(right)).astype(float)
    cons = oracle.to_conserved(h, prim)
    dx = 1.0 / n
    t = 0.0
    c = h.light_speed
    while t < t_end:
        prim, code, _ = oracle.try_to_primitive(h, cons)
        assert (code == 0).all()
        lam = oracle.max_signal_speed(h, prim)
        dt = min(cfl * dx / lam.max(), t_end - t)
        pe = np.vstack([prim[:1], prim, prim[-1:]])
        F, _ = oracle.intercell_flux(h, pe[:-1], pe[1:], 0.0, 0.0, 0)
        cons = cons - dt / dx * (F[1:] - F[:-1]) / c
        t += dt
    prim, code, _ = oracle.try_to_primitive(h, cons)
    return x, prim


def sod_exact(x, t, gamma, left, right):
    """Exact Riemann solution of the Euler equations (Toro ch. 4)."""
    (rl, ul, pl), (rr, ur, pr) = left, right
    cl, cr = np.sqrt(gamma * pl / rl), np.sqrt(gamma * pr / rr)

    def fk(p, rk, pk, ck):
        if p > pk:
            A, B = 2 / ((gamma + 1) * rk), (gamma - 1) / (gamma + 1) * pk
            return (p - pk) * np.sqrt(A / (p + B))
        return 2 * ck / (gamma - 1) * ((p / pk) ** ((gamma - 1) / (2 * gamma)) - 1)

    lo, hi = 1e-12, 10 * max(pl, pr)
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if fk(mid, rl, pl, cl) + fk(mid, rr, pr, cr) + (ur - ul) > 0:
            hi = mid
        else:
            lo = mid
    ps = 0.5 * (lo + hi)
    us = 0.5 * (ul + ur) + 0.5 * (fk(ps, rr, pr, cr) - fk(ps, rl, pl, cl))
    g = gamma
    out = np.zeros((x.size, 3))
    for k, xi in enumerate(x):
        s = (xi - 0.5) / t
        if s < us:
            # left rarefaction (ps < pl in Sod)
            csl = cl * (ps / pl) ** ((g - 1) / (2 * g))
            if s < ul - cl:
                out[k] = (rl, ul, pl)
            elif s > us - csl:
                out[k] = (rl * (ps / pl) ** (1 / g), us, ps)
            else:
                u = 2 / (g + 1) * (cl + (g - 1) / 2 * ul + s)
                c = 2 / (g + 1) * (cl + (g - 1) / 2 * (ul - s))
                out[k] = (rl * (c / cl) ** (2 / (g - 1)), u, pl * (c / cl) ** (2 * g / (g - 1)))
        else:
            # right shock
            sr = ur + cr * np.sqrt((g + 1) / (2 * g) * ps / pr + (g - 1) / (2 * g))
            if s > sr:
                out[k] = (rr, ur, pr)
            else:
                rs = rr * (ps / pr + (g - 1) / (g + 1)) / ((g - 1) / (g + 1) * ps / pr + 1)
                out[k] = (rs, us, ps)
    return out, ps, us


@pytest.mark.parametrize("solver", [oracle.HLLE, oracle.HLLC])
def test_sod_shock_tube(solver):
    kat = GOLDEN["sod_toro_test1"]
    h = oracle.hydro(oracle.EULER, gamma=kat["gamma"], riemann=solver)
    L, R = kat["left"], kat["right"]
    x, P = godunov_1d(h, [L[0], L[1], 0.0, L[2]], [R[0], R[1], 0.0, R[2]], 800, kat["t"])
    ex, ps, us = sod_exact(x, kat["t"], kat["gamma"], L, R)
    # the exact solver reproduces the published star state
    assert abs(ps - kat["p_star"]) < 1e-5 and abs(us - kat["u_star"]) < 1e-5
    assert np.mean(np.abs(P[:, 0] - ex[:, 0])) < 1.5e-2
    star = (x > 0.55) & (x < 0.8)
    assert np.allclose(P[star, 3], kat["p_star"], rtol=1e-2)
    assert np.allclose(P[star, 1], kat["u_star"], rtol=1e-2)


@pytest.mark.parametrize("solver", [oracle.HLLE, oracle.HLLC])
def test_srhd_blast_wave_1(solver):
    kat = GOLDEN["srhd_blast_wave_1"]
    h = oracle.hydro(oracle.SRHD, gamma=kat["gamma"], riemann=solver, light_speed=1.0)
    L, R = kat["left"], kat["right"]
    x, P = godunov_1d(h, [L[0], L[1], 0.0, L[2]], [R[0], R[1], 0.0, R[2]], 1600, kat["t"])
    plateau = (x > 0.72) & (x < 0.77)
    p = P[plateau, 3].mean()
    v = (P[plateau, 1] / np.sqrt(1 + P[plateau, 1] ** 2)).mean()
    assert abs(p - kat["p_star"]) / kat["p_star"] < 0.03, p
    assert abs(v - kat["v_star"]) / kat["v_star"] < 0.02, v
