// knc_oracle.cpp -- CPU restatement of the per-block RK stage of the Clemson
// Kilonova Code (KNC), used ONLY as test infrastructure: the parity checker
// for the HIP path and the CPU baseline timed by bench.py. Nothing in the
// product library links, loads or calls this file.
//
// What it restates (reference paths are relative to the app-kilonova tree):
//   src/scheme.rs:16-173   try_advance_rk   (one RK stage over every block)
//   src/scheme.rs:212-240  advance          (dt once, then RK steps)
//   src/state.rs:76-94     BlockState::try_to_primitive
//   src/state.rs:143-162   State::time_step (both branches)
//   src/state.rs:179-245   WeightedAverage  (RK convex combination)
//   src/mesh.rs:40-46,119-134,157-161   geometry formulas
//   src/physics/relativistic_hydro.rs:60-138, newtonian_hydro.rs:52-106
// The per-zone arithmetic of the reference lives in the un-vendored crates
// hydro-srhd@25bf21ec, hydro-euler@fc929f9b and godunov-core@d8dd524a; it
// is restated here from the published methods named at each function.
// Parity status of each function is listed in the header comment of the
// Python wrapper (oracle/oracle.py) and pinned by tests/test_oracle_*.py.
//
// Build: g++ -O2 -ffp-contract=off (no fast-math): every + - * / sqrt is an
// IEEE-754 correctly rounded double operation, in the order written here.

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <limits>
#include <thread>
#include <vector>

#include "knc_model_oracle.hpp"

namespace {

// ---------------------------------------------------------------------------
// Parameters (field meaning follows relativistic_hydro.rs:13-36 and
// newtonian_hydro.rs:13-28; light_speed is src/physics/mod.rs:7 for SRHD, 1
// for Euler, which leaves Euler fluxes unscaled as newtonian_hydro.rs:96-106)
// ---------------------------------------------------------------------------
struct Hydro {
    int32_t system;             // 0 = Euler, 1 = SRHD
    int32_t riemann_solver;     // 0 = HLLE, 1 = HLLC
    int32_t runge_kutta_order;  // 1, 2, 3
    int32_t adaptive_time_step; // SRHD only; Euler is always adaptive
    double gamma_law_index;
    double plm_theta;
    double cfl_number;
    double light_speed;
};

enum { SYS_EULER = 0, SYS_SRHD = 1 };
enum { HLLE = 0, HLLC = 1 };
enum {
    OK = 0,
    NEGATIVE_DENSITY = 1,
    NEGATIVE_PRESSURE = 2,
    NEGATIVE_ENERGY_DENSITY = 3,
    ROOT_FINDER_FAILED = 4,
};
enum { BND_FIXED = 0, BND_OUTFLOW = 1 };

const int C2P_MAX_ITER = 50;
const double C2P_TOL = 1e-10;
const double PI = 3.14159265358979323846;

struct Prim { double rho, ur, uq, p; };   // SRHD u = Gamma*beta; Euler u = v
struct Cons { double d, sr, sq, tau; };   // SRHD (D,S_r,S_q,tau); Euler (rho,m_r,m_q,E)

// ---------------------------------------------------------------------------
// c2p, SRHD.  Newton-Raphson on the pressure (Marti & Mueller, Living Rev.
// Relativ. 2003, sec. 2.3).  With Q = tau + D + p and den = Q^2 - S^2:
//   rho h = den/Q,  rho = D sqrt(den)/Q,  v^2 = S^2/Q^2,
//   f(p)  = (gamma-1)(rho h - rho - p) - p,   f'(p) = v^2 c_s^2 - 1,
//   c_s^2 = gamma p/(rho h).
// Multiplying f by Q leaves one division per iteration:
//   F = Q f = (gamma-1)(den - D sqrt(den) - pQ) - pQ,
//   dp = f/f' = F den / (gamma p S^2 - den Q).
// The loop stops when |dp| <= 1e-10 (tau + D + p), a test relative to Q
// because the rounding noise of f scales with Q (relative to p alone it
// stalls for cold or Gamma~100 gas).  Quadratic convergence then leaves the
// iterate within the rounding noise of f, ~1.5e-16 (tau + D), of the root
// (measured over bomb_jet, wind_shock and jet_in_cloud states against a
// long-double root; a 3e-10 stop already leaves 9e-16), which
// c2p_srhd_polish below relies on.  A 1e-12 stop costs a further iteration
// in a third of the zones for no gain.  The converged iterate then goes
// through c2p_srhd_polish.
// Error mapping follows relativistic_hydro.rs:68-88: D<0 and tau<0 are
// errors, a negative recovered pressure is floored to 1e-3 rho.
// ---------------------------------------------------------------------------
// Canonical final steps.  A converged iterate is only defined to within the
// rounding noise of f, ~eps (tau + D + p); different starting guesses land at
// different points of that band.  Rounding the iterate to the power-of-two
// grid 2^(ilogb(tau + D) - 33) (which depends on the state only) and taking
// one Newton step makes the result a function of the state alone, except
// when two iterates straddle a grid midpoint (probability ~ eps 2^33 ~ 1e-6).
// A second rounding, to the finer grid 2^(ilogb(tau + D) - 40), and Newton
// step maps those rare pairs together again except with probability ~5e-4,
// so two starting guesses disagree in about one zone in 10^9.
double c2p_srhd_newton_step(double D, double S2, double E, double gamma, double p, int shift)
{
    const int e = std::ilogb(E) - shift;
    p = std::ldexp(std::rint(std::ldexp(p, -e)), e);
    const double Q = E + p;
    const double den = Q * Q - S2;
    if (!(den > 0.0)) return p;
    const double sq = std::sqrt(den);
    const double pQ = p * Q;
    const double F = (gamma - 1.0) * (den - D * sq - pQ) - pQ;
    return p - F * den / (gamma * p * S2 - den * Q);
}

double c2p_srhd_polish(double D, double S2, double E, double gamma, double p)
{
    if (!(E > 0.0)) return p;
    p = c2p_srhd_newton_step(D, S2, E, gamma, p, 33);
    return c2p_srhd_newton_step(D, S2, E, gamma, p, 40);
}

int c2p_srhd_newton(double D, double S2, double tau, double gamma, double p,
                    double* p_out, int* iters)
{
    const double E = tau + D;
    const double gm1 = gamma - 1.0;
    for (int n = 1; n <= C2P_MAX_ITER; ++n) {
        const double Q = E + p;
        const double den = Q * Q - S2;
        if (!(den > 0.0)) { *iters = n; return ROOT_FINDER_FAILED; }
        const double sq = std::sqrt(den);
        const double pQ = p * Q;
        const double F = gm1 * (den - D * sq - pQ) - pQ;
        const double dp = F * den / (gamma * p * S2 - den * Q);
        p = p - dp;
        if (!(std::fabs(p) < 1e300)) { *iters = n; return ROOT_FINDER_FAILED; }
        if (std::fabs(dp) <= C2P_TOL * (E + p)) {
            *iters = n + 2;
            *p_out = c2p_srhd_polish(D, S2, E, gamma, p);
            return OK;
        }
    }
    *iters = C2P_MAX_ITER;
    return ROOT_FINDER_FAILED;
}

// Primitive from the converged pressure: rho = D sqrt(den)/Q, u = S/sqrt(den)
// (u = Gamma v with Gamma = Q/sqrt(den)); *rW receives 1/Gamma = sqrt(den)/Q,
// formed as den/(Q sqrt(den)).
int c2p_srhd(const Cons& u, double gamma, double guess, Prim* out, int* iters, double* rW)
{
    *iters = 0;
    if (u.d < 0.0) return NEGATIVE_DENSITY;
    if (u.tau < 0.0) return NEGATIVE_ENERGY_DENSITY;
    const double S2 = u.sr * u.sr + u.sq * u.sq;
    double p = 0.0;
    int rc = c2p_srhd_newton(u.d, S2, u.tau, gamma, guess, &p, iters);
    if (rc != OK) return rc;
    const double Q = (u.tau + u.d) + p;
    const double den = Q * Q - S2;
    const double sq = std::sqrt(den);
    // one division: r = 1/(Q sqrt(den)), 1/Gamma = sqrt(den)/Q = den r, 1/sqrt(den) = Q r
    const double r = 1.0 / (Q * sq);
    const double rsq = Q * r;
    *rW = den * r;
    out->rho = u.d * *rW;
    out->ur = u.sr * rsq;
    out->uq = u.sq * rsq;
    out->p = p;
    if (p < 0.0) { out->p = 1e-3 * out->rho; return NEGATIVE_PRESSURE; }
    return OK;
}

// c2p, Euler: closed form; only D<0 is an error (newtonian_hydro.rs:60-65).
int c2p_euler(const Cons& u, double gamma, Prim* out)
{
    if (u.d < 0.0) return NEGATIVE_DENSITY;
    const double rho = u.d;
    const double rr = 1.0 / rho;
    const double vr = u.sr * rr;
    const double vq = u.sq * rr;
    const double ek = (0.5 * rho) * (vr * vr + vq * vq);
    out->rho = rho;
    out->ur = vr;
    out->uq = vq;
    out->p = (u.tau - ek) * (gamma - 1.0);
    return OK;
}

// p2c (to_conserved), closed form.
Cons p2c(const Hydro& h, const Prim& p)
{
    const double g = h.gamma_law_index;
    Cons u;
    if (h.system == SYS_EULER) {
        u.d = p.rho;
        u.sr = p.rho * p.ur;
        u.sq = p.rho * p.uq;
        u.tau = p.p / (g - 1.0) + (0.5 * p.rho) * (p.ur * p.ur + p.uq * p.uq);
    } else {
        const double W = std::sqrt(1.0 + (p.ur * p.ur + p.uq * p.uq));
        const double rhoh = p.rho + p.p * (g / (g - 1.0));
        const double hw = rhoh * W;
        u.d = p.rho * W;
        u.sr = hw * p.ur;
        u.sq = hw * p.uq;
        u.tau = (hw * W - p.p) - u.d;
    }
    return u;
}

// Physical flux along axis (0 radial, 1 polar) plus the passive scalar:
// U_s = s*W, F_s = U_s*v_n.
// The outer speeds also as fractions lm = nm/den, lp = np/den (den > 0), the
// form the HLLE flux combines with one division.
struct FluxState {
    Cons u, f;
    double us, fs, vn, lm, lp, p, ut;
    double nm, np, den;
};

FluxState flux_state(const Hydro& h, const Prim& P, double s, int axis)
{
    const double g = h.gamma_law_index;
    FluxState st;
    st.u = p2c(h, P);
    const double un = axis == 0 ? P.ur : P.uq;
    st.p = P.p;
    st.ut = 0.0;
    if (h.system == SYS_EULER) {
        st.vn = un;
        st.us = s;
        const double cs = std::sqrt(g * P.p / P.rho);
        st.lm = st.vn - cs;
        st.lp = st.vn + cs;
        st.nm = st.lm;
        st.np = st.lp;
        st.den = 1.0;
        st.f.d = st.u.d * st.vn;
        st.f.sr = st.u.sr * st.vn;
        st.f.sq = st.u.sq * st.vn;
        if (axis == 0) st.f.sr = st.f.sr + P.p; else st.f.sq = st.f.sq + P.p;
        st.f.tau = (st.u.tau + P.p) * st.vn;
        st.fs = st.us * st.vn;
        return st;
    }
    const double uu = P.ur * P.ur + P.uq * P.uq;
    const double W = std::sqrt(1.0 + uu);
    const double rhoh = P.rho + P.p * (g / (g - 1.0));
    const double ut = axis == 0 ? P.uq : P.ur;
    // Outer characteristic speeds of Marti & Mueller (2003) eq. (72),
    // numerator and denominator multiplied by W^2 rho h.  With gp = gamma p
    // and e = rho h - gp:
    //   denominator  W^2 rho h (1 - v^2 c_s^2) = rho h + u^2 e,
    //   root         W^2 rho h c_s sqrt((1-v^2)(1 - v^2 c_s^2 - v_n^2 (1 - c_s^2)))
    //                  = sqrt(gp (rho h + u_t^2 e)),
    //   first term   W^2 rho h v_n (1 - c_s^2) = u_n W e,
    // so one division serves both speeds (HLLE keeps the fractions).
    const double gp = g * P.p;
    const double e = rhoh - gp;
    const double Y = rhoh + uu * e;
    const double rY = 1.0 / Y;
    const double k0 = std::sqrt(gp * (rhoh + (ut * ut) * e));
    const double b = (un * W) * e;
    st.nm = b - k0;
    st.np = b + k0;
    st.den = Y;
    st.lm = st.nm * rY;
    st.lp = st.np * rY;
    st.vn = un / W;   // HLLC only
    st.us = s * W;
    // Physical fluxes with v_n = u_n/W cancelled against the W of U:
    //   F_D = rho u_n,  F_S = rho h u u_n + p n,  F_tau = (tau + p) v_n = (rho h W - rho) u_n,
    //   F_s = s W v_n = s u_n.
    st.f.d = P.rho * un;
    st.f.sr = (rhoh * P.ur) * un;
    st.f.sq = (rhoh * P.uq) * un;
    if (axis == 0) st.f.sr = st.f.sr + P.p; else st.f.sq = st.f.sq + P.p;
    st.f.tau = (rhoh * W - P.rho) * un;
    st.fs = s * un;
    return st;
}

inline double cn(const Cons& u, int axis) { return axis == 0 ? u.sr : u.sq; }

// HLLE / HLLC Riemann solver with passive scalar.  Returns the unscaled
// flux; the adapter multiplies by light_speed (relativistic_hydro.rs:123-134).
// HLLE: Einfeldt / Harten-Lax-van Leer with am = min(0,.), ap = max(0,.).
// HLLC SRHD: Mignone & Bodo, MNRAS 364 (2005) 126, eqs. (16)-(18).
// HLLC Euler: Toro, Riemann Solvers (2009), sec. 10.4, eqs. (10.37)-(10.39).
void riemann(const Hydro& h, const Prim& PL, const Prim& PR, double sL, double sR,
             int axis, Cons* F, double* Fs)
{
    const FluxState L = flux_state(h, PL, sL, axis);
    const FluxState R = flux_state(h, PR, sR, axis);

    if (h.riemann_solver == HLLE) {
        // a0 = min(lm_L, lm_R, 0) = n0/d0, a1 = max(lp_L, lp_R, 0) = n1/d1,
        // selected by cross-multiplication; the HLLE flux
        // (a1 F_L - a0 F_R + a1 a0 (U_R - U_L)) / (a1 - a0) multiplied through
        // by d0 d1.
        const bool left_m = L.nm * R.den <= R.nm * L.den;
        const double nm = left_m ? L.nm : R.nm, dm = left_m ? L.den : R.den;
        const bool left_p = L.np * R.den >= R.np * L.den;
        const double np = left_p ? L.np : R.np, dp = left_p ? L.den : R.den;
        const double n0 = nm < 0.0 ? nm : 0.0, d0 = nm < 0.0 ? dm : 1.0;
        const double n1 = np > 0.0 ? np : 0.0, d1 = np > 0.0 ? dp : 1.0;
        const double c1 = n1 * d0, c0 = n0 * d1, ab = n1 * n0;
        const double rd = 1.0 / (c1 - c0);
        F->d   = (L.f.d   * c1 - R.f.d   * c0 + (R.u.d   - L.u.d)   * ab) * rd;
        F->sr  = (L.f.sr  * c1 - R.f.sr  * c0 + (R.u.sr  - L.u.sr)  * ab) * rd;
        F->sq  = (L.f.sq  * c1 - R.f.sq  * c0 + (R.u.sq  - L.u.sq)  * ab) * rd;
        F->tau = (L.f.tau * c1 - R.f.tau * c0 + (R.u.tau - L.u.tau) * ab) * rd;
        *Fs    = (L.fs    * c1 - R.fs    * c0 + (R.us    - L.us)    * ab) * rd;
        return;
    }
    const double am = std::min(L.lm, R.lm);
    const double ap = std::max(L.lp, R.lp);
    if (am >= 0.0) { *F = L.f; *Fs = L.fs; return; }
    if (ap <= 0.0) { *F = R.f; *Fs = R.fs; return; }

    double lstar;
    double pstar = 0.0;
    if (h.system == SYS_SRHD) {
        const double rd = 1.0 / (ap - am);
        const double ab = ap * am;
        const double ud = (R.u.d * ap - L.u.d * am + L.f.d - R.f.d) * rd;
        const double ut = (R.u.tau * ap - L.u.tau * am + L.f.tau - R.f.tau) * rd;
        const double um = (cn(R.u, axis) * ap - cn(L.u, axis) * am + cn(L.f, axis) - cn(R.f, axis)) * rd;
        const double fd = (L.f.d * ap - R.f.d * am + (R.u.d - L.u.d) * ab) * rd;
        const double ft = (L.f.tau * ap - R.f.tau * am + (R.u.tau - L.u.tau) * ab) * rd;
        const double fm = (cn(L.f, axis) * ap - cn(R.f, axis) * am + (cn(R.u, axis) - cn(L.u, axis)) * ab) * rd;
        const double Ehll = ut + ud;
        const double FEhll = ft + fd;
        const double b = Ehll + fm;
        const double disc = b * b - 4.0 * FEhll * um;
        lstar = 2.0 * um / (b + std::sqrt(disc));
        pstar = fm - FEhll * lstar;
    } else {
        const double dl = L.u.d * (am - L.vn);
        const double dr = R.u.d * (ap - R.vn);
        lstar = (R.p - L.p + dl * L.vn - dr * R.vn) / (dl - dr);
    }
    const bool left = lstar >= 0.0;
    const FluxState& K = left ? L : R;
    const double lam = left ? am : ap;
    const double lv = lam - K.vn;
    const double r = 1.0 / (lam - lstar);
    Cons us;
    double uss;
    if (h.system == SYS_SRHD) {
        const double fac = lv * r;
        const double EK = K.u.tau + K.u.d;
        us.d = K.u.d * fac;
        const double snst = (cn(K.u, axis) * lv + pstar - K.p) * r;
        const double stst = (axis == 0 ? K.u.sq : K.u.sr) * fac;
        const double Est = (EK * lv + pstar * lstar - K.p * K.vn) * r;
        if (axis == 0) { us.sr = snst; us.sq = stst; } else { us.sq = snst; us.sr = stst; }
        us.tau = Est - us.d;
        uss = K.us * fac;
    } else {
        const double fac = K.u.d * lv * r;
        const double rhoK = K.u.d;
        const double vt = (axis == 0 ? K.u.sq : K.u.sr) / rhoK;
        const double en = K.u.tau / rhoK + (lstar - K.vn) * (lstar + K.p / (rhoK * lv));
        us.d = fac;
        if (axis == 0) { us.sr = fac * lstar; us.sq = fac * vt; }
        else           { us.sq = fac * lstar; us.sr = fac * vt; }
        us.tau = fac * en;
        uss = K.us * (lv * r);
    }
    F->d   = K.f.d   + lam * (us.d   - K.u.d);
    F->sr  = K.f.sr  + lam * (us.sr  - K.u.sr);
    F->sq  = K.f.sq  + lam * (us.sq  - K.u.sq);
    F->tau = K.f.tau + lam * (us.tau - K.u.tau);
    *Fs    = K.fs    + lam * (uss    - K.us);
}

// Spherical-polar geometric source terms per unit volume at the zone centre
// (r, theta), given 1/r; mass and energy sources are zero.
//   SRHD:  S_r = (2p + rho h u_q^2)/r,  S_q = (p cot - rho h u_r u_q)/r
//   Euler: S_r = (2p + rho v_q^2)/r,    S_q = (p cot - rho v_r v_q)/r
Cons sources(const Hydro& h, const Prim& P, double rinv, double cot)
{
    const double g = h.gamma_law_index;
    const double rhoh = h.system == SYS_EULER ? P.rho : P.rho + P.p * (g / (g - 1.0));
    Cons s;
    s.d = 0.0;
    s.sr = (2.0 * P.p + rhoh * (P.uq * P.uq)) * rinv;
    s.sq = (P.p * cot - rhoh * (P.ur * P.uq)) * rinv;
    s.tau = 0.0;
    return s;
}

// ---------------------------------------------------------------------------
// Literal forms of the crates' arithmetic.  The functions above restate the
// un-vendored crates with quotients rewritten for speed (one division per
// Newton step, per face and per flux; reciprocal radii; speeds kept as
// fractions), identically on the device.  The forms below write the same
// quantities as the published methods state them, every quotient divided
// out, so that a Sim built with `literal` set measures how far two
// restatements of the crates' arithmetic move the state:
//   sources divided by the centre radius;
//   wave speeds of Marti & Mueller (2003) eq. (72) in velocity form, the
//     HLLE flux (a+ F_L - a- F_R + a+ a- (U_R - U_L)) / (a+ - a-), and the
//     HLLC star states (Mignone & Bodo 2005, Toro 2009 sec. 10.4), each
//     quotient divided.
// The expressions of the reference's own code (face areas, volumes, U/dV,
// M/dV/Gamma) are written as the reference writes them in both modes.
// ---------------------------------------------------------------------------
// Parts of the literal mode (a Sim's `literal` is a mask of them), both
// restating the un-vendored crates' arithmetic in its textbook form.
enum {
    LIT_SOURCES = 4,    // sources divided by r
    LIT_FLUX = 8,       // wave speeds and HLLE/HLLC fluxes with every quotient divided
};

struct FluxLit {
    Cons u, f;
    double us, fs, vn, lm, lp, p;
};

FluxLit flux_state_literal(const Hydro& h, const Prim& P, double s, int axis)
{
    const double g = h.gamma_law_index;
    FluxLit st;
    st.u = p2c(h, P);
    st.p = P.p;
    const double un = axis == 0 ? P.ur : P.uq;
    if (h.system == SYS_EULER) {
        st.vn = un;
        st.us = s;
        const double cs = std::sqrt(g * P.p / P.rho);
        st.lm = st.vn - cs;
        st.lp = st.vn + cs;
    } else {
        const double W = std::sqrt(1.0 + P.ur * P.ur + P.uq * P.uq);
        const double v2 = (P.ur * P.ur + P.uq * P.uq) / (W * W);
        const double enthalpy = 1.0 + g * P.p / ((g - 1.0) * P.rho);
        const double cs2 = g * P.p / (P.rho * enthalpy);
        const double cs = std::sqrt(cs2);
        st.vn = un / W;
        const double root = std::sqrt((1.0 - v2) * (1.0 - v2 * cs2 - st.vn * st.vn * (1.0 - cs2)));
        const double den = 1.0 - v2 * cs2;
        st.lm = (st.vn * (1.0 - cs2) - cs * root) / den;
        st.lp = (st.vn * (1.0 - cs2) + cs * root) / den;
        st.us = s * W;
    }
    // F = U v_n + p n, F_tau = (tau + p) v_n (F_E = E v_n + p v_n for Euler), F_s = U_s v_n
    st.f.d = st.u.d * st.vn;
    st.f.sr = st.u.sr * st.vn;
    st.f.sq = st.u.sq * st.vn;
    if (axis == 0) st.f.sr = st.f.sr + P.p; else st.f.sq = st.f.sq + P.p;
    st.f.tau = (st.u.tau + P.p) * st.vn;
    st.fs = st.us * st.vn;
    return st;
}

void riemann_literal(const Hydro& h, const Prim& PL, const Prim& PR, double sL, double sR, int axis, Cons* F,
                     double* Fs)
{
    const FluxLit L = flux_state_literal(h, PL, sL, axis);
    const FluxLit R = flux_state_literal(h, PR, sR, axis);
    if (h.riemann_solver == HLLE) {
        const double am = std::min(std::min(L.lm, R.lm), 0.0);
        const double ap = std::max(std::max(L.lp, R.lp), 0.0);
        const double w = ap - am;
        F->d   = (ap * L.f.d   - am * R.f.d   + ap * am * (R.u.d   - L.u.d))   / w;
        F->sr  = (ap * L.f.sr  - am * R.f.sr  + ap * am * (R.u.sr  - L.u.sr))  / w;
        F->sq  = (ap * L.f.sq  - am * R.f.sq  + ap * am * (R.u.sq  - L.u.sq))  / w;
        F->tau = (ap * L.f.tau - am * R.f.tau + ap * am * (R.u.tau - L.u.tau)) / w;
        *Fs    = (ap * L.fs    - am * R.fs    + ap * am * (R.us    - L.us))    / w;
        return;
    }
    const double am = std::min(L.lm, R.lm);
    const double ap = std::max(L.lp, R.lp);
    if (am >= 0.0) { *F = L.f; *Fs = L.fs; return; }
    if (ap <= 0.0) { *F = R.f; *Fs = R.fs; return; }
    const double w = ap - am;
    double lstar, pstar = 0.0;
    if (h.system == SYS_SRHD) {
        // HLL state and flux of E = tau + D and of the normal momentum
        const double EL = L.u.tau + L.u.d, ER = R.u.tau + R.u.d;
        const double FEL = L.f.tau + L.f.d, FER = R.f.tau + R.f.d;
        const double E = (ap * ER - am * EL + FEL - FER) / w;
        const double m = (ap * cn(R.u, axis) - am * cn(L.u, axis) + cn(L.f, axis) - cn(R.f, axis)) / w;
        const double FE = (ap * FEL - am * FER + ap * am * (ER - EL)) / w;
        const double Fm = (ap * cn(L.f, axis) - am * cn(R.f, axis) + ap * am * (cn(R.u, axis) - cn(L.u, axis))) / w;
        // root of FE l^2 - (E + Fm) l + m = 0 with the minus sign (Mignone &
        // Bodo eq. 18), written 2m / (b + sqrt(b^2 - 4 FE m)), defined at FE = 0
        const double b = E + Fm;
        lstar = 2.0 * m / (b + std::sqrt(b * b - 4.0 * FE * m));
        pstar = Fm - FE * lstar;
    } else {
        lstar = (R.p - L.p + L.u.d * L.vn * (am - L.vn) - R.u.d * R.vn * (ap - R.vn)) /
                (L.u.d * (am - L.vn) - R.u.d * (ap - R.vn));
    }
    const bool left = lstar >= 0.0;
    const FluxLit& K = left ? L : R;
    const double lam = left ? am : ap;
    Cons us;
    double uss;
    if (h.system == SYS_SRHD) {
        const double EK = K.u.tau + K.u.d;
        us.d = K.u.d * (lam - K.vn) / (lam - lstar);
        const double mn = (cn(K.u, axis) * (lam - K.vn) + pstar - K.p) / (lam - lstar);
        const double mt = (axis == 0 ? K.u.sq : K.u.sr) * (lam - K.vn) / (lam - lstar);
        const double Est = (EK * (lam - K.vn) + pstar * lstar - K.p * K.vn) / (lam - lstar);
        if (axis == 0) { us.sr = mn; us.sq = mt; } else { us.sq = mn; us.sr = mt; }
        us.tau = Est - us.d;
        uss = K.us * (lam - K.vn) / (lam - lstar);
    } else {
        const double rhoK = K.u.d;
        const double fac = rhoK * (lam - K.vn) / (lam - lstar);
        const double vt = (axis == 0 ? K.u.sq : K.u.sr) / rhoK;
        const double en = K.u.tau / rhoK + (lstar - K.vn) * (lstar + K.p / (rhoK * (lam - K.vn)));
        us.d = fac;
        if (axis == 0) { us.sr = fac * lstar; us.sq = fac * vt; }
        else           { us.sq = fac * lstar; us.sr = fac * vt; }
        us.tau = fac * en;
        uss = K.us * (lam - K.vn) / (lam - lstar);
    }
    F->d   = K.f.d   + lam * (us.d   - K.u.d);
    F->sr  = K.f.sr  + lam * (us.sr  - K.u.sr);
    F->sq  = K.f.sq  + lam * (us.sq  - K.u.sq);
    F->tau = K.f.tau + lam * (us.tau - K.u.tau);
    *Fs    = K.fs    + lam * (uss    - K.us);
}

// Sources at the centre (r, theta), divided by r.
Cons sources_literal(const Hydro& h, const Prim& P, double r, double theta)
{
    const double g = h.gamma_law_index;
    const double rhoh = h.system == SYS_EULER ? P.rho : P.rho * (1.0 + g * P.p / ((g - 1.0) * P.rho));
    const double cot = std::cos(theta) / std::sin(theta);
    Cons s;
    s.d = 0.0;
    s.sr = (2.0 * P.p + rhoh * P.uq * P.uq) / r;
    s.sq = (P.p * cot - rhoh * P.ur * P.uq) / r;
    s.tau = 0.0;
    return s;
}

// Lorentz factor of a recovered primitive for s = M/dV/Gamma:
// lorentz_factor_squared().sqrt() (relativistic_hydro.rs:162-164); 1 for
// Euler (newtonian_hydro.rs:130-132).
double lorentz_factor(const Hydro& h, const Prim& P)
{
    return h.system == SYS_EULER ? 1.0 : std::sqrt(1.0 + P.ur * P.ur + P.uq * P.uq);
}

// Maximum signal speed (unscaled).  SRHD: relativistic sum of |v| and c_s,
// an upper bound of every characteristic speed; Euler: |v| + c_s.
double max_signal_speed(const Hydro& h, const Prim& P)
{
    const double g = h.gamma_law_index;
    if (h.system == SYS_EULER) {
        return std::sqrt(P.ur * P.ur + P.uq * P.uq) + std::sqrt(g * P.p / P.rho);
    }
    const double uu = P.ur * P.ur + P.uq * P.uq;
    const double rhoh = P.rho + P.p * (g / (g - 1.0));
    const double cs = std::sqrt(g * P.p / rhoh);
    const double v = std::sqrt(uu / (1.0 + uu));
    return (v + cs) / (1.0 + v * cs);
}

// Generalised-minmod PLM difference (godunov-core plm_gradient):
// minmod((b-a) theta, (c-a)/2, (c-b) theta).
// All three candidate slopes positive <=> their minimum is positive; all
// negative <=> their maximum is negative; otherwise the slope is zero.
inline double plm(double theta, double a, double b, double c)
{
    const double x = (b - a) * theta;
    const double y = (c - a) * 0.5;
    const double z = (c - b) * theta;
    const double lo = std::fmin(x, std::fmin(y, z));
    const double hi = std::fmax(x, std::fmax(y, z));
    return lo > 0.0 ? lo : (hi < 0.0 ? hi : 0.0);
}

// Primitive + scalar concentration, the element of the extended arrays
// pe/se of scheme.rs:76-77.
struct PS { Prim P; double s; };

inline PS plm5(double th, const PS& a, const PS& b, const PS& c)
{
    PS g;
    g.P.rho = plm(th, a.P.rho, b.P.rho, c.P.rho);
    g.P.ur = plm(th, a.P.ur, b.P.ur, c.P.ur);
    g.P.uq = plm(th, a.P.uq, b.P.uq, c.P.uq);
    g.P.p = plm(th, a.P.p, b.P.p, c.P.p);
    g.s = plm(th, a.s, b.s, c.s);
    return g;
}
inline PS add_half(const PS& a, const PS& g)
{
    return PS{{a.P.rho + g.P.rho * 0.5, a.P.ur + g.P.ur * 0.5, a.P.uq + g.P.uq * 0.5, a.P.p + g.P.p * 0.5}, a.s + g.s * 0.5};
}
inline PS sub_half(const PS& a, const PS& g)
{
    return PS{{a.P.rho - g.P.rho * 0.5, a.P.ur - g.P.ur * 0.5, a.P.uq - g.P.uq * 0.5, a.P.p - g.P.p * 0.5}, a.s - g.s * 0.5};
}

// ---------------------------------------------------------------------------
// Geometry of one block: the 2-D arrays of GridGeometry (mesh.rs:21-29),
// zone by zone with the expressions of mesh.rs:119-134,157-161 as written
// there (the frustum area pi (s0+s1) sqrt(ds^2+dz^2) from the corners, the
// volume 2 pi (r1^3 - r0^3) / 3 * dcost), plus 1/r and cot theta at the
// centre for the sources.
// ---------------------------------------------------------------------------
struct RadialRow {            // per radial index, from the vertices
    double volA;              // ((2 pi)(r1^3 - r0^3))/3
    double rc, rinv;          // sqrt(r0 r1), 1/rc
};

struct PolarCol {             // per polar index, from the vertices
    double dcost;             // -(cos q1 - cos q0)
    double cot;               // cos(qc)/sin(qc), qc = (q0+q1)/2
};

RadialRow radial_row(double r0, double r1)
{
    RadialRow t;
    t.volA = 2.0 * PI * ((r1 * r1) * r1 - (r0 * r0) * r0) / 3.0;
    t.rc = std::sqrt(r0 * r1);
    t.rinv = 1.0 / t.rc;
    return t;
}

PolarCol polar_col(double q0, double q1)
{
    PolarCol t;
    t.dcost = -(std::cos(q1) - std::cos(q0));
    const double qc = 0.5 * (q0 + q1);
    t.cot = std::cos(qc) / std::sin(qc);
    return t;
}

struct Block {
    int nr = 0, nq = 0;
    int literal = 0;                 // LIT_* parts of the crates' arithmetic in their literal forms
    std::vector<double> rv, qv;      // vertices (nr+1), (nq+1)
    std::vector<double> Ar, Aq, dV;  // (nr+1,nq), (nr,nq+1), (nr,nq)
    std::vector<double> rinv, cot, dl; // 1/rc, cot(qc), linear dim per zone
    std::vector<double> rc, qc;      // zone centre (src/mesh.rs:157-161)
    std::vector<Cons> U;             // conserved * dV, (nr,nq) row-major, theta fastest
    std::vector<double> M;           // scalar mass
};

// face_area and cell_volume of src/mesh.rs:119-134, as written there:
// c = (r, theta) corners, powi(3) as r * r * r.
double face_area_literal(double r0, double q0, double r1, double q1)
{
    const double s0 = r0 * std::sin(q0);
    const double s1 = r1 * std::sin(q1);
    const double z0 = r0 * std::cos(q0);
    const double z1 = r1 * std::cos(q1);
    const double ds = s1 - s0;
    const double dz = z1 - z0;
    return PI * (s0 + s1) * std::sqrt(ds * ds + dz * dz);
}

double cell_volume_literal(double r0, double q0, double r1, double q1)
{
    const double dcost = -(std::cos(q1) - std::cos(q0));
    return 2.0 * PI * (r1 * r1 * r1 - r0 * r0 * r0) / 3.0 * dcost;
}

void build_geometry(Block& b)
{
    const int nr = b.nr, nq = b.nq;
    b.Ar.assign((size_t)(nr + 1) * nq, 0.0);
    b.Aq.assign((size_t)nr * (nq + 1), 0.0);
    b.dV.assign((size_t)nr * nq, 0.0);
    b.rinv.assign((size_t)nr * nq, 0.0);
    b.rc.assign((size_t)nr * nq, 0.0);
    b.qc.assign((size_t)nr * nq, 0.0);
    b.cot.assign((size_t)nr * nq, 0.0);
    b.dl.assign((size_t)nr * nq, 0.0);
    std::vector<RadialRow> R(nr);
    std::vector<PolarCol> C(nq);
    for (int i = 0; i < nr; ++i) R[i] = radial_row(b.rv[i], b.rv[i + 1]);
    for (int j = 0; j < nq; ++j) C[j] = polar_col(b.qv[j], b.qv[j + 1]);
    for (int i = 0; i <= nr; ++i)
        for (int j = 0; j < nq; ++j)   // zone(i,j).face_area_r()
            b.Ar[(size_t)i * nq + j] = face_area_literal(b.rv[i], b.qv[j], b.rv[i], b.qv[j + 1]);
    for (int i = 0; i < nr; ++i)
        for (int j = 0; j <= nq; ++j)  // zone(i,j).face_area_q()
            b.Aq[(size_t)i * (nq + 1) + j] = face_area_literal(b.rv[i], b.qv[j], b.rv[i + 1], b.qv[j]);
    for (int i = 0; i < nr; ++i)
        for (int j = 0; j < nq; ++j) {
            const size_t k = (size_t)i * nq + j;
            b.dV[k] = cell_volume_literal(b.rv[i], b.qv[j], b.rv[i + 1], b.qv[j + 1]);   // zone(i,j).volume()
            b.rc[k] = R[i].rc;
            b.qc[k] = 0.5 * (b.qv[j] + b.qv[j + 1]);
            b.rinv[k] = R[i].rinv;
            b.cot[k] = C[j].cot;
            const double dr = b.rv[i + 1] - b.rv[i];
            const double dq = b.qv[j + 1] - b.qv[j];
            b.dl[k] = std::min(dr, dq * b.rv[i]);
        }
}

struct Boundary {
    int mode = BND_OUTFLOW;
    std::vector<PS> rows;   // 2*nq, rows in radial order (inner: -2,-1; outer: nr,nr+1)
};

struct ErrRec {
    int code = OK;
    int block = 0, i = 0, j = 0;
    double r = 0, theta = 0, value = 0;
    Cons u{0, 0, 0, 0};
};

struct C2PStats {
    std::atomic<long long> zones{0}, iterations{0}, floors{0};
};

// Ghost rows evaluated at each stage's input time, as BlockState::from_model
// on the boundary blocks at state.time (src/scheme.rs:46-52): fills prim
// (2*nq*4) and scalar (2*nq) for side 0 (inner) or 1 (outer).
typedef int (*BoundarySource)(void* user, int side, double time, int nq, double* prim, double* scalar);

struct Sim {
    Hydro h;
    std::vector<Block> blocks;   // radial order
    Boundary inner, outer;
    BoundarySource source[2] = {nullptr, nullptr};
    void* source_user[2] = {nullptr, nullptr};
    // model boundaries: ghost blocks of the mesh from the initial model
    bool model_on[2] = {false, false};
    om::Model model[2]{};
    std::vector<double> model_table[2];
    om::Mesh mesh[2]{};
    int first_index = 0;          // block index of blocks[0]
    double time = 0.0;
    long long iteration = 0;
    int nthreads = 1;
    int literal = 0;              // LIT_* parts of blocks, births and model rows in literal form
    C2PStats stats;
};

void parallel_for(int n, int nthreads, const std::function<void(int)>& fn)
{
    if (nthreads <= 1 || n <= 1) {
        for (int k = 0; k < n; ++k) fn(k);
        return;
    }
    std::atomic<int> next{0};
    std::vector<std::thread> pool;
    const int nt = std::min(nthreads, n);
    for (int t = 0; t < nt; ++t)
        pool.emplace_back([&]() {
            for (int k = next++; k < n; k = next++) fn(k);
        });
    for (auto& th : pool) th.join();
}

// Error ordering: lowest (block, i, j) wins, so the result is deterministic
// (the reference reports whichever task failed first, scheme.rs:162-166).
inline bool err_less(const ErrRec& a, const ErrRec& b)
{
    if (a.block != b.block) return a.block < b.block;
    if (a.i != b.i) return a.i < b.i;
    return a.j < b.j;
}

// BlockState::try_to_primitive (state.rs:76-94) plus the scalar
// concentration s = M/dV/Gamma (scheme.rs:35), for one block.
ErrRec block_primitives(Sim& sim, int bi, const std::vector<Cons>& U, const std::vector<double>& M,
                        std::vector<PS>& out)
{
    const Block& b = sim.blocks[bi];
    const size_t n = (size_t)b.nr * b.nq;
    out.resize(n);
    ErrRec err;
    long long iters_sum = 0, floors = 0;
    for (size_t k = 0; k < n; ++k) {
        const double dv = b.dV[k];
        const Cons u{U[k].d / dv, U[k].sr / dv, U[k].sq / dv, U[k].tau / dv};   // state.rs:85
        int rc, it = 0;
        double value = 0.0;
        double rW = 1.0;
        if (sim.h.system == SYS_EULER) {
            rc = c2p_euler(u, sim.h.gamma_law_index, &out[k].P);
            value = u.d;
        } else {
            rc = c2p_srhd(u, sim.h.gamma_law_index, 0.0, &out[k].P, &it, &rW);
            value = rc == NEGATIVE_DENSITY ? u.d : u.tau;
        }
        iters_sum += it;
        if (rc == NEGATIVE_PRESSURE) { ++floors; rc = OK; }
        if (rc != OK) {
            ErrRec e;
            e.code = rc; e.block = bi; e.i = (int)(k / b.nq); e.j = (int)(k % b.nq);
            e.r = b.rc[k]; e.theta = 0.5 * (b.qv[e.j] + b.qv[e.j + 1]); e.value = value; e.u = u;
            if (err.code == OK || err_less(e, err)) err = e;
            out[k] = PS{{0, 0, 0, 0}, 0};
            continue;
        }
        // s = M/dV/Gamma (scheme.rs:35)
        out[k].s = M[k] / dv / lorentz_factor(sim.h, out[k].P);
    }
    sim.stats.zones += (long long)n;
    sim.stats.iterations += iters_sum;
    sim.stats.floors += floors;
    return err;
}

// One RK stage over all blocks: try_advance_rk (scheme.rs:16-173) for the
// 2-D and 1-D branches.  Uout = Uin + dt L(Uin); then, when bweight > 0,
// the weighted average Uout*(1-b) + U0*b (state.rs:179-212) is applied.
ErrRec stage(Sim& sim, const std::vector<std::vector<Cons>>& Uin, const std::vector<std::vector<double>>& Min,
             const std::vector<std::vector<Cons>>* U0, const std::vector<std::vector<double>>* M0, double bweight,
             double dt, std::vector<std::vector<Cons>>& Uout, std::vector<std::vector<double>>& Mout)
{
    const Hydro& h = sim.h;
    const int nb = (int)sim.blocks.size();
    std::vector<std::vector<PS>> prim(nb);
    std::vector<ErrRec> errs(nb);
    parallel_for(nb, sim.nthreads, [&](int b) { errs[b] = block_primitives(sim, b, Uin[b], Min[b], prim[b]); });
    ErrRec first;
    for (int b = 0; b < nb; ++b)
        if (errs[b].code != OK) { first = errs[b]; break; }
    if (first.code != OK) return first;

    Uout.resize(nb);
    Mout.resize(nb);
    // Work items: (block, row chunk).
    const int chunk = 16;
    std::vector<std::pair<int, int>> items;
    for (int b = 0; b < nb; ++b) {
        Uout[b].resize(Uin[b].size());
        Mout[b].resize(Min[b].size());
        for (int i0 = 0; i0 < sim.blocks[b].nr; i0 += chunk) items.push_back({b, i0});
    }
    const double th = h.plm_theta;
    const double c = h.light_speed;
    parallel_for((int)items.size(), sim.nthreads, [&](int w) {
        const int b = items[w].first, i0 = items[w].second;
        const Block& B = sim.blocks[b];
        const int nr = B.nr, nq = B.nq;
        const int i1 = std::min(nr, i0 + chunk);
        // pe row accessor: extended index k in [0, nr+4) -> own row k-2, or
        // the last two rows of the inner neighbour, or the first two rows of
        // the outer neighbour (scheme.rs:76-77), or the boundary rows.
        auto pe = [&](int k, int j) -> const PS& {
            const int i = k - 2;
            if (i >= 0 && i < nr) return prim[b][(size_t)i * nq + j];
            if (i < 0) {
                if (b > 0) {
                    const int nl = sim.blocks[b - 1].nr;
                    return prim[b - 1][(size_t)(nl + i) * nq + j];
                }
                if (sim.inner.mode == BND_OUTFLOW) return prim[b][j];
                return sim.inner.rows[(size_t)(i + 2) * nq + j];
            }
            if (b + 1 < nb) return prim[b + 1][(size_t)(i - nr) * nq + j];
            if (sim.outer.mode == BND_OUTFLOW) return prim[b][(size_t)(nr - 1) * nq + j];
            return sim.outer.rows[(size_t)(i - nr) * nq + j];
        };
        // Radial slopes gx[k] = plm(pe[k], pe[k+1], pe[k+2]) for k in [i0, i1+2)
        // and radial fluxes at faces f in [i0, i1] (scheme.rs:79-97).
        const int nf = i1 - i0 + 1;
        std::vector<PS> gx((size_t)(nf + 1) * nq);
        for (int k = i0; k <= i1 + 1; ++k)
            for (int j = 0; j < nq; ++j)
                gx[(size_t)(k - i0) * nq + j] = plm5(th, pe(k, j), pe(k + 1, j), pe(k + 2, j));
        std::vector<Cons> Fx((size_t)nf * nq);
        std::vector<double> Gx((size_t)nf * nq);
        for (int f = i0; f <= i1; ++f)
            for (int j = 0; j < nq; ++j) {
                const PS L = add_half(pe(f + 1, j), gx[(size_t)(f - i0) * nq + j]);
                const PS R = sub_half(pe(f + 2, j), gx[(size_t)(f + 1 - i0) * nq + j]);
                Cons F; double g;
                if (B.literal & LIT_FLUX) riemann_literal(h, L.P, R.P, L.s, R.s, 0, &F, &g);
                else riemann(h, L.P, R.P, L.s, R.s, 0, &F, &g);
                const double A = B.Ar[(size_t)f * nq + j];
                Cons& o = Fx[(size_t)(f - i0) * nq + j];
                o.d = F.d * c * A; o.sr = F.sr * c * A; o.sq = F.sq * c * A; o.tau = F.tau * c * A;
                Gx[(size_t)(f - i0) * nq + j] = g * c * A;
            }
        std::vector<PS> gy(nq);
        std::vector<Cons> Fy(nq + 1);
        std::vector<double> Gy(nq + 1);
        for (int i = i0; i < i1; ++i) {
            const int k = i + 2;
            if (nq > 1) {
                // Polar slopes with zero gradient at the pole rows (scheme.rs:107-110).
                gy[0] = PS{{0, 0, 0, 0}, 0};
                gy[nq - 1] = PS{{0, 0, 0, 0}, 0};
                for (int j = 1; j + 1 < nq; ++j) gy[j] = plm5(th, pe(k, j - 1), pe(k, j), pe(k, j + 1));
                Fy[0] = Cons{0, 0, 0, 0}; Gy[0] = 0.0;
                Fy[nq] = Cons{0, 0, 0, 0}; Gy[nq] = 0.0;
                for (int j = 1; j < nq; ++j) {
                    const PS L = add_half(pe(k, j - 1), gy[j - 1]);
                    const PS R = sub_half(pe(k, j), gy[j]);
                    Cons F; double g;
                    if (B.literal & LIT_FLUX) riemann_literal(h, L.P, R.P, L.s, R.s, 1, &F, &g);
                    else riemann(h, L.P, R.P, L.s, R.s, 1, &F, &g);
                    const double A = B.Aq[(size_t)i * (nq + 1) + j];
                    Fy[j] = Cons{F.d * c * A, F.sr * c * A, F.sq * c * A, F.tau * c * A};
                    Gy[j] = g * c * A;
                }
            }
            for (int j = 0; j < nq; ++j) {
                const size_t z = (size_t)i * nq + j;
                const Prim& P = prim[b][z].P;
                const Cons S = (B.literal & LIT_SOURCES) ? sources_literal(h, P, B.rc[z], B.qc[z]) : sources(h, P, B.rinv[z], B.cot[z]);
                const double dv = B.dV[z];
                const Cons sc{S.d * c * dv, S.sr * c * dv, S.sq * c * dv, S.tau * c * dv};
                const Cons& a = Fx[(size_t)(i - i0) * nq + j];
                const Cons& bb = Fx[(size_t)(i + 1 - i0) * nq + j];
                const double ga = Gx[(size_t)(i - i0) * nq + j];
                const double gb = Gx[(size_t)(i + 1 - i0) * nq + j];
                Cons du;
                double ds;
                if (nq == 1) {   // scheme.rs:99-105
                    du.d = (sc.d - (bb.d - a.d)) * dt;
                    du.sr = (sc.sr - (bb.sr - a.sr)) * dt;
                    du.sq = (sc.sq - (bb.sq - a.sq)) * dt;
                    du.tau = (sc.tau - (bb.tau - a.tau)) * dt;
                    ds = (gb - ga) * -dt;
                } else {         // scheme.rs:136-149
                    const Cons& cc = Fy[j];
                    const Cons& dd = Fy[j + 1];
                    du.d = (sc.d - (bb.d - a.d) - (dd.d - cc.d)) * dt;
                    du.sr = (sc.sr - (bb.sr - a.sr) - (dd.sr - cc.sr)) * dt;
                    du.sq = (sc.sq - (bb.sq - a.sq) - (dd.sq - cc.sq)) * dt;
                    du.tau = (sc.tau - (bb.tau - a.tau) - (dd.tau - cc.tau)) * dt;
                    ds = ((gb - ga) + (Gy[j + 1] - Gy[j])) * -dt;
                }
                const Cons& u = Uin[b][z];
                Cons un{u.d + du.d, u.sr + du.sr, u.sq + du.sq, u.tau + du.tau};
                double mn = Min[b][z] + ds;
                if (bweight > 0.0) {
                    const double w1 = -bweight + 1.0;
                    const Cons& u0 = (*U0)[b][z];
                    un = Cons{un.d * w1 + u0.d * bweight, un.sr * w1 + u0.sr * bweight,
                              un.sq * w1 + u0.sq * bweight, un.tau * w1 + u0.tau * bweight};
                    mn = mn * w1 + (*M0)[b][z] * bweight;
                }
                Uout[b][z] = un;
                Mout[b][z] = mn;
            }
        }
    });
    return ErrRec{};
}

// State::time_step (state.rs:143-162).
ErrRec time_step(Sim& sim, double* dt)
{
    const Hydro& h = sim.h;
    if (h.system == SYS_SRHD && !h.adaptive_time_step) {
        const Block& b = sim.blocks.front();
        *dt = h.cfl_number * (b.rv[1] - b.rv[0]) / h.light_speed;
        return ErrRec{};
    }
    const int nb = (int)sim.blocks.size();
    std::vector<double> mins(nb, std::numeric_limits<double>::max());
    std::vector<ErrRec> errs(nb);
    parallel_for(nb, sim.nthreads, [&](int b) {
        std::vector<PS> prim;
        errs[b] = block_primitives(sim, b, sim.blocks[b].U, sim.blocks[b].M, prim);
        if (errs[b].code != OK) return;
        double m = std::numeric_limits<double>::max();
        const Block& B = sim.blocks[b];
        for (size_t k = 0; k < prim.size(); ++k)
            m = std::min(m, B.dl[k] / (max_signal_speed(h, prim[k].P) * h.light_speed));
        mins[b] = m;
    });
    double m = std::numeric_limits<double>::max();
    for (int b = 0; b < nb; ++b) {
        if (errs[b].code != OK) return errs[b];
        m = std::min(m, mins[b]);
    }
    *dt = m * h.cfl_number;
    return ErrRec{};
}

// RungeKuttaOrder::try_advance with the SSP weights of godunov-core:
// RK1: U1; RK2: avg(L(L(U0)), U0; 1/2); RK3: b = 3/4 then 1/3.
// Refreshes the ghost rows of the sides that have a boundary source, at
// the time of the state the next stage starts from.
// Ghost rows of a model side at time t: BlockState::from_model on the two
// rows of mesh block first-1 (side 0) or last+1 (side 1) next to the live
// mesh, then try_to_primitive and s = M/dV/Gamma (src/scheme.rs:30-52),
// with U/dV and M/dV/Gamma divided as the reference divides.
ErrRec model_rows(Sim& sim, int side, double t, std::vector<PS>& rows)
{
    const om::Mesh& mesh = sim.mesh[side];
    const int nq = sim.blocks.front().nq;
    const std::vector<double>& qv = sim.blocks.front().qv;
    const int index = side ? sim.first_index + (int)sim.blocks.size() : sim.first_index - 1;
    std::vector<double> rv((size_t)mesh.block_size + 1);
    om::block_vertices(mesh, index, rv.data(), nullptr);
    const int i0 = side ? 0 : mesh.block_size - 2;
    rows.assign((size_t)2 * nq, PS{{0, 0, 0, 0}, 0});
    ErrRec err;
    for (int i = 0; i < 2; ++i) {
        const RadialRow R = radial_row(rv[i0 + i], rv[i0 + i + 1]);
        for (int j = 0; j < nq; ++j) {
            const PolarCol C = polar_col(qv[j], qv[j + 1]);
            double P[4], X = 0.0;
            ErrRec e;
            e.block = index;
            e.i = i0 + i;
            e.j = j;
            e.r = R.rc;
            e.theta = 0.5 * (qv[j] + qv[j + 1]);
            if (!om::primitive_at(sim.model[side], R.rc, 0.5 * (qv[j] + qv[j + 1]), t, P, &X)) {
                e.code = 14;
                if (err.code == OK) err = e;
                continue;
            }
            const Cons c = p2c(sim.h, Prim{P[0], P[1], P[2], P[3]});
            const double dv = R.volA * C.dcost;   // cell_volume (mesh.rs:119-123)
            const Cons U{c.d * dv, c.sr * dv, c.sq * dv, c.tau * dv};
            const double M = U.d * X;
            const Cons u{U.d / dv, U.sr / dv, U.sq / dv, U.tau / dv};
            PS q{{0, 0, 0, 0}, 0};
            double rW = 1.0;
            int rc, it = 0;
            if (sim.h.system == SYS_EULER) {
                rc = c2p_euler(u, sim.h.gamma_law_index, &q.P);
                e.value = u.d;
            } else {
                rc = c2p_srhd(u, sim.h.gamma_law_index, 0.0, &q.P, &it, &rW);
                e.value = rc == NEGATIVE_DENSITY ? u.d : u.tau;
            }
            if (rc == NEGATIVE_PRESSURE) rc = OK;
            if (rc != OK) {
                e.code = rc;
                e.u = u;
                if (err.code == OK) err = e;
                continue;
            }
            q.s = M / dv / lorentz_factor(sim.h, q.P);
            rows[(size_t)i * nq + j] = q;
        }
    }
    return err;
}

// Refreshes the ghost rows of the sides with a boundary source or model,
// at the time of the state the next stage starts from.
ErrRec refresh_boundaries(Sim& sim, double t)
{
    for (int side = 0; side < 2; ++side) {
        Boundary& bd = side == 0 ? sim.inner : sim.outer;
        const int nq = sim.blocks.empty() ? 0 : sim.blocks.front().nq;
        if (sim.model_on[side]) {
            ErrRec e = model_rows(sim, side, t, bd.rows);
            if (e.code != OK) return e;
            bd.mode = BND_FIXED;
            continue;
        }
        if (!sim.source[side]) continue;
        std::vector<double> prim((size_t)8 * nq), scalar((size_t)2 * nq);
        if (sim.source[side](sim.source_user[side], side, t, nq, prim.data(), scalar.data())) {
            ErrRec e;
            e.code = 10;
            return e;
        }
        bd.mode = BND_FIXED;
        bd.rows.resize((size_t)2 * nq);
        for (int k = 0; k < 2 * nq; ++k)
            bd.rows[k] = PS{{prim[4 * k], prim[4 * k + 1], prim[4 * k + 2], prim[4 * k + 3]}, scalar[k]};
    }
    return ErrRec{};
}

ErrRec rk_step(Sim& sim, double dt)
{
    const int nb = (int)sim.blocks.size();
    std::vector<std::vector<Cons>> U0(nb), U1, U2;
    std::vector<std::vector<double>> M0(nb), M1, M2;
    for (int b = 0; b < nb; ++b) { U0[b] = sim.blocks[b].U; M0[b] = sim.blocks[b].M; }
    ErrRec e;
    const double t0 = sim.time;
    double t;
    const int order = sim.h.runge_kutta_order;
    // each stage's ghost rows come from the model at its input state's time:
    // t0, t0 + dt, and (RK3) the averaged time of the second stage's output
    if ((e = refresh_boundaries(sim, t0)).code) return e;
    if (order == 1) {
        e = stage(sim, U0, M0, nullptr, nullptr, 0.0, dt, U1, M1);
        if (e.code) return e;
        t = t0 + dt;
    } else if (order == 2) {
        e = stage(sim, U0, M0, nullptr, nullptr, 0.0, dt, U1, M1);
        if (e.code) return e;
        if ((e = refresh_boundaries(sim, t0 + dt)).code) return e;
        e = stage(sim, U1, M1, &U0, &M0, 0.5, dt, U2, M2);
        if (e.code) return e;
        U1.swap(U2); M1.swap(M2);
        const double t2 = (t0 + dt) + dt;
        t = t2 * (-0.5 + 1.0) + t0 * 0.5;
    } else {
        e = stage(sim, U0, M0, nullptr, nullptr, 0.0, dt, U1, M1);
        if (e.code) return e;
        if ((e = refresh_boundaries(sim, t0 + dt)).code) return e;
        e = stage(sim, U1, M1, &U0, &M0, 0.75, dt, U2, M2);
        if (e.code) return e;
        const double t1 = t0 + dt;
        const double t2 = (t1 + dt) * (-0.75 + 1.0) + t0 * 0.75;
        const double b3 = 1.0 / 3.0;
        if ((e = refresh_boundaries(sim, t2)).code) return e;
        e = stage(sim, U2, M2, &U0, &M0, b3, dt, U1, M1);
        if (e.code) return e;
        t = (t2 + dt) * (-b3 + 1.0) + t0 * b3;
    }
    for (int b = 0; b < nb; ++b) { sim.blocks[b].U.swap(U1[b]); sim.blocks[b].M.swap(M1[b]); }
    sim.time = t;
    sim.iteration += 1;
    return ErrRec{};
}

}  // namespace

// ===========================================================================
// extern "C" surface for ctypes (tests and bench.py's CPU baseline only).
// ===========================================================================
extern "C" {

struct ko_error {
    int32_t code, block, i, j;
    double r, theta, value;
    double conserved[4];
};

static void put_err(const ErrRec& e, ko_error* out)
{
    if (!out) return;
    out->code = e.code; out->block = e.block; out->i = e.i; out->j = e.j;
    out->r = e.r; out->theta = e.theta; out->value = e.value;
    out->conserved[0] = e.u.d; out->conserved[1] = e.u.sr; out->conserved[2] = e.u.sq; out->conserved[3] = e.u.tau;
}

// Per-zone trait methods over arrays (traits.rs:41-136).
int ko_try_to_primitive(const Hydro* h, int64_t n, const double* cons, const double* guess,
                        double* prim, int32_t* code, int32_t* iters)
{
    for (int64_t k = 0; k < n; ++k) {
        const Cons u{cons[4 * k], cons[4 * k + 1], cons[4 * k + 2], cons[4 * k + 3]};
        Prim P{0, 0, 0, 0};
        int rc, it = 0;
        double rW = 1.0;
        if (h->system == SYS_EULER) rc = c2p_euler(u, h->gamma_law_index, &P);
        else rc = c2p_srhd(u, h->gamma_law_index, guess ? guess[k] : 0.0, &P, &it, &rW);
        prim[4 * k] = P.rho; prim[4 * k + 1] = P.ur; prim[4 * k + 2] = P.uq; prim[4 * k + 3] = P.p;
        if (code) code[k] = rc;
        if (iters) iters[k] = it;
    }
    return 0;
}

int ko_to_conserved(const Hydro* h, int64_t n, const double* prim, double* cons)
{
    for (int64_t k = 0; k < n; ++k) {
        const Cons u = p2c(*h, Prim{prim[4 * k], prim[4 * k + 1], prim[4 * k + 2], prim[4 * k + 3]});
        cons[4 * k] = u.d; cons[4 * k + 1] = u.sr; cons[4 * k + 2] = u.sq; cons[4 * k + 3] = u.tau;
    }
    return 0;
}

// Scaled by light_speed exactly as intercell_flux (relativistic_hydro.rs:133).
int ko_intercell_flux(const Hydro* h, int64_t n, const double* pl, const double* pr, const double* sl,
                      const double* sr, int32_t axis, double* flux, double* sflux)
{
    for (int64_t k = 0; k < n; ++k) {
        const Prim L{pl[4 * k], pl[4 * k + 1], pl[4 * k + 2], pl[4 * k + 3]};
        const Prim R{pr[4 * k], pr[4 * k + 1], pr[4 * k + 2], pr[4 * k + 3]};
        Cons F; double g;
        riemann(*h, L, R, sl[k], sr[k], axis, &F, &g);
        const double c = h->light_speed;
        flux[4 * k] = F.d * c; flux[4 * k + 1] = F.sr * c; flux[4 * k + 2] = F.sq * c; flux[4 * k + 3] = F.tau * c;
        sflux[k] = g * c;
    }
    return 0;
}

int ko_wavespeeds(const Hydro* h, int64_t n, const double* prim, int32_t axis, double* lm, double* lp)
{
    for (int64_t k = 0; k < n; ++k) {
        const FluxState st = flux_state(*h, Prim{prim[4 * k], prim[4 * k + 1], prim[4 * k + 2], prim[4 * k + 3]}, 0.0, axis);
        lm[k] = st.lm; lp[k] = st.lp;
    }
    return 0;
}

int ko_geometrical_source_terms(const Hydro* h, int64_t n, const double* prim, const double* r,
                                const double* theta, double* out)
{
    for (int64_t k = 0; k < n; ++k) {
        const double cot = std::cos(theta[k]) / std::sin(theta[k]);
        const Cons s = sources(*h, Prim{prim[4 * k], prim[4 * k + 1], prim[4 * k + 2], prim[4 * k + 3]}, 1.0 / r[k], cot);
        const double c = h->light_speed;
        out[4 * k] = s.d * c; out[4 * k + 1] = s.sr * c; out[4 * k + 2] = s.sq * c; out[4 * k + 3] = s.tau * c;
    }
    return 0;
}

int ko_max_signal_speed(const Hydro* h, int64_t n, const double* prim, double* out)
{
    for (int64_t k = 0; k < n; ++k)
        out[k] = max_signal_speed(*h, Prim{prim[4 * k], prim[4 * k + 1], prim[4 * k + 2], prim[4 * k + 3]}) * h->light_speed;
    return 0;
}

int ko_plm_gradient(double theta, int64_t n, const double* a, const double* b, const double* c, double* out)
{
    for (int64_t k = 0; k < n; ++k) out[k] = plm(theta, a[k], b[k], c[k]);
    return 0;
}

int ko_geometry(int nr, int nq, const double* rv, const double* qv, double* Ar, double* Aq, double* dV,
                double* rc, double* cot, double* dl)
{
    Block b;
    b.nr = nr; b.nq = nq;
    b.rv.assign(rv, rv + nr + 1);
    b.qv.assign(qv, qv + nq + 1);
    build_geometry(b);
    std::memcpy(Ar, b.Ar.data(), b.Ar.size() * 8);
    std::memcpy(Aq, b.Aq.data(), b.Aq.size() * 8);
    std::memcpy(dV, b.dV.data(), b.dV.size() * 8);
    if (rc) std::memcpy(rc, b.rc.data(), b.rc.size() * 8);
    if (cot) std::memcpy(cot, b.cot.data(), b.cot.size() * 8);
    if (dl) std::memcpy(dl, b.dl.data(), b.dl.size() * 8);
    return 0;
}

// Whole-mesh simulation: a radially ordered chain of blocks.
// BlockProducts::try_from_block_state (src/products.rs:44-66) for one block:
// try_to_primitive of U/dV per zone (the NegativePressure floor accepted),
// written in AnyPrimitive order (velocity_r, velocity_q, mass_density,
// gas_pressure; src/physics/mod.rs:36-75 and Hydrodynamics::any,
// relativistic_hydro.rs:114-121), and scalar = scalar_mass / lab_frame_mass
// (M / D).  Returns 0, or the first failure (row-major) in err.
int ko_products(const Hydro* h, int nr, int nq, const double* rv, const double* qv, const double* cons,
                const double* scalar_mass, double* prim, double* scalar, ko_error* err)
{
    for (int i = 0; i < nr; ++i) {
        const RadialRow R = radial_row(rv[i], rv[i + 1]);
        for (int j = 0; j < nq; ++j) {
            const PolarCol C = polar_col(qv[j], qv[j + 1]);
            const size_t k = (size_t)i * nq + j;
            const double dv = R.volA * C.dcost;   // cell_volume (mesh.rs:119-123)
            const Cons u{cons[4 * k] / dv, cons[4 * k + 1] / dv, cons[4 * k + 2] / dv, cons[4 * k + 3] / dv};
            Prim P{0, 0, 0, 0};
            int rc, it = 0;
            double rW = 1.0;
            if (h->system == SYS_EULER) rc = c2p_euler(u, h->gamma_law_index, &P);
            else rc = c2p_srhd(u, h->gamma_law_index, 0.0, &P, &it, &rW);
            if (rc == NEGATIVE_PRESSURE) rc = OK;
            if (rc != OK) {
                ErrRec e;
                e.code = rc;
                e.i = i;
                e.j = j;
                e.r = R.rc;
                e.theta = 0.5 * (qv[j] + qv[j + 1]);
                e.value = (h->system == SYS_EULER || rc == NEGATIVE_DENSITY) ? u.d : u.tau;
                e.u = u;
                put_err(e, err);
                return rc;
            }
            prim[4 * k] = P.ur;
            prim[4 * k + 1] = P.uq;
            prim[4 * k + 2] = P.rho;
            prim[4 * k + 3] = P.p;
            scalar[k] = scalar_mass[k] / cons[4 * k];
        }
    }
    return 0;
}

void* ko_sim_create(const Hydro* h, int nthreads)
{
    Sim* s = new Sim();
    s->h = *h;
    s->nthreads = nthreads < 1 ? 1 : nthreads;
    return s;
}

void ko_sim_destroy(void* p) { delete (Sim*)p; }

// Switches the LIT_* parts of a simulation (its blocks, births and model
// rows) between the rewritten forms the device uses and the literal forms.
int ko_sim_set_literal(void* p, int mask)
{
    Sim* s = (Sim*)p;
    s->literal = mask & (LIT_SOURCES | LIT_FLUX);
    for (Block& b : s->blocks) {
        b.literal = s->literal;
        build_geometry(b);
    }
    return 0;
}



// Appends a block at the outer end of the chain.  conserved: (nr,nq,4)
// volume-integrated, as BlockState::conserved (state.rs:26-30).
int ko_sim_push_block(void* p, int nr, int nq, const double* rv, const double* qv, const double* cons,
                      const double* scalar_mass)
{
    Sim* s = (Sim*)p;
    Block b;
    b.nr = nr; b.nq = nq;
    b.literal = s->literal;
    b.rv.assign(rv, rv + nr + 1);
    b.qv.assign(qv, qv + nq + 1);
    build_geometry(b);
    const size_t n = (size_t)nr * nq;
    b.U.resize(n);
    b.M.assign(scalar_mass, scalar_mass + n);
    for (size_t k = 0; k < n; ++k) b.U[k] = Cons{cons[4 * k], cons[4 * k + 1], cons[4 * k + 2], cons[4 * k + 3]};
    s->blocks.push_back(std::move(b));
    return 0;
}

// Removes the innermost block (add_remove_blocks, scheme.rs:191-194).
int ko_sim_pop_inner_block(void* p)
{
    Sim* s = (Sim*)p;
    if (s->blocks.empty()) return -1;
    s->blocks.erase(s->blocks.begin());
    return 0;
}

int ko_sim_num_blocks(void* p) { return (int)((Sim*)p)->blocks.size(); }

// Boundary rows: mode 0 = fixed primitive rows (prim: 2*nq*4, scalar: 2*nq,
// radial order), mode 1 = zero-gradient outflow.
int ko_sim_set_boundary(void* p, int side, int mode, int nq, const double* prim, const double* scalar)
{
    Sim* s = (Sim*)p;
    Boundary& bd = side == 0 ? s->inner : s->outer;
    bd.mode = mode;
    bd.rows.clear();
    if (mode == BND_FIXED) {
        bd.rows.resize((size_t)2 * nq);
        for (int k = 0; k < 2 * nq; ++k)
            bd.rows[k] = PS{{prim[4 * k], prim[4 * k + 1], prim[4 * k + 2], prim[4 * k + 3]}, scalar[k]};
    }
    return 0;
}

// Ghost rows of `side` from a source evaluated at every stage's input time
// (NULL: back to the rows set by ko_sim_set_boundary).
int ko_sim_set_boundary_source(void* p, int side, BoundarySource fn, void* user)
{
    Sim* s = (Sim*)p;
    s->source[side] = fn;
    s->source_user[side] = user;
    return 0;
}

// Model boundary of `side` from a mesh and an initial model (the table is
// copied); m = NULL clears it.
int ko_sim_set_boundary_model(void* p, int side, const om::Model* m, const om::Mesh* mesh)
{
    Sim& sim = *(Sim*)p;
    if (side < 0 || side > 1) return 1;
    if (!m) {
        sim.model_on[side] = false;
        return 0;
    }
    sim.model[side] = *m;
    sim.mesh[side] = *mesh;
    const int rows = m->kind == 5 ? m->table_rows : 0;
    sim.model_table[side].assign(m->table, m->table + (size_t)4 * rows);
    sim.model[side].table = sim.model_table[side].data();
    sim.model[side].table_rows = rows;
    sim.model_on[side] = true;
    return 0;
}

// Block index of the innermost block (for the ghost blocks of models).
int ko_sim_set_first_index(void* p, int index)
{
    ((Sim*)p)->first_index = index;
    return 0;
}

// BlockState::from_model of mesh block `index` at time t
// (src/state.rs:55-69): vertices rv (block_size+1), qv (num_polar_zones+1),
// conserved (nr, nq, 4) = to_conserved(P) dV, scalar_mass = D dV X.
// Returns -1, or the first zone where the model cannot be evaluated.
long long ko_block_from_model(const Hydro* h, const om::Model* m, const om::Mesh* mesh, int index, double t,
                              double* rv, double* qv, double* cons, double* scalar_mass)
{
    om::block_vertices(*mesh, index, rv, qv);
    const int nr = mesh->block_size, nq = mesh->num_polar_zones;
    long long bad = -1;
    for (int i = 0; i < nr; ++i) {
        const RadialRow R = radial_row(rv[i], rv[i + 1]);
        for (int j = 0; j < nq; ++j) {
            const PolarCol C = polar_col(qv[j], qv[j + 1]);
            const size_t k = (size_t)i * nq + j;
            double P[4] = {0, 0, 0, 0}, X = 0.0;
            if (!om::primitive_at(*m, R.rc, 0.5 * (qv[j] + qv[j + 1]), t, P, &X) && bad < 0) bad = (long long)k;
            const Cons c = p2c(*h, Prim{P[0], P[1], P[2], P[3]});
            const double dv = R.volA * C.dcost;
            const double d = c.d * dv;
            cons[4 * k] = d;
            cons[4 * k + 1] = c.sr * dv;
            cons[4 * k + 2] = c.sq * dv;
            cons[4 * k + 3] = c.tau * dv;
            scalar_mass[k] = d * X;
        }
    }
    return bad;
}

// Replaces the state of block bi (same shape).
int ko_sim_set_block(void* p, int bi, const double* cons, const double* scalar_mass)
{
    Sim& sim = *(Sim*)p;
    if (bi < 0 || bi >= (int)sim.blocks.size()) return 1;
    Block& b = sim.blocks[bi];
    const size_t n = (size_t)b.nr * b.nq;
    for (size_t k = 0; k < n; ++k) {
        b.U[k] = Cons{cons[4 * k], cons[4 * k + 1], cons[4 * k + 2], cons[4 * k + 3]};
        b.M[k] = scalar_mass[k];
    }
    return 0;
}

int ko_sim_get_block(void* p, int bi, double* cons, double* scalar_mass)
{
    Sim* s = (Sim*)p;
    const Block& b = s->blocks[bi];
    for (size_t k = 0; k < b.U.size(); ++k) {
        cons[4 * k] = b.U[k].d; cons[4 * k + 1] = b.U[k].sr; cons[4 * k + 2] = b.U[k].sq; cons[4 * k + 3] = b.U[k].tau;
    }
    std::memcpy(scalar_mass, b.M.data(), b.M.size() * 8);
    return 0;
}

int ko_sim_set_time(void* p, double t) { ((Sim*)p)->time = t; return 0; }
double ko_sim_time(void* p) { return ((Sim*)p)->time; }
long long ko_sim_iteration(void* p) { return ((Sim*)p)->iteration; }

int ko_sim_time_step(void* p, double* dt, ko_error* err)
{
    ErrRec e = time_step(*(Sim*)p, dt);
    put_err(e, err);
    return e.code;
}

// nsteps RK steps with a fixed dt (the fold loop of scheme.rs:228-238).
int ko_sim_advance(void* p, double dt, int nsteps, ko_error* err)
{
    Sim* s = (Sim*)p;
    for (int n = 0; n < nsteps; ++n) {
        ErrRec e = rk_step(*s, dt);
        if (e.code != OK) { put_err(e, err); return e.code; }
    }
    put_err(ErrRec{}, err);
    return 0;
}

// scheme::advance without excision: dt from time_step, then fold steps.
int ko_sim_scheme_advance(void* p, int fold, double* dt_out, ko_error* err)
{
    double dt = 0.0;
    int rc = ko_sim_time_step(p, &dt, err);
    if (rc) return rc;
    if (dt_out) *dt_out = dt;
    return ko_sim_advance(p, dt, fold, err);
}

int ko_sim_c2p_stats(void* p, long long* zones, long long* iterations, long long* floors)
{
    Sim* s = (Sim*)p;
    *zones = s->stats.zones; *iterations = s->stats.iterations; *floors = s->stats.floors;
    return 0;
}

}  // extern "C"
