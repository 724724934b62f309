// knc_zone.hpp -- per-zone FP64 arithmetic of the KNC hot path, for gfx950.
//
// Device counterparts of the per-zone methods of the reference's
// `Hydrodynamics` trait (src/traits.rs:41-136) as implemented by
// `RelativisticHydro` (src/physics/relativistic_hydro.rs:42-143) and
// `NewtonianHydro` (src/physics/newtonian_hydro.rs:34-111).  The crates that
// hold the reference arithmetic (hydro-srhd@25bf21ec, hydro-euler@fc929f9b,
// godunov-core@d8dd524a) are not vendored; every formula below is written
// from the published method it names.  Only IEEE-754 double + - * / sqrt and
// comparisons are used, in a fixed order, and the library is compiled with
// -ffp-contract=off, so a zone's result does not depend on tiling, launch
// shape or batching.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace knc {

enum { SYS_EULER = 0, SYS_SRHD = 1 };
enum { RS_HLLE = 0, RS_HLLC = 1 };
enum {
    E_OK = 0,
    E_NEGATIVE_DENSITY = 1,
    E_NEGATIVE_PRESSURE = 2,
    E_NEGATIVE_ENERGY_DENSITY = 3,
    E_ROOT_FINDER_FAILED = 4,
};

constexpr int C2P_MAX_ITER = 50;
constexpr double C2P_TOL = 1e-10;
constexpr double KNC_PI = 3.14159265358979323846;

// Parameters read by the per-zone functions.  gfac = gamma/(gamma-1) is
// computed once on the host with the same expression the formulas use.
struct ZoneParams {
    double gamma;
    double gm1;        // gamma - 1
    double gfac;       // gamma / (gamma - 1)
    double plm_theta;
    double c;          // light speed scale (1 for Euler)
};

struct Prim { double rho, ur, uq, p; };   // SRHD u = Gamma*beta; Euler u = v
struct Cons { double d, sr, sq, tau; };   // SRHD (D,S_r,S_q,tau); Euler (rho,m_r,m_q,E)
struct PS { Prim P; double s; };          // primitive + scalar concentration rho*X

__host__ __device__ inline double dmin(double a, double b) { return (b < a) ? b : a; }
__host__ __device__ inline double dmax(double a, double b) { return (a < b) ? b : a; }

// ---------------------------------------------------------------------------
// Correctly rounded FP64 square root.  For x in [2^-767, DBL_MAX] it is the
// coupled Newton iteration for sqrt(x) and 1/(2 sqrt(x)) (Goldschmidt 1964;
// Markstein, IBM J. Res. Dev. 34 (1990) 111) started from the hardware's
// v_rsq_f64: one joint correction g, h <- g + g r, h + h r with
// r = 1/2 - h g, then two corrections of the root alone, g <- g + (x - g^2) h,
// the last of which rounds once.  This is the operation sequence the
// compiler's sqrt performs on such x; its range scaling (x < 2^-767), the
// zero / infinity pass-through and their selects (seven of its seventeen
// VALU instructions) are taken only by the lanes outside that range.
// Every result is the host's sqrt to the bit (test_fp64_arithmetic_rounds_like_host).
__device__ inline double sqrt_rn_core(double x)
{
    const double y = __builtin_amdgcn_rsq(x);
    double g = x * y;
    double h = y * 0.5;
    const double r = fma(-h, g, 0.5);
    g = fma(g, r, g);
    h = fma(h, r, h);
    double d = fma(-g, g, x);
    g = fma(d, h, g);
    d = fma(-g, g, x);
    return fma(d, h, g);
}

// The high 32 bits of x.  The range tests of the guards below compare them
// with 32-bit literals: a compare of x itself with an FP64 constant reads the
// constant from a pair of scalar registers, which the stage kernel spills.
__host__ __device__ inline uint32_t hi_word(double x) { return (uint32_t)(__builtin_bit_cast(uint64_t, x) >> 32); }

// The range in which sqrt_rn_core is the correctly rounded root,
// [2^-767, DBL_MAX]: the high words 0x10000000 .. 0x7FEFFFFF (zeros,
// subnormals, negatives, infinities and NaN fall outside).
__device__ inline bool sqrt_rn_core_exact(double x) { return hi_word(x) - 0x10000000u <= 0x6FEFFFFFu; }

__device__ inline double sqrt_rn(double x)
{
    double g = sqrt_rn_core(x);
    if (__builtin_expect(!sqrt_rn_core_exact(x), 0)) g = __builtin_sqrt(x);
    return g;
}

// sqrt_rn of a Lorentz factor's square 1 + u^2, which is at least 1 unless
// it is NaN or +inf: one compare (high word at least 0x7FF00000: +inf, NaN,
// or a negative x, which 1 + u^2 is not) selects the compiler's square root
// for those two.
__device__ inline bool sqrt_rn_ge1_outside(double x) { return hi_word(x) >= 0x7FF00000u; }

__device__ inline double sqrt_rn_ge1(double x)
{
    double g = sqrt_rn_core(x);
    if (__builtin_expect(sqrt_rn_ge1_outside(x), 0)) g = __builtin_sqrt(x);
    return g;
}

// x / d, correctly rounded, from y = 1/d correctly rounded: q = x y is within
// one ulp of x/d, the residual x - q d is exact under fma, and q + (x - q d) y
// rounds to x/d (Markstein, IBM J. Res. Dev. 34 (1990) 111, theorem 1).  The
// residual is formed as -(q d - x), which keeps the sign of a zero quotient
// for d > 0.  Three operations instead of a division when y is shared or
// tabulated.  The theorem needs every intermediate in the normal range:
// for d in [2^-400, 2^400] (cell volumes, checked where their tables are
// built) that holds when |q| is in [2^-500, 2^500), the high words of |q|
// 0x20B00000 .. 0x5F2FFFFF, and lanes outside it (tiny scalar tails, zeros
// excepted, infinities, NaN) divide.  Equal to the host's x / d bit for bit
// (test_fp64_arithmetic_rounds_like_host).
__host__ __device__ inline bool div_rcp_outside(double r) { return (hi_word(r) & 0x7FFFFFFFu) - 0x20B00000u >= 0x3E800000u; }

__host__ __device__ inline double div_rcp(double x, double d, double y)
{
    const double q = x * y;
    const double t = fma(q, d, -x);
    double r = fma(-t, y, q);
    if (__builtin_expect(div_rcp_outside(r) && x != 0.0, 0)) r = x / d;
    return r;
}

// The range of d for which div_rcp's three-operation path is exact.
constexpr double DIV_RCP_DMIN = 0x1p-400;
constexpr double DIV_RCP_DMAX = 0x1p400;

// Unguarded forms for a group of operations that takes one branch: each
// returns the fast path's value and sets *slow where that lane needs the
// guarded form (sqrt_rn, sqrt_rn_ge1, div_rcp), which the caller then
// recomputes for the whole group.  Rare lanes only; no lane's result changes.
__device__ inline double sqrt_rn_fast(double x, bool* slow)
{
    *slow = *slow || !sqrt_rn_core_exact(x);
    return sqrt_rn_core(x);
}

__device__ inline double sqrt_rn_ge1_fast(double x, bool* slow)
{
    *slow = *slow || sqrt_rn_ge1_outside(x);
    return sqrt_rn_core(x);
}

__device__ inline double div_rcp_fast(double x, double d, double y, bool* slow)
{
    const double q = x * y;
    const double t = fma(q, d, -x);
    const double r = fma(-t, y, q);
    *slow = *slow || (div_rcp_outside(r) && x != 0.0);
    return r;
}

// ---------------------------------------------------------------------------
// c2p, SRHD: Newton-Raphson on p (Marti & Mueller, Living Rev. Relativ. 6
// (2003) 7, sec. 2.3).  Q = tau + D + p, den = Q^2 - S^2:
//   rho h = den/Q,  rho = D sqrt(den)/Q,  v^2 = S^2/Q^2,
//   f = (gamma-1)(rho h - rho - p) - p,  f' = v^2 c_s^2 - 1,  c_s^2 = gamma p/(rho h).
// With F = Q f = (gamma-1)(den - D sqrt(den) - pQ) - pQ the step is
// dp = F den / (gamma p S^2 - den Q): one division per iteration.
// Converged when |dp| <= 1e-10 (tau + D + p): quadratic convergence then
// leaves the iterate within the rounding noise of f, ~1.5e-16 (tau + D), of
// the root, which c2p_srhd_polish relies on.  Error
// mapping follows relativistic_hydro.rs:68-88.
// ---------------------------------------------------------------------------
// Canonical final steps: round the converged iterate to the power-of-two
// grid 2^(ilogb(tau + D) - 33), which depends on the state only, and take a
// Newton step; then the same on the grid 2^(ilogb(tau + D) - 40).  The
// recovered pressure then does not depend on the starting guess except in
// about one zone in 10^9 (the iterate is only defined to ~eps (tau + D + p)).
__device__ inline double c2p_srhd_newton_step(double D, double S2, double E, const ZoneParams& zp, double p,
                                              int shift)
{
    const int e = ilogb(E) - shift;
    p = ldexp(rint(ldexp(p, -e)), e);
    const double Q = E + p;
    const double den = Q * Q - S2;
    double sq;
    if (__builtin_expect(!sqrt_rn_core_exact(den), 0)) {
        if (!(den > 0.0)) return p;
        sq = __builtin_sqrt(den);
    } else {
        sq = sqrt_rn_core(den);
    }
    const double pQ = p * Q;
    const double F = zp.gm1 * (den - D * sq - pQ) - pQ;
    return p - F * den / (zp.gamma * p * S2 - den * Q);
}

__device__ inline double c2p_srhd_polish(double D, double S2, double E, const ZoneParams& zp, double p)
{
    if (!(E > 0.0)) return p;
    p = c2p_srhd_newton_step(D, S2, E, zp, p, 33);
    return c2p_srhd_newton_step(D, S2, E, zp, p, 40);
}

// The loop only iterates; lanes leave it as they converge or fail, and the
// polish step runs once per wavefront after the loop rather than once per
// distinct convergence iteration.
__device__ inline int c2p_srhd_newton(double D, double S2, double tau, const ZoneParams& zp,
                                      double p, double* p_out, int* iters)
{
    const double E = tau + D;
    int code = E_ROOT_FINDER_FAILED;
    int n = 1;
    for (; n <= C2P_MAX_ITER; ++n) {
        const double Q = E + p;
        const double den = Q * Q - S2;
        // one branch for the exits and the square root's slow lanes
        double sq;
        if (__builtin_expect(!sqrt_rn_core_exact(den), 0)) {
            if (!(den > 0.0)) break;
            sq = __builtin_sqrt(den);
        } else {
            sq = sqrt_rn_core(den);
        }
        const double pQ = p * Q;
        const double F = zp.gm1 * (den - D * sq - pQ) - pQ;
        const double dp = F * den / (zp.gamma * p * S2 - den * Q);
        p = p - dp;
        const bool finite = fabs(p) < 1e300;
        if (!finite || fabs(dp) <= C2P_TOL * (E + p)) {
            if (finite) code = E_OK;
            break;
        }
    }
    if (code != E_OK) {
        *iters = n > C2P_MAX_ITER ? C2P_MAX_ITER : n;
        return code;
    }
    *iters = n + 2;
    *p_out = c2p_srhd_polish(D, S2, E, zp, p);
    return E_OK;
}

// Primitive from a converged pressure: rho = D sqrt(den)/Q, u = S/sqrt(den);
// *rW receives 1/Gamma = sqrt(den)/Q.  Returns E_NEGATIVE_PRESSURE with the
// 1e-3 rho floor applied (relativistic_hydro.rs:79-81).
__device__ inline int c2p_srhd_finish(const Cons& u, double S2, double p, Prim* out, double* rW)
{
    const double Q = (u.tau + u.d) + p;
    const double den = Q * Q - S2;
    const double sq = sqrt_rn(den);
    const double r = 1.0 / (Q * sq);
    const double rsq = Q * r;
    *rW = den * r;
    out->rho = u.d * *rW;
    out->ur = u.sr * rsq;
    out->uq = u.sq * rsq;
    out->p = p;
    if (p < 0.0) { out->p = 1e-3 * out->rho; return E_NEGATIVE_PRESSURE; }
    return E_OK;
}

__device__ inline int c2p_euler(const Cons& u, const ZoneParams& zp, Prim* out)
{
    if (u.d < 0.0) return E_NEGATIVE_DENSITY;
    const double rho = u.d;
    const double rr = 1.0 / rho;
    const double vr = u.sr * rr;
    const double vq = u.sq * rr;
    const double ek = (0.5 * rho) * (vr * vr + vq * vq);
    out->rho = rho;
    out->ur = vr;
    out->uq = vq;
    out->p = (u.tau - ek) * zp.gm1;
    return E_OK;
}

// SRHD conserved state from the primitive and its Lorentz factor
// W = sqrt(1 + u_r^2 + u_q^2).
__device__ inline Cons p2c_srhd(const Prim& p, double W, const ZoneParams& zp)
{
    Cons u;
    const double rhoh = p.rho + p.p * zp.gfac;
    const double hw = rhoh * W;
    u.d = p.rho * W;
    u.sr = hw * p.ur;
    u.sq = hw * p.uq;
    u.tau = (hw * W - p.p) - u.d;
    return u;
}

template <int SYS>
__device__ inline Cons p2c(const Prim& p, const ZoneParams& zp)
{
    Cons u;
    if (SYS == SYS_EULER) {
        u.d = p.rho;
        u.sr = p.rho * p.ur;
        u.sq = p.rho * p.uq;
        u.tau = p.p / zp.gm1 + (0.5 * p.rho) * (p.ur * p.ur + p.uq * p.uq);
    } else {
        u = p2c_srhd(p, sqrt_rn_ge1(1.0 + (p.ur * p.ur + p.uq * p.uq)), zp);
    }
    return u;
}

// Conserved state, physical flux and outer wave speeds along one axis.  The
// speeds also come as fractions lm = nm/den, lp = np/den with den > 0: the
// HLLE flux combines them with a single division.
struct FluxState {
    Cons u, f;
    double us, fs, vn, lm, lp, p;
    double nm, np, den;
};

// FAST: the square roots take their fast path and *slow records the lanes
// that need the guarded form (the caller recomputes those with FAST false).
template <int SYS, int RS, int AXIS, bool FAST = false>
__device__ inline FluxState flux_state(const Prim& P, double s, const ZoneParams& zp, bool* slow = nullptr)
{
    FluxState st;
    const double un = AXIS == 0 ? P.ur : P.uq;
    st.p = P.p;
    if (SYS == SYS_EULER) {
        st.u = p2c<SYS>(P, zp);
        st.vn = un;
        st.us = s;
        const double cs = FAST ? sqrt_rn_fast(zp.gamma * P.p / P.rho, slow) : sqrt_rn(zp.gamma * P.p / P.rho);
        st.lm = st.vn - cs;
        st.lp = st.vn + cs;
        st.nm = st.lm;
        st.np = st.lp;
        st.den = 1.0;
        st.f.d = st.u.d * st.vn;
        st.f.sr = st.u.sr * st.vn;
        st.f.sq = st.u.sq * st.vn;
        if (AXIS == 0) st.f.sr = st.f.sr + P.p; else st.f.sq = st.f.sq + P.p;
        st.f.tau = (st.u.tau + P.p) * st.vn;
        st.fs = st.us * st.vn;
        return st;
    }
    const double uu = P.ur * P.ur + P.uq * P.uq;
    const double W = FAST ? sqrt_rn_ge1_fast(1.0 + uu, slow) : sqrt_rn_ge1(1.0 + uu);
    st.u = p2c_srhd(P, W, zp);
    const double rhoh = P.rho + P.p * zp.gfac;
    const double ut = AXIS == 0 ? P.uq : P.ur;
    // Outer characteristic speeds of Marti & Mueller (2003) eq. (72) with
    // numerator and denominator multiplied by W^2 rho h (gp = gamma p,
    // e = rho h - gp): lambda = (u_n W e -/+ sqrt(gp (rho h + u_t^2 e))) /
    // (rho h + u^2 e).  HLLE keeps them as fractions; HLLC divides once for
    // both speeds.
    const double gp = zp.gamma * P.p;
    const double e = rhoh - gp;
    const double Y = rhoh + uu * e;
    const double k0 = FAST ? sqrt_rn_fast(gp * (rhoh + (ut * ut) * e), slow) : sqrt_rn(gp * (rhoh + (ut * ut) * e));
    const double b = (un * W) * e;
    st.nm = b - k0;
    st.np = b + k0;
    st.den = Y;
    if (RS == RS_HLLC) {
        const double rY = 1.0 / Y;
        st.lm = st.nm * rY;
        st.lp = st.np * rY;
    }
    st.vn = RS == RS_HLLC ? un / W : 0.0;
    st.us = s * W;
    // Fluxes with v_n = u_n/W cancelled against the W of U:
    // F_D = rho u_n, F_S = rho h u u_n + p n, F_tau = (rho h W - rho) u_n, F_s = s u_n.
    st.f.d = P.rho * un;
    st.f.sr = (rhoh * P.ur) * un;
    st.f.sq = (rhoh * P.uq) * un;
    if (AXIS == 0) st.f.sr = st.f.sr + P.p; else st.f.sq = st.f.sq + P.p;
    st.f.tau = (rhoh * W - P.rho) * un;
    st.fs = s * un;
    return st;
}

template <int AXIS> __device__ inline double cnorm(const Cons& u) { return AXIS == 0 ? u.sr : u.sq; }
template <int AXIS> __device__ inline double ctang(const Cons& u) { return AXIS == 0 ? u.sq : u.sr; }

// The two square roots of an SRHD face state and the outer speeds they give
// (flux_state's W, rho h and nm/den, np/den): all the HLLE flux needs beyond
// the primitive itself.
struct SrhdSpeeds {
    double W, rhoh, nm, np, den;
};

template <int AXIS, bool FAST>
__device__ inline SrhdSpeeds srhd_speeds(const Prim& P, const ZoneParams& zp, bool* slow = nullptr)
{
    SrhdSpeeds v;
    const double un = AXIS == 0 ? P.ur : P.uq;
    const double ut = AXIS == 0 ? P.uq : P.ur;
    const double uu = P.ur * P.ur + P.uq * P.uq;
    v.W = FAST ? sqrt_rn_ge1_fast(1.0 + uu, slow) : sqrt_rn_ge1(1.0 + uu);
    v.rhoh = P.rho + P.p * zp.gfac;
    const double gp = zp.gamma * P.p;
    const double e = v.rhoh - gp;
    const double k0 = FAST ? sqrt_rn_fast(gp * (v.rhoh + (ut * ut) * e), slow) : sqrt_rn(gp * (v.rhoh + (ut * ut) * e));
    const double b = (un * v.W) * e;
    v.nm = b - k0;
    v.np = b + k0;
    v.den = v.rhoh + uu * e;
    return v;
}

// SRHD HLLE flux from the primitives and their speeds, each conserved and
// flux component formed where it is consumed (the expressions and their
// order are flux_state's and the HLLE branch of riemann's, so the bits are
// the same); only the speeds stay live across the whole face.
template <int AXIS>
__device__ inline void riemann_hlle_srhd(const PS& Lq, const PS& Rq, const ZoneParams& zp, Cons* F, double* Fs)
{
    bool slow = false;
    SrhdSpeeds L = srhd_speeds<AXIS, true>(Lq.P, zp, &slow);
    SrhdSpeeds R = srhd_speeds<AXIS, true>(Rq.P, zp, &slow);
    if (__builtin_expect(slow, 0)) {
        L = srhd_speeds<AXIS, false>(Lq.P, zp);
        R = srhd_speeds<AXIS, false>(Rq.P, zp);
    }
    const bool left_m = L.nm * R.den <= R.nm * L.den;
    const double nm = left_m ? L.nm : R.nm, dm = left_m ? L.den : R.den;
    const bool left_p = L.np * R.den >= R.np * L.den;
    const double np = left_p ? L.np : R.np, dp = left_p ? L.den : R.den;
    const double n0 = nm < 0.0 ? nm : 0.0, d0 = nm < 0.0 ? dm : 1.0;
    const double n1 = np > 0.0 ? np : 0.0, d1 = np > 0.0 ? dp : 1.0;
    const double c1 = n1 * d0, c0 = n0 * d1, ab = n1 * n0;
    const double rd = 1.0 / (c1 - c0);
    const Prim& a = Lq.P;
    const Prim& b = Rq.P;
    const double una = AXIS == 0 ? a.ur : a.uq, unb = AXIS == 0 ? b.ur : b.uq;
    // p2c_srhd: d = rho W, s = (rho h W) u, tau = ((rho h W) W - p) - d
    const double da = a.rho * L.W, db = b.rho * R.W;
    F->d = ((a.rho * una) * c1 - (b.rho * unb) * c0 + (db - da) * ab) * rd;
    const double hwa = L.rhoh * L.W, hwb = R.rhoh * R.W;
    double fa = (L.rhoh * a.ur) * una, fb = (R.rhoh * b.ur) * unb;
    if (AXIS == 0) { fa = fa + a.p; fb = fb + b.p; }
    F->sr = (fa * c1 - fb * c0 + (hwb * b.ur - hwa * a.ur) * ab) * rd;
    fa = (L.rhoh * a.uq) * una;
    fb = (R.rhoh * b.uq) * unb;
    if (AXIS == 1) { fa = fa + a.p; fb = fb + b.p; }
    F->sq = (fa * c1 - fb * c0 + (hwb * b.uq - hwa * a.uq) * ab) * rd;
    F->tau = (((L.rhoh * L.W - a.rho) * una) * c1 - ((R.rhoh * R.W - b.rho) * unb) * c0 +
              (((hwb * R.W - b.p) - db) - ((hwa * L.W - a.p) - da)) * ab) * rd;
    *Fs = ((Lq.s * una) * c1 - (Rq.s * unb) * c0 + (Rq.s * R.W - Lq.s * L.W) * ab) * rd;
}

// HLLE / HLLC Godunov flux with passive scalar (intercell_flux,
// relativistic_hydro.rs:123-134, newtonian_hydro.rs:96-102), unscaled.
// HLLE: Harten-Lax-van Leer-Einfeldt.  HLLC SRHD: Mignone & Bodo, MNRAS 364
// (2005) 126, eqs. (16)-(18).  HLLC Euler: Toro (2009) sec. 10.4.
template <int SYS, int RS, int AXIS>
__device__ inline void riemann(const PS& Lq, const PS& Rq, const ZoneParams& zp, Cons* F, double* Fs)
{
    if (SYS == SYS_SRHD && RS == RS_HLLE) {
        riemann_hlle_srhd<AXIS>(Lq, Rq, zp, F, Fs);
        return;
    }
    // both states' square roots on their fast path, one branch for the rare
    // lanes that need the guarded forms
    bool slow = false;
    FluxState L = flux_state<SYS, RS, AXIS, true>(Lq.P, Lq.s, zp, &slow);
    FluxState R = flux_state<SYS, RS, AXIS, true>(Rq.P, Rq.s, zp, &slow);
    if (__builtin_expect(slow, 0)) {
        L = flux_state<SYS, RS, AXIS>(Lq.P, Lq.s, zp);
        R = flux_state<SYS, RS, AXIS>(Rq.P, Rq.s, zp);
    }

    if (RS == RS_HLLE) {
        // a0 = min(lm_L, lm_R, 0) = n0/d0 and a1 = max(lp_L, lp_R, 0) = n1/d1,
        // chosen by cross-multiplication (positive denominators); the flux
        // (a1 F_L - a0 F_R + a1 a0 (U_R - U_L)) / (a1 - a0) is multiplied
        // through by d0 d1, leaving one division.
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
    const double am = dmin(L.lm, R.lm);
    const double ap = dmax(L.lp, R.lp);
    if (am >= 0.0) { *F = L.f; *Fs = L.fs; return; }
    if (ap <= 0.0) { *F = R.f; *Fs = R.fs; return; }

    double lstar;
    double pstar = 0.0;
    if (SYS == SYS_SRHD) {
        const double rd = 1.0 / (ap - am);
        const double ab = ap * am;
        const double ud = (R.u.d * ap - L.u.d * am + L.f.d - R.f.d) * rd;
        const double ut = (R.u.tau * ap - L.u.tau * am + L.f.tau - R.f.tau) * rd;
        const double um = (cnorm<AXIS>(R.u) * ap - cnorm<AXIS>(L.u) * am + cnorm<AXIS>(L.f) - cnorm<AXIS>(R.f)) * rd;
        const double fd = (L.f.d * ap - R.f.d * am + (R.u.d - L.u.d) * ab) * rd;
        const double ft = (L.f.tau * ap - R.f.tau * am + (R.u.tau - L.u.tau) * ab) * rd;
        const double fm = (cnorm<AXIS>(L.f) * ap - cnorm<AXIS>(R.f) * am + (cnorm<AXIS>(R.u) - cnorm<AXIS>(L.u)) * ab) * rd;
        const double Ehll = ut + ud;
        const double FEhll = ft + fd;
        const double b = Ehll + fm;
        const double disc = b * b - 4.0 * FEhll * um;
        lstar = 2.0 * um / (b + sqrt_rn(disc));
        pstar = fm - FEhll * lstar;
    } else {
        const double dl = L.u.d * (am - L.vn);
        const double dr = R.u.d * (ap - R.vn);
        lstar = (R.p - L.p + dl * L.vn - dr * R.vn) / (dl - dr);
    }
    const bool left = lstar >= 0.0;
    const FluxState K = left ? L : R;
    const double lam = left ? am : ap;
    const double lv = lam - K.vn;
    const double r = 1.0 / (lam - lstar);
    Cons us;
    double uss;
    if (SYS == SYS_SRHD) {
        const double fac = lv * r;
        const double EK = K.u.tau + K.u.d;
        us.d = K.u.d * fac;
        const double snst = (cnorm<AXIS>(K.u) * lv + pstar - K.p) * r;
        const double stst = ctang<AXIS>(K.u) * fac;
        const double Est = (EK * lv + pstar * lstar - K.p * K.vn) * r;
        if (AXIS == 0) { us.sr = snst; us.sq = stst; } else { us.sq = snst; us.sr = stst; }
        us.tau = Est - us.d;
        uss = K.us * fac;
    } else {
        const double fac = K.u.d * lv * r;
        const double rhoK = K.u.d;
        const double vt = ctang<AXIS>(K.u) / rhoK;
        const double en = K.u.tau / rhoK + (lstar - K.vn) * (lstar + K.p / (rhoK * lv));
        us.d = fac;
        if (AXIS == 0) { us.sr = fac * lstar; us.sq = fac * vt; }
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

// Spherical-polar geometric sources per unit volume at the zone centre
// (geometrical_source_terms, relativistic_hydro.rs:136-138,
// newtonian_hydro.rs:104-106), given 1/r; mass and energy sources are zero.
template <int SYS>
__device__ inline void sources(const Prim& P, double rinv, double cot, const ZoneParams& zp, double* sr, double* sq)
{
    const double rhoh = SYS == SYS_EULER ? P.rho : P.rho + P.p * zp.gfac;
    *sr = (2.0 * P.p + rhoh * (P.uq * P.uq)) * rinv;
    *sq = (P.p * cot - rhoh * (P.ur * P.uq)) * rinv;
}

// max_signal_speed (relativistic_hydro.rs:98-100 scales it by c; the scale is
// applied by the caller).  SRHD: relativistic sum of |v| and c_s.
template <int SYS>
__device__ inline double max_signal_speed(const Prim& P, const ZoneParams& zp)
{
    if (SYS == SYS_EULER)
        return sqrt_rn(P.ur * P.ur + P.uq * P.uq) + sqrt_rn(zp.gamma * P.p / P.rho);
    const double uu = P.ur * P.ur + P.uq * P.uq;
    const double rhoh = P.rho + P.p * zp.gfac;
    const double cs = sqrt_rn(zp.gamma * P.p / rhoh);
    const double v = sqrt_rn(uu / (1.0 + uu));
    return (v + cs) / (1.0 + v * cs);
}

// Generalised minmod (godunov-core piecewise_linear::plm_gradient): all
// three candidate slopes positive <=> their minimum is positive, all
// negative <=> their maximum is negative, otherwise zero (v_min/v_max_f64).
__device__ inline double plm(double th, double a, double b, double c)
{
    const double x = (b - a) * th;
    const double y = (c - a) * 0.5;
    const double z = (c - b) * th;
    const double lo = fmin(x, fmin(y, z));
    const double hi = fmax(x, fmax(y, z));
    // lo > 0 ? lo : (hi < 0 ? hi : 0) as the median of lo, 0 and hi
    // (lo <= hi), max(lo, min(hi, 0)), plus +0 so that a zero comes out +0
    // as the select returns it: one min, one max and one add instead of two
    // compares and four 32-bit selects, the same bits for every input.
    return fmax(lo, fmin(hi, 0.0)) + 0.0;
}

__device__ inline PS plm5(double th, const PS& a, const PS& b, const PS& c)
{
    PS g;
    g.P.rho = plm(th, a.P.rho, b.P.rho, c.P.rho);
    g.P.ur = plm(th, a.P.ur, b.P.ur, c.P.ur);
    g.P.uq = plm(th, a.P.uq, b.P.uq, c.P.uq);
    g.P.p = plm(th, a.P.p, b.P.p, c.P.p);
    g.s = plm(th, a.s, b.s, c.s);
    return g;
}

__device__ inline PS add_half(const PS& a, const PS& g)
{
    PS r;
    r.P.rho = a.P.rho + g.P.rho * 0.5;
    r.P.ur = a.P.ur + g.P.ur * 0.5;
    r.P.uq = a.P.uq + g.P.uq * 0.5;
    r.P.p = a.P.p + g.P.p * 0.5;
    r.s = a.s + g.s * 0.5;
    return r;
}

__device__ inline PS sub_half(const PS& a, const PS& g)
{
    PS r;
    r.P.rho = a.P.rho - g.P.rho * 0.5;
    r.P.ur = a.P.ur - g.P.ur * 0.5;
    r.P.uq = a.P.uq - g.P.uq * 0.5;
    r.P.p = a.P.p - g.P.p * 0.5;
    r.s = a.s - g.s * 0.5;
    return r;
}

}  // namespace knc
