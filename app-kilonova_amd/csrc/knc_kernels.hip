// knc_kernels.hip -- gfx950 kernels of the KNC per-block RK stage.
//
// k_stage: one RK stage (src/scheme.rs:16-173) fused end to end.  A
// workgroup owns a work item (rows [i0,i1) of one block) and one thread per
// polar column; it marches outward in r.  Each step it recovers the
// primitives of one new row (c2p), forms the radial PLM slope and the radial
// Riemann flux of the face behind it (everything radial stays in the
// thread's registers), the polar slope and face states of the row behind,
// and finishes the zone row two behind: polar Riemann fluxes from face
// states exchanged through LDS, geometric sources, flux divergence, the
// update U + dt L(U) and, on the last RK stage, the weighted average with
// U^n (src/state.rs:179-212), with one workgroup barrier per row.  Each
// zone's c2p runs once per work item plus two halo rows on either side;
// nothing but U (and the per-zone geometry) is read from HBM, and only U is
// written.
//
// k_prim: the c2p pass of the two-pass stage form, which with DT also forms
// the CFL minima of the run's first stage.  k_dt: the CFL reduction of
// State::time_step (src/state.rs:143-162) as a pass of its own: c2p,
// dl / (max_signal_speed * c), wavefront min by cross-lane shuffles, one
// partial minimum per workgroup, reduced by k_dt_finish.
#include <hip/hip_runtime.h>
#include <float.h>

#include "knc_device.hpp"
#include "knc_kernels.hpp"
#include "knc_mlp.hpp"
#include "knc_zone.hpp"

namespace knc {

// Loads and stores through the global address space.  Pointers taken from
// the block table in device memory are generic; accessing them as global
// lets the compiler emit global_load/global_store, where flat accesses would
// also count against the LDS wait counter and serialise the row prefetch
// with the LDS exchanges.
__device__ inline double gld(const double* p, int64_t k)
{
    return ((const __attribute__((address_space(1))) double*)p)[k];
}

__device__ inline float gldf(const float* p, int64_t k)
{
    return ((const __attribute__((address_space(1))) float*)p)[k];
}

__device__ inline void gst(double* p, int64_t k, double v)
{
    ((__attribute__((address_space(1))) double*)p)[k] = v;
}

__device__ inline void gstf(float* p, int64_t k, float v)
{
    ((__attribute__((address_space(1))) float*)p)[k] = v;
}

// The value in vector registers: the compiler treats it as per-lane from
// here on.  Uniform doubles that the stage reads in every row then occupy
// VGPRs, of which it has spare, instead of scalar registers, of which it
// has none.
__device__ inline double in_vgpr(double x)
{
    asm volatile("" : "+v"(x));
    return x;
}

// Issue priority of a stage wave, set once per row of the march.  A CU holds
// two stage workgroups, and its SIMDs issue for the older of two ready
// waves: left to that, the workgroup dispatched first runs ahead and the
// other finishes alone, one wave per SIMD, for the last fifth of the stage.
// Instead the two take turns: the priority bit alternates on the 100 MHz
// wall clock every 2^11 ticks (20 us), in opposite phase for the first and
// second half of the grid (on MI355X the dispatcher places blockIdx and
// blockIdx + gridDim/2 on one CU when the grid fills two workgroups per CU,
// as a per-workgroup timing of the stage showed; where it does not, the
// hardware's age order decides within each half of the work item), and a
// wave in the first half of its work item ranks above one in the second
// half.  Priorities order issue only; no result depends on them.
__device__ inline void march_priority(int done, int total, unsigned int phase)
{
    const bool turn = (((unsigned int)(wall_clock64() >> 11)) ^ phase) & 1u;
    if (2 * done < total) {
        if (turn) __builtin_amdgcn_s_setprio(3);
        else __builtin_amdgcn_s_setprio(2);
    } else {
        if (turn) __builtin_amdgcn_s_setprio(1);
        else __builtin_amdgcn_s_setprio(0);
    }
}

// c2p of one zone (BlockState::try_to_primitive of U/dV, src/state.rs:76-94,
// and the scalar concentration s = M/dV/Gamma of src/scheme.rs:35, each
// quotient correctly rounded as the reference divides).  u is the
// volume-integrated state, dv the cell volume and rdv = 1/dv correctly
// rounded.
template <int SYS, bool STATS, bool GUESS>
__device__ inline PS zone_primitive(const StageArgs& a, double ud, double usr, double usq, double utau,
                                    double um, double dv, double rdv, int tag, int col, unsigned int* lds_hist,
                                    double guess)
{
    // U/dV on the three-operation path, one branch for lanes outside its range
    bool slow = false;
    Cons u{div_rcp_fast(ud, dv, rdv, &slow), div_rcp_fast(usr, dv, rdv, &slow), div_rcp_fast(usq, dv, rdv, &slow),
           div_rcp_fast(utau, dv, rdv, &slow)};
    if (__builtin_expect(slow, 0))
        u = Cons{div_rcp(ud, dv, rdv), div_rcp(usr, dv, rdv), div_rcp(usq, dv, rdv), div_rcp(utau, dv, rdv)};
    PS out;
    int code;
    double value = u.d;
    double rW = 1.0;
    if (SYS == SYS_EULER) {
        code = c2p_euler(u, a.zp, &out.P);
    } else {
        int it = 0;
        if (u.d < 0.0) {
            code = E_NEGATIVE_DENSITY;
        } else if (u.tau < 0.0) {
            code = E_NEGATIVE_ENERGY_DENSITY;
            value = u.tau;
        } else {
            const double S2 = u.sr * u.sr + u.sq * u.sq;
            double p = 0.0;
            if (GUESS) {
                // Newton from the learned guess; on failure restart from the
                // plain guess (p = 0) and count the fallback.
                code = c2p_srhd_newton(u.d, S2, u.tau, a.zp, guess, &p, &it);
                if (code != E_OK) {
                    int it2 = 0;
                    code = c2p_srhd_newton(u.d, S2, u.tau, a.zp, 0.0, &p, &it2);
                    if (STATS) {
                        atomicAdd(&a.stats->fallbacks, 1ull);
                        atomicAdd(&a.stats->learned_iters, (unsigned long long)it2);
                    }
                    it += it2;
                }
            } else {
                code = c2p_srhd_newton(u.d, S2, u.tau, a.zp, 0.0, &p, &it);
            }
            if (code == E_OK) {
                code = c2p_srhd_finish(u, S2, p, &out.P, &rW);
                if (code == E_NEGATIVE_PRESSURE) {
                    if (STATS) atomicAdd(&a.stats->floors, 1ull);
                    code = E_OK;
                }
            } else {
                value = u.tau;
            }
        }
        if (STATS) atomicAdd(&lds_hist[it < HIST_BINS ? it : HIST_BINS - 1], 1u);
    }
    if (code != E_OK) {
        record_error(a.err, a.step, tag, col, code, value, u);
        out.P = Prim{0.0, 0.0, 0.0, 0.0};
        out.s = 0.0;
        return out;
    }
    // lorentz_factor = lorentz_factor_squared().sqrt() (relativistic_hydro.rs:162-164), 1 for Euler
    bool slow_s = false;
    double m = div_rcp_fast(um, dv, rdv, &slow_s);
    double W = SYS == SYS_EULER ? 1.0 : sqrt_rn_ge1_fast(1.0 + out.P.ur * out.P.ur + out.P.uq * out.P.uq, &slow_s);
    if (__builtin_expect(slow_s, 0)) {
        m = div_rcp(um, dv, rdv);
        if (SYS != SYS_EULER) W = sqrt_rn_ge1(1.0 + out.P.ur * out.P.ur + out.P.uq * out.P.uq);
    }
    out.s = SYS == SYS_EULER ? m : m / W;
    return out;
}

// Where row k (relative to block B; -2 <= k < nr+2) comes from: the block's
// own rows, the last/first two rows of the inner/outer neighbour
// (src/scheme.rs:73-77), or the boundary rows.  k is uniform across the
// workgroup, so the resolution and every branch on it are uniform.
struct RowSrc {
    const double* base;   // first plane at (row, 0): U planes, or boundary prim planes
    const float* guess;   // guess row (block rows of guess-reading stages)
    const double* rdv;    // 1/dV row (conserved rows)
    double volA;          // volA of the row (conserved rows): dV = volA dcost
    int32_t np;           // plane stride (nr nq < 2^26, or 2 nq for boundary rows)
    int32_t tag;          // (block order << 16) | row, or -1 for primitive rows (no c2p)
};

__device__ inline RowSrc fixed_rows(const double* rows, int k2, int nq)
{
    RowSrc r;
    r.base = rows + (int64_t)k2 * nq;
    r.guess = nullptr;
    r.rdv = nullptr;
    r.volA = 1.0;
    r.np = 2 * nq;
    r.tag = -1;
    return r;
}

// Row k of block S: its conserved planes (c2p in the stage), or with PRIMS
// the primitive planes a c2p pass already wrote (no c2p in the stage).
template <bool PRIMS>
__device__ inline RowSrc block_row(const StageArgs& a, const BlockDesc& S, int k)
{
    RowSrc r;
    r.base = (PRIMS ? S.prim : slot_of(S, a.in_slot)) + (int64_t)k * S.nq;
    r.guess = guess_plane(S, a.guess_in) + (int64_t)k * S.nq;
    r.rdv = zg_rdv(S) + (int64_t)k * S.nq;
    r.volA = PRIMS ? 1.0 : gld(rg_volA(S), k);
    r.np = S.nr * S.nq;
    r.tag = PRIMS ? -1 : (S.order << 16) | k;
    return r;
}

template <bool PRIMS>
__device__ inline RowSrc resolve_row(const StageArgs& a, const BlockDesc& B, int k)
{
    if (k < 0) {
        if (B.left >= 0) {
            const BlockDesc N = a.blocks[B.left];
            return block_row<PRIMS>(a, N, N.nr + k);
        }
        if (a.bnd.inner_mode == BND_FIXED) return fixed_rows(a.bnd.inner_rows, k + 2, B.nq);
        return block_row<PRIMS>(a, B, 0);
    }
    if (k >= B.nr) {
        if (B.right >= 0) {
            const BlockDesc N = a.blocks[B.right];
            return block_row<PRIMS>(a, N, k - B.nr);
        }
        if (a.bnd.outer_mode == BND_FIXED) return fixed_rows(a.bnd.outer_rows, k - B.nr, B.nq);
        return block_row<PRIMS>(a, B, B.nr - 1);
    }
    return block_row<PRIMS>(a, B, k);
}

// One zone as loaded from a row: the five planes, 1/dV for conserved rows
// and, for guess-reading stages, the guessed log2(p/D).
struct Raw5 {
    double v[5];
    double rdv;
    float g;
};

template <bool GUESS, bool PRIMS>
__device__ inline Raw5 load_raw(const RowSrc& r, int j)
{
    const double* p = r.base + j;
    Raw5 x;
#pragma unroll
    for (int f = 0; f < 5; ++f) x.v[f] = gld(p, f * r.np);
    x.rdv = !PRIMS && r.tag >= 0 ? gld(r.rdv, j) : 1.0;
    x.g = GUESS && r.tag >= 0 ? gldf(r.guess, j) : 0.0f;
    return x;
}

// Primitive of a loaded row value.
template <int SYS, bool STATS, bool GUESS>
__device__ inline PS raw_primitive(const StageArgs& a, const RowSrc& r, const Raw5& x, int j, unsigned int* lds_hist)
{
    if (r.tag < 0) {
        PS q;
        q.P = Prim{x.v[0], x.v[1], x.v[2], x.v[3]};
        q.s = x.v[4];
        return q;
    }
    const double dv = r.volA * gld(a.th.dcost, j);
    const double guess = GUESS ? mlp_guess_pressure(x.v[0] * x.rdv, x.g) : 0.0;
    return zone_primitive<SYS, STATS, GUESS>(a, x.v[0], x.v[1], x.v[2], x.v[3], x.v[4], dv, x.rdv, r.tag, j,
                                             lds_hist, guess);
}

__device__ inline void lds_put(double* L, int W, int j, const PS& q)
{
    L[j] = q.P.rho;
    L[W + j] = q.P.ur;
    L[2 * W + j] = q.P.uq;
    L[3 * W + j] = q.P.p;
    L[4 * W + j] = q.s;
}

__device__ inline PS lds_get(const double* L, int W, int j)
{
    PS q;
    q.P = Prim{L[j], L[W + j], L[2 * W + j], L[3 * W + j]};
    q.s = L[4 * W + j];
    return q;
}

struct Flux5 { double d, sr, sq, tau, s; };

// (F * c) * A, the order of intercell_flux's scaling followed by the face
// area product (relativistic_hydro.rs:133, scheme.rs:96-97,127-128).
__device__ inline Flux5 scaled(const Cons& F, double Fs, double c, double A)
{
    return Flux5{F.d * c * A, F.sr * c * A, F.sq * c * A, F.tau * c * A, Fs * c * A};
}

// Stage kernel.  NT threads, one polar column each; a work item is rows
// [i0,i1) x columns [j0,j1) of one block (j1 - j0 <= NT).  When the item does
// not span every column, its halo columns j0-2, j0-1, j0 and j1-2 .. j1+1 are
// recovered once per item for all its rows, with the fluxes through its
// right polar face j1 and the left states of its left polar face j0, into
// LDS before the march.  Step k of the march recovers row k, forms the
// radial face below row k-1, the polar slopes and face states of row k-1
// and the polar fluxes and update of row k-2, with one workgroup barrier:
//   * row k's primitives go to an LDS row for row k's slopes in step k+1;
//   * row k-1's slopes use its neighbours' primitives published in step
//     k-1; each column keeps its lower face's right state P - g/2 for step
//     k+1 and publishes its upper face's left state P + g/2 to an LDS row;
//   * row k-2's polar fluxes combine the left states published in step k-1
//     with the right states kept from it, so each zone's slope and face
//     states are formed once;
//   * the barrier publishes the fluxes of row k-2 (and what steps k+1 and
//     k+2 will read); each LDS array alternates between two rows, whose
//     previous contents were read before the barrier of step k-1.
constexpr int MAX_ROWS = 32;
constexpr int NHALO_F = 7;   // j0-2, j0-1, j0 | j1-2, j1-1, j1, j1+1
constexpr int NHALO = 6;     // the warp-specialised form: j0-2, j0-1 | j1-2, j1-1, j1, j1+1

template <int SYS, bool STATS, bool GUESS, bool PRIMS>
__device__ inline PS column_primitive(const StageArgs& a, const BlockDesc& B, int row, int col, unsigned int* lds_hist)
{
    const RowSrc src = block_row<PRIMS>(a, B, row);
    return raw_primitive<SYS, STATS, GUESS>(a, src, load_raw<GUESS && !PRIMS, PRIMS>(src, col), col, lds_hist);
}

__device__ inline void lds_put5(double* d, const PS& q)
{
    d[0] = q.P.rho; d[1] = q.P.ur; d[2] = q.P.uq; d[3] = q.P.p; d[4] = q.s;
}

__device__ inline PS lds_get5(const double* d) { return PS{{d[0], d[1], d[2], d[3]}, d[4]}; }

// PRIMS: the primitives come from the planes written by k_prim (the two-pass
// stage); otherwise the stage recovers them itself (fused stage).
template <int SYS, int RS, bool STATS, int NT, bool GUESS, bool PRIMS>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(2))) k_stage(StageArgs a)
{
    constexpr int WP = NT + 3;                 // primitive row: j0-2, j0-1 | NT own | j1
    __shared__ double ldsPP[2][5 * WP];      // primitive rows, alternating by row parity
    __shared__ double ldsPL[2][5 * NT];      // left states of each column's upper polar face, by row parity
    __shared__ double ldsF[2][5 * NT];       // polar fluxes of each column's lower face, by step parity
    __shared__ double ldsH[NHALO_F * MAX_ROWS * 5];
    __shared__ double ldsR[MAX_ROWS * 5];      // flux through face j1 per row
    __shared__ double ldsL[MAX_ROWS * 5];      // left state of face j0 (column j0-1's upper face) per row
    __shared__ unsigned int lds_hist[HIST_BINS];
    const int tid = threadIdx.x;
    if (STATS) {
        for (int b = tid; b < HIST_BINS; b += NT) lds_hist[b] = 0u;
        __syncthreads();
    }
    const int nitems = a.nitems_ptr ? *a.nitems_ptr : a.nitems;
    // uniform doubles read in every row: vector registers (see in_vgpr)
    a.zp.gamma = in_vgpr(a.zp.gamma);
    a.zp.gm1 = in_vgpr(a.zp.gm1);
    a.zp.gfac = in_vgpr(a.zp.gfac);
    a.zp.plm_theta = in_vgpr(a.zp.plm_theta);
    a.zp.c = in_vgpr(a.zp.c);
    const double dt = in_vgpr(a.dt_dev ? *a.dt_dev : a.dt);
    const double bw = in_vgpr(a.bweight);
    const double th = a.zp.plm_theta;
    const double c = a.zp.c;
    const bool avg = a.avg_slot >= 0;
    const double w1 = -bw + 1.0;
    const PS zero{{0.0, 0.0, 0.0, 0.0}, 0.0};
    const unsigned int phase = 2u * blockIdx.x >= gridDim.x ? 1u : 0u;

    for (int w = blockIdx.x; w < nitems; w += gridDim.x) {
        const WorkItem it = a.items[w];
        const BlockDesc B = a.blocks[it.block];
        const int nq = B.nq;
        const int j0 = it.j0, j1 = it.j1;
        const int nrow = it.i1 - it.i0;
        const int64_t np = (int64_t)B.nr * nq;
        const double* area_r = zg_ar(B);
        const double* area_q = zg_aq(B);
        const double* Uin = slot_of(B, a.in_slot);
        double* Uout = slot_of(B, a.out_slot);
        const double* U0 = slot_of(B, avg ? a.avg_slot : a.in_slot);
        const bool act = j0 + tid < j1;
        const int j = act ? j0 + tid : j1 - 1;
        const bool left_halo = j0 > 0;
        const bool right_halo = j1 < nq;

        if (left_halo || right_halo) {
            // Halo primitives of this item's rows, then the right-face fluxes
            // and the left-face left states.
            __syncthreads();   // the previous item's reads of ldsH, ldsR and ldsL are done
            // uniform trip count: every lane reaches raw_primitive together
            for (int z0 = 0; z0 < NHALO_F * nrow; z0 += NT) {
                const int z = z0 + tid;
                const int zz = z < NHALO_F * nrow ? z : 0;
                const int h = zz % NHALO_F, rr = zz / NHALO_F;
                const int col = h < 3 ? j0 - 2 + h : j1 - 5 + h;
                const bool need = z < NHALO_F * nrow && (h < 3 ? left_halo : right_halo) && col >= 0 && col < nq;
                const PS q0 = column_primitive<SYS, STATS, GUESS, PRIMS>(a, B, it.i0 + rr, need ? col : j, lds_hist);
                if (z < NHALO_F * nrow) lds_put5(ldsH + (h * MAX_ROWS + rr) * 5, need ? q0 : zero);
            }
            __syncthreads();
            for (int rr = tid; rr < nrow; rr += NT) {
                if (right_halo) {
                    PS q[4];
#pragma unroll
                    for (int h = 0; h < 4; ++h) q[h] = lds_get5(ldsH + ((3 + h) * MAX_ROWS + rr) * 5);
                    // face j1 between columns j1-1 (q[1]) and j1 (q[2])
                    PS gl = zero, gr = zero;
                    if (j1 - 1 >= 1) gl = plm5(th, q[0], q[1], q[2]);
                    if (j1 + 1 < nq) gr = plm5(th, q[1], q[2], q[3]);
                    Cons G;
                    double Gs;
                    riemann<SYS, RS, 1>(add_half(q[1], gl), sub_half(q[2], gr), a.zp, &G, &Gs);
                    const Flux5 f = scaled(G, Gs, c, gld(zg_aq(B), (int64_t)(it.i0 + rr) * (nq + 1) + j1));
                    double* d = ldsR + rr * 5;
                    d[0] = f.d; d[1] = f.sr; d[2] = f.sq; d[3] = f.tau; d[4] = f.s;
                }
                if (left_halo) {
                    // column j0-1's slope and its upper face's left state
                    PS q[3];
#pragma unroll
                    for (int h = 0; h < 3; ++h) q[h] = lds_get5(ldsH + (h * MAX_ROWS + rr) * 5);
                    PS g = zero;
                    if (j0 - 1 >= 1 && j0 < nq) g = plm5(th, q[0], q[1], q[2]);
                    lds_put5(ldsL + rr * 5, add_half(q[1], g));
                }
            }
            // the march's barriers publish ldsR and ldsL before their first reads
        }

        // One step of the march.  The rows and fluxes it carries to the next
        // step are passed by role (the step's Pm2, Pm1, new Pk, ...), so that
        // the HLLE loop below, two steps per iteration, swaps the roles of
        // the values carried from one step to the next instead of copying
        // them.
        auto march_step = [&](const int k, const PS& Pm2, const PS& Pm1, PS& Pk, PS& gm2, const Flux5& Fm, Flux5& Fk,
                              const PS& PRm, PS& PRk, const Raw5& cur, Raw5& nxt, const Raw5& rm2, const RowSrc& src,
                              RowSrc& nsrc) {
            march_priority(k - (it.i0 - 2), nrow + 4, phase);
            // Prefetch the next row and this step's update operands.
            nsrc = src;
            if (k + 1 < it.i1 + 2) {
                nsrc = resolve_row<PRIMS>(a, B, k + 1);
                nxt = load_raw<GUESS && !PRIMS, PRIMS>(nsrc, j);
            }
            // This step's face areas: radial face k-1 and the polar faces of
            // row k-2, consumed after the c2p.
            const double ar = k >= it.i0 + 1 ? gld(area_r, (int64_t)(k - 1) * nq + j) : 0.0;
            const bool update = k >= it.i0 + 2;
            const double aqr = update ? gld(area_q, (int64_t)(k - 2) * (nq + 1) + j) : 0.0;
            Raw5 ui, u0;
            if (update) {
                const int64_t o = (int64_t)(k - 2) * nq + j;
                if (PRIMS) {
#pragma unroll
                    for (int f = 0; f < 5; ++f) ui.v[f] = gld(Uin, o + f * np);
                } else {
                    ui = rm2;
                }
                if (avg) {
#pragma unroll
                    for (int f = 0; f < 5; ++f) u0.v[f] = gld(U0, o + f * np);
                }
            }
            Pk = raw_primitive<SYS, STATS, GUESS>(a, src, cur, j, lds_hist);
            if (GUESS && !PRIMS && a.write_guess && act && k >= it.i0 && k < it.i1) {
                // this item owns row k: keep its pressure as the next stage's guess
                gstf(guess_plane(B, a.guess_out), (int64_t)k * nq + j, guess_log2(Pk.P.p, cur.v[0] * cur.rdv));
            }
            Fk = Flux5{0.0, 0.0, 0.0, 0.0, 0.0};
            if (k >= it.i0) {
                const PS g = plm5(th, Pm2, Pm1, Pk);                 // gx[k-1]
                if (k >= it.i0 + 1) {
                    // Radial face k-1, between rows k-2 and k-1 (scheme.rs:81-97).
                    Cons F;
                    double Fs;
                    riemann<SYS, RS, 0>(add_half(Pm2, gm2), sub_half(Pm1, g), a.zp, &F, &Fs);
                    Fk = scaled(F, Fs, c, ar);
                }
                gm2 = g;
            }
            // Row k-1: polar slope (scheme.rs:107-110, zero at the first and
            // last columns) and the states of its two polar faces.  Its
            // neighbours' primitives were published in step k-1.
            PRk = zero;
            if (k >= it.i0 + 1 && k <= it.i1) {
                const double* L = ldsPP[(k - 1) & 1];
                PS gy = zero;
                if (j >= 1 && j + 1 < nq) gy = plm5(th, lds_get(L, WP, tid + 1), Pm1, lds_get(L, WP, tid + 3));
                lds_put(ldsPL[(k - 1) & 1], NT, tid, add_half(Pm1, gy));
                PRk = sub_half(Pm1, gy);
            }
            // Row r = k-2: flux through the lower polar face of column j
            // (scheme.rs:118-128), between column j-1's upper-face left state
            // (published in step k-1) and this column's lower-face right state.
            Flux5 Fy{0.0, 0.0, 0.0, 0.0, 0.0};
            if (update) {
                if (act && j >= 1) {
                    // thread 0 with j >= 1 has a left halo (j0 > 0)
                    const PS Pl = tid == 0 ? lds_get5(ldsL + (k - 2 - it.i0) * 5)
                                           : lds_get(ldsPL[(k - 2) & 1], NT, tid - 1);
                    Cons G;
                    double Gs;
                    riemann<SYS, RS, 1>(Pl, PRm, a.zp, &G, &Gs);
                    Fy = scaled(G, Gs, c, aqr);
                }
                double* F = ldsF[k & 1];
                F[tid] = Fy.d;
                F[NT + tid] = Fy.sr;
                F[2 * NT + tid] = Fy.sq;
                F[3 * NT + tid] = Fy.tau;
                F[4 * NT + tid] = Fy.s;
            }
            // Row k's primitives, for row k's slopes in step k+1, with the
            // item's halo columns j0-2, j0-1 and j1 of that row.
            if (k >= it.i0 && k < it.i1) {
                const int rr = k - it.i0;
                double* Lp = ldsPP[k & 1];
                lds_put(Lp, WP, tid + 2, Pk);
                if (tid < 2 && left_halo) lds_put(Lp, WP, tid, lds_get5(ldsH + (tid * MAX_ROWS + rr) * 5));
                if (tid == NT - 1 && right_halo)
                    lds_put(Lp, WP, NT + 2, lds_get5(ldsH + (5 * MAX_ROWS + rr) * 5));
            }
            __syncthreads();
            if (update) {
                const int r = k - 2, rr = r - it.i0;
                const PS& P = Pm2;
                Flux5 Fn{0.0, 0.0, 0.0, 0.0, 0.0};
                if (j + 1 < nq) {
                    if (j + 1 < j1) {
                        const double* F = ldsF[k & 1];
                        Fn = Flux5{F[tid + 1], F[NT + tid + 1], F[2 * NT + tid + 1], F[3 * NT + tid + 1],
                                   F[4 * NT + tid + 1]};
                    } else {
                        const double* d = ldsR + rr * 5;
                        Fn = Flux5{d[0], d[1], d[2], d[3], d[4]};
                    }
                }
                if (act) {
                    // Sources x dV and the update (scheme.rs:130-158).
                    const double dv = gld(rg_volA(B), r) * gld(a.th.dcost, j);
                    double ssr, ssq;
                    sources<SYS>(P.P, gld(rg_rinv(B), r), gld(a.th.cot, j), a.zp, &ssr, &ssq);
                    const double sc0 = 0.0 * c * dv;
                    const double sc1 = ssr * c * dv;
                    const double sc2 = ssq * c * dv;
                    const double sc3 = 0.0 * c * dv;
                    const double du0 = (sc0 - (Fk.d - Fm.d) - (Fn.d - Fy.d)) * dt;
                    const double du1 = (sc1 - (Fk.sr - Fm.sr) - (Fn.sr - Fy.sr)) * dt;
                    const double du2 = (sc2 - (Fk.sq - Fm.sq) - (Fn.sq - Fy.sq)) * dt;
                    const double du3 = (sc3 - (Fk.tau - Fm.tau) - (Fn.tau - Fy.tau)) * dt;
                    const double ds = ((Fk.s - Fm.s) + (Fn.s - Fy.s)) * -dt;
                    double v0 = ui.v[0] + du0;
                    double v1 = ui.v[1] + du1;
                    double v2 = ui.v[2] + du2;
                    double v3 = ui.v[3] + du3;
                    double v4 = ui.v[4] + ds;
                    if (avg) {
                        v0 = v0 * w1 + u0.v[0] * bw;
                        v1 = v1 * w1 + u0.v[1] * bw;
                        v2 = v2 * w1 + u0.v[2] * bw;
                        v3 = v3 * w1 + u0.v[3] * bw;
                        v4 = v4 * w1 + u0.v[4] * bw;
                    }
                    double* o = Uout + (int64_t)r * nq + j;
                    gst(o, 0, v0);
                    gst(o, np, v1);
                    gst(o, 2 * np, v2);
                    gst(o, 3 * np, v3);
                    gst(o, 4 * np, v4);
                }
            }
        };
        PS X0, X1, X2, gm2, R0, R1;
        Flux5 F0, F1;
        RowSrc S0 = resolve_row<PRIMS>(a, B, it.i0 - 2), S1 = S0;
        // raw rows: the update operand (row k-2), row k-1, the c2p row k and
        // the prefetched row k+1; the fused stage keeps the rows it loaded
        // for the c2p, row k-2 being also the update operand U of step k
        Raw5 RA, RB, C0 = load_raw<GUESS && !PRIMS, PRIMS>(S0, j), C1;
        int k = it.i0 - 2;
        if constexpr (RS == RS_HLLC && (GUESS || PRIMS)) {
            // HLLC's star states leave no registers for a second step in
            // flight beside a guess plane or the two-pass form's operands:
            // one step per iteration, the carried values copied
            for (; k < it.i1 + 2; ++k) {
                march_step(k, X0, X1, X2, gm2, F0, F1, R0, R1, C0, C1, RA, S0, S1);
                F0 = F1;
                X0 = X1;
                X1 = X2;
                R0 = R1;
                if (!PRIMS) {
                    RA = RB;
                    RB = C0;
                }
                C0 = C1;
                S0 = S1;
            }
        } else {
            for (; k + 1 < it.i1 + 2; k += 2) {
                march_step(k, X0, X1, X2, gm2, F0, F1, R0, R1, C0, C1, RA, S0, S1);
                march_step(k + 1, X1, X2, X0, gm2, F1, F0, R1, R0, C1, RA, RB, S1, S0);
                // back to the first step's roles: rows k and k+1 as Pm2 and Pm1,
                // raw rows k, k+1, k+2 as update operand, row k-1 and c2p row
                X1 = X0;
                X0 = X2;
                if (!PRIMS) {
                    RB = C1;
                    const Raw5 t = C0;
                    C0 = RA;
                    RA = t;
                } else {
                    C0 = RA;
                }
            }
            if (k < it.i1 + 2) march_step(k, X0, X1, X2, gm2, F0, F1, R0, R1, C0, C1, RA, S0, S1);
        }
    }
    if (STATS) {
        __syncthreads();
        for (int b = tid; b < HIST_BINS; b += NT)
            if (lds_hist[b]) atomicAdd(&a.stats->hist[b], (unsigned long long)lds_hist[b]);
    }
}

// ---------------------------------------------------------------------------
// Stage form 4: the flux pass of the two-pass stage with one (work item,
// row) unit per workgroup pass instead of a march down the item's rows.  A
// unit reads the primitives of rows r-2 .. r+2 that k_prim wrote, forms the
// radial slopes of rows r-1, r and r+1 and both radial faces of row r (each
// face is formed by the two rows that share it, from the same operands, so
// with the same bits), the polar slopes and faces of row r through LDS, and
// the update.  It reads five primitive rows per zone and solves a third
// Riemann problem per zone, so the march is faster wherever it fills the
// GPU; on meshes of few zones and in per-block launches the rows as units
// fill it, and the autotuner chooses between the two.  The expressions and
// their order are k_stage's.
// ---------------------------------------------------------------------------
__device__ inline PS row_prim(const RowSrc& src, int col)
{
    const Raw5 x = load_raw<false, true>(src, col);
    return PS{{x.v[0], x.v[1], x.v[2], x.v[3]}, x.v[4]};
}

template <int SYS, int RS, int NT>
__global__ void __launch_bounds__(NT) k_stage_rows(StageArgs a)
{
    constexpr int WP = NT + 4;                 // row r: j0-2, j0-1 | NT own | j1, j1+1
    __shared__ double ldsP[5 * WP];
    __shared__ double ldsPL[5 * NT];           // left state of each column's upper polar face
    __shared__ double ldsF[5 * NT];            // polar flux through each column's lower face
    const int tid = threadIdx.x;
    const double dt = a.dt_dev ? *a.dt_dev : a.dt;
    const double bw = a.bweight;
    const double th = a.zp.plm_theta;
    const double c = a.zp.c;
    const bool avg = a.avg_slot >= 0;
    const double w1 = -bw + 1.0;
    const PS zero{{0.0, 0.0, 0.0, 0.0}, 0.0};
    const int nunits = (a.nitems_ptr ? *a.nitems_ptr : a.nitems) * a.rows;
    for (int u = blockIdx.x; u < nunits; u += gridDim.x) {
        const WorkItem it = a.items[u / a.rows];
        const int r = it.i0 + u % a.rows;
        if (r >= it.i1) continue;   // uniform over the workgroup
        const BlockDesc B = a.blocks[it.block];
        const int nq = B.nq;
        const int j0 = it.j0, j1 = it.j1, n = j1 - j0;
        const int64_t np = (int64_t)B.nr * nq;
        const bool act = tid < n;
        const int j = act ? j0 + tid : j1 - 1;
        PS P[5];
#pragma unroll
        for (int m = 0; m < 5; ++m) P[m] = row_prim(resolve_row<true>(a, B, r - 2 + m), j);
        // Radial slopes (scheme.rs:99-105) and the faces below and above row
        // r (scheme.rs:81-97), as the march forms radial face k-1 in step k.
        const PS g0 = plm5(th, P[0], P[1], P[2]);
        const PS g1 = plm5(th, P[1], P[2], P[3]);
        const PS g2 = plm5(th, P[2], P[3], P[4]);
        Cons F;
        double Fs;
        riemann<SYS, RS, 0>(add_half(P[1], g0), sub_half(P[2], g1), a.zp, &F, &Fs);
        const Flux5 Fm = scaled(F, Fs, c, gld(zg_ar(B), (int64_t)r * nq + j));
        riemann<SYS, RS, 0>(add_half(P[2], g1), sub_half(P[3], g2), a.zp, &F, &Fs);
        const Flux5 Fk = scaled(F, Fs, c, gld(zg_ar(B), (int64_t)(r + 1) * nq + j));
        // Row r and the item's halo columns j0-2, j0-1, j1, j1+1 into LDS.
        __syncthreads();   // the previous unit's reads of the LDS rows are done
        if (act) lds_put(ldsP, WP, tid + 2, P[2]);
        if (tid < 4) {
            const int col = tid < 2 ? j0 - 2 + tid : j1 + tid - 2;
            if (col >= 0 && col < nq) lds_put(ldsP, WP, tid < 2 ? tid : n + tid, row_prim(block_row<true>(a, B, r), col));
        }
        __syncthreads();
        // Polar slope of column j (scheme.rs:107-110, zero at the first and
        // last columns) and the states of its two polar faces.
        PS gy = zero;
        if (j >= 1 && j + 1 < nq) gy = plm5(th, lds_get(ldsP, WP, tid + 1), P[2], lds_get(ldsP, WP, tid + 3));
        lds_put(ldsPL, NT, tid, add_half(P[2], gy));
        const PS PR = sub_half(P[2], gy);
        __syncthreads();
        // Flux through the lower polar face of column j (scheme.rs:118-128).
        Flux5 Fy{0.0, 0.0, 0.0, 0.0, 0.0};
        if (act && j >= 1) {
            PS Pl;
            if (tid == 0) {
                // column j0-1 of the left halo: its slope and its upper face's left state
                PS gl = zero;
                if (j0 - 1 >= 1 && j0 < nq) gl = plm5(th, lds_get(ldsP, WP, 0), lds_get(ldsP, WP, 1), P[2]);
                Pl = add_half(lds_get(ldsP, WP, 1), gl);
            } else {
                Pl = lds_get(ldsPL, NT, tid - 1);
            }
            Cons G;
            double Gs;
            riemann<SYS, RS, 1>(Pl, PR, a.zp, &G, &Gs);
            Fy = scaled(G, Gs, c, gld(zg_aq(B), (int64_t)r * (nq + 1) + j));
        }
        ldsF[tid] = Fy.d;
        ldsF[NT + tid] = Fy.sr;
        ldsF[2 * NT + tid] = Fy.sq;
        ldsF[3 * NT + tid] = Fy.tau;
        ldsF[4 * NT + tid] = Fy.s;
        __syncthreads();
        if (act) {
            Flux5 Fn{0.0, 0.0, 0.0, 0.0, 0.0};
            if (j + 1 < nq) {
                if (j + 1 < j1) {
                    Fn = Flux5{ldsF[tid + 1], ldsF[NT + tid + 1], ldsF[2 * NT + tid + 1], ldsF[3 * NT + tid + 1],
                               ldsF[4 * NT + tid + 1]};
                } else {
                    // face j1, between this column and column j1 of the right halo
                    const PS Q = lds_get(ldsP, WP, n + 2);
                    PS gr = zero;
                    if (j1 + 1 < nq) gr = plm5(th, P[2], Q, lds_get(ldsP, WP, n + 3));
                    Cons G;
                    double Gs;
                    riemann<SYS, RS, 1>(add_half(P[2], gy), sub_half(Q, gr), a.zp, &G, &Gs);
                    Fn = scaled(G, Gs, c, gld(zg_aq(B), (int64_t)r * (nq + 1) + j1));
                }
            }
            // Sources x dV and the update (scheme.rs:130-158).
            const double dv = gld(rg_volA(B), r) * gld(a.th.dcost, j);
            double ssr, ssq;
            sources<SYS>(P[2].P, gld(rg_rinv(B), r), gld(a.th.cot, j), a.zp, &ssr, &ssq);
            const double sc0 = 0.0 * c * dv;
            const double sc1 = ssr * c * dv;
            const double sc2 = ssq * c * dv;
            const double sc3 = 0.0 * c * dv;
            const double du0 = (sc0 - (Fk.d - Fm.d) - (Fn.d - Fy.d)) * dt;
            const double du1 = (sc1 - (Fk.sr - Fm.sr) - (Fn.sr - Fy.sr)) * dt;
            const double du2 = (sc2 - (Fk.sq - Fm.sq) - (Fn.sq - Fy.sq)) * dt;
            const double du3 = (sc3 - (Fk.tau - Fm.tau) - (Fn.tau - Fy.tau)) * dt;
            const double ds = ((Fk.s - Fm.s) + (Fn.s - Fy.s)) * -dt;
            const int64_t o = (int64_t)r * nq + j;
            const double* Uin = slot_of(B, a.in_slot);
            double v0 = gld(Uin, o) + du0;
            double v1 = gld(Uin, o + np) + du1;
            double v2 = gld(Uin, o + 2 * np) + du2;
            double v3 = gld(Uin, o + 3 * np) + du3;
            double v4 = gld(Uin, o + 4 * np) + ds;
            if (avg) {
                const double* U0 = slot_of(B, a.avg_slot);
                v0 = v0 * w1 + gld(U0, o) * bw;
                v1 = v1 * w1 + gld(U0, o + np) * bw;
                v2 = v2 * w1 + gld(U0, o + 2 * np) * bw;
                v3 = v3 * w1 + gld(U0, o + 3 * np) * bw;
                v4 = v4 * w1 + gld(U0, o + 4 * np) * bw;
            }
            double* out = slot_of(B, a.out_slot) + o;
            gst(out, 0, v0);
            gst(out, np, v1);
            gst(out, 2 * np, v2);
            gst(out, 3 * np, v3);
            gst(out, 4 * np, v4);
        }
    }
}

// ---------------------------------------------------------------------------
// Warp-specialised stage (stage form 3): the same stage as k_stage, with the
// work of a work item's column split between two kinds of wavefront.
//   * flux waves (threads [0, NT), one polar column each) march like k_stage
//     without the c2p: radial PLM slope and radial Riemann flux, the polar
//     slopes and Riemann fluxes of the row two behind, sources and update;
//   * c2p waves (threads [NT, 3NT/2)) recover the primitives of the next row,
//     two columns per lane (col and col + 64 of their 128-column half), so
//     each lane runs two independent Newton chains.
// Primitive rows pass from the c2p waves to the flux waves through a ring
// of four LDS rows, and the polar fluxes through two alternating LDS rows,
// with one workgroup barrier per row: in step k the c2p waves write row k
// while the flux waves use row k-1 (radial) and row k-3 (polar and update).
// Every per-zone operation is k_stage's, in the same order, so the results
// are the same bits; only the schedule differs.  The radial halo rows of an
// item and its halo columns are recovered as in k_stage.
constexpr int WS_RING = 4;
constexpr int WS_MAX_ROWS = 32;

template <int SYS, int RS, bool STATS, int NT, bool GUESS>
__global__ void __launch_bounds__(NT + NT / 2) __attribute__((amdgpu_waves_per_eu(3)))
k_stage_ws(StageArgs a)
{
    constexpr int NTH = NT + NT / 2;
    constexpr int WP = NT + 3;                 // primitive row: 2 left halo, NT own, 1 right halo
    __shared__ double ldsP[WS_RING][5 * WP];   // primitive rows, row k in slot k % WS_RING
    __shared__ double ldsF[2][5 * NT];         // polar face fluxes, alternating by step parity
    __shared__ double ldsH[NHALO * WS_MAX_ROWS * 5];
    __shared__ double ldsR[WS_MAX_ROWS * 5];   // flux through face j1 per row
    __shared__ unsigned int lds_hist[HIST_BINS];
    const int tid = threadIdx.x;
    // the role is uniform over a wavefront (NT is a multiple of 64); reading it
    // from the first lane tells the compiler so, and each wave branches to one role
    const bool flux_role = __builtin_amdgcn_readfirstlane(tid < NT ? 1 : 0) != 0;
    if (STATS) {
        for (int b = tid; b < HIST_BINS; b += NTH) lds_hist[b] = 0u;
        __syncthreads();
    }
    const int nitems = a.nitems_ptr ? *a.nitems_ptr : a.nitems;
    const double dt = a.dt_dev ? *a.dt_dev : a.dt;
    const double th = a.zp.plm_theta;
    const double c = a.zp.c;
    const bool avg = a.avg_slot >= 0;
    const double bw = a.bweight;
    const double w1 = -bw + 1.0;
    const PS zero{{0.0, 0.0, 0.0, 0.0}, 0.0};
    // c2p lanes: columns cl and cl + 64 of their wave's 128-column half
    const int ct = tid - NT;
    const int cl = (ct >> 6) * 128 + (ct & 63);

    for (int w = blockIdx.x; w < nitems; w += gridDim.x) {
        const WorkItem it = a.items[w];
        const BlockDesc B = a.blocks[it.block];
        const int nq = B.nq;
        const int j0 = it.j0, j1 = it.j1;
        const int nrow = it.i1 - it.i0;
        const int64_t np = (int64_t)B.nr * nq;
        const bool left_halo = j0 > 0;
        const bool right_halo = j1 < nq;

        __syncthreads();   // the previous item's last reads of the LDS rows are done
        if (left_halo || right_halo) {
            // Halo primitives of this item's rows (every thread), then the right-face fluxes.
            for (int z0 = 0; z0 < NHALO * nrow; z0 += NTH) {
                const int z = z0 + tid;
                const int zz = z < NHALO * nrow ? z : 0;
                const int h = zz % NHALO, rr = zz / NHALO;
                const int col = h < 2 ? j0 - 2 + h : j1 - 4 + h;
                const bool need = z < NHALO * nrow && (h < 2 ? left_halo : right_halo) && col >= 0 && col < nq;
                const PS q0 = column_primitive<SYS, STATS, GUESS, false>(a, B, it.i0 + rr, need ? col : j1 - 1,
                                                                         lds_hist);
                if (z < NHALO * nrow) {
                    const PS q = need ? q0 : zero;
                    double* d = ldsH + (h * WS_MAX_ROWS + rr) * 5;
                    d[0] = q.P.rho; d[1] = q.P.ur; d[2] = q.P.uq; d[3] = q.P.p; d[4] = q.s;
                }
            }
            __syncthreads();
            if (right_halo) {
                for (int rr = tid; rr < nrow; rr += NTH) {
                    PS q[4];
#pragma unroll
                    for (int h = 0; h < 4; ++h) {
                        const double* d = ldsH + ((2 + h) * WS_MAX_ROWS + rr) * 5;
                        q[h] = PS{{d[0], d[1], d[2], d[3]}, d[4]};
                    }
                    PS gl = zero, gr = zero;
                    if (j1 - 1 >= 1) gl = plm5(th, q[0], q[1], q[2]);
                    if (j1 + 1 < nq) gr = plm5(th, q[1], q[2], q[3]);
                    Cons G;
                    double Gs;
                    riemann<SYS, RS, 1>(add_half(q[1], gl), sub_half(q[2], gr), a.zp, &G, &Gs);
                    const Flux5 f = scaled(G, Gs, c, gld(zg_aq(B), (int64_t)(it.i0 + rr) * (nq + 1) + j1));
                    double* d = ldsR + rr * 5;
                    d[0] = f.d; d[1] = f.sr; d[2] = f.sq; d[3] = f.tau; d[4] = f.s;
                }
            }
            // the march's first barrier publishes ldsR
        }

        if (flux_role) {
            // ---- flux waves -------------------------------------------------
            const double* area_r = zg_ar(B);
            const double* area_q = zg_aq(B);
            const double* Uin = slot_of(B, a.in_slot);
            double* Uout = slot_of(B, a.out_slot);
            const double* U0 = slot_of(B, avg ? a.avg_slot : a.in_slot);
            const bool act = j0 + tid < j1;
            const int j = act ? j0 + tid : j1 - 1;
            PS Pm2, Pm1, gm2;
            Flux5 Fm, Fk;
            Raw5 ui, u0;
            for (int k = it.i0 - 2; k <= it.i1 + 2; ++k) {
                const int kk = k - 1;   // the row whose primitives the last barrier published
                const int r = kk - 2;   // the zone row updated in this step
                const bool live = kk >= it.i0 - 2;
                const bool update = kk >= it.i0 + 2;
                PS Pk = zero;
                if (live) {
                    // Halo columns of row kk-1 (polar two steps from now) into its ring row.
                    const int rh = kk - 1 - it.i0;
                    if (rh >= 0 && rh < nrow) {
                        double* L = ldsP[(kk - 1) & (WS_RING - 1)];
                        if (tid < 2 && left_halo) {
                            const double* d = ldsH + (tid * WS_MAX_ROWS + rh) * 5;
                            lds_put(L, WP, tid, PS{{d[0], d[1], d[2], d[3]}, d[4]});
                        }
                        if (tid == NT - 1 && right_halo) {
                            const double* d = ldsH + (4 * WS_MAX_ROWS + rh) * 5;
                            lds_put(L, WP, NT + 2, PS{{d[0], d[1], d[2], d[3]}, d[4]});
                        }
                    }
                    if (update) {
                        // Zone row r: polar slopes and fluxes (scheme.rs:107-128).
                        const double aqr = gld(area_q, (int64_t)r * (nq + 1) + j);
                        const double* L = ldsP[r & (WS_RING - 1)];
                        const int lp = tid + 2;
                        const PS& P = Pm2;
                        PS gy = zero;
                        if (j >= 1 && j + 1 < nq) gy = plm5(th, lds_get(L, WP, lp - 1), P, lds_get(L, WP, lp + 1));
                        Flux5 Fy{0.0, 0.0, 0.0, 0.0, 0.0};
                        if (act && j >= 1) {
                            const PS Pl = lds_get(L, WP, lp - 1);
                            PS gl = zero;
                            if (j >= 2) gl = plm5(th, lds_get(L, WP, lp - 2), Pl, P);
                            Cons G;
                            double Gs;
                            riemann<SYS, RS, 1>(add_half(Pl, gl), sub_half(P, gy), a.zp, &G, &Gs);
                            Fy = scaled(G, Gs, c, aqr);
                        }
                        double* F = ldsF[k & 1];
                        F[tid] = Fy.d;
                        F[NT + tid] = Fy.sr;
                        F[2 * NT + tid] = Fy.sq;
                        F[3 * NT + tid] = Fy.tau;
                        F[4 * NT + tid] = Fy.s;
                    }
                    Pk = lds_get(ldsP[kk & (WS_RING - 1)], WP, tid + 2);
                    if (kk >= it.i0) {
                        const PS g = plm5(th, Pm2, Pm1, Pk);                 // gx[kk-1]
                        if (kk >= it.i0 + 1) {
                            // Radial face kk-1, between rows kk-2 and kk-1 (scheme.rs:81-97).
                            const double ar = gld(area_r, (int64_t)(kk - 1) * nq + j);
                            Cons F;
                            double Fs;
                            riemann<SYS, RS, 0>(add_half(Pm2, gm2), sub_half(Pm1, g), a.zp, &F, &Fs);
                            Fk = scaled(F, Fs, c, ar);
                        }
                        gm2 = g;
                    }
                    if (update) {
                        // the update operands, in flight across the barrier
                        const int64_t o = (int64_t)r * nq + j;
#pragma unroll
                        for (int f = 0; f < 5; ++f) ui.v[f] = gld(Uin, o + f * np);
                        if (avg) {
#pragma unroll
                            for (int f = 0; f < 5; ++f) u0.v[f] = gld(U0, o + f * np);
                        }
                    }
                }
                __syncthreads();   // step k: ring row k (c2p waves) and the polar fluxes of row r
                if (live) {
                    if (update) {
                        const int rr = r - it.i0;
                        const double* F = ldsF[k & 1];
                        const Flux5 Fy{F[tid], F[NT + tid], F[2 * NT + tid], F[3 * NT + tid], F[4 * NT + tid]};
                        Flux5 Fn{0.0, 0.0, 0.0, 0.0, 0.0};
                        if (j + 1 < nq) {
                            if (j + 1 < j1) {
                                Fn = Flux5{F[tid + 1], F[NT + tid + 1], F[2 * NT + tid + 1], F[3 * NT + tid + 1],
                                           F[4 * NT + tid + 1]};
                            } else {
                                const double* d = ldsR + rr * 5;
                                Fn = Flux5{d[0], d[1], d[2], d[3], d[4]};
                            }
                        }
                        if (act) {
                            // Sources x dV and the update (scheme.rs:130-158).
                            const PS& P = Pm2;
                            const double dv = gld(rg_volA(B), r) * gld(a.th.dcost, j);
                            double ssr, ssq;
                            sources<SYS>(P.P, gld(rg_rinv(B), r), gld(a.th.cot, j), a.zp, &ssr, &ssq);
                            const double sc0 = 0.0 * c * dv;
                            const double sc1 = ssr * c * dv;
                            const double sc2 = ssq * c * dv;
                            const double sc3 = 0.0 * c * dv;
                            const double du0 = (sc0 - (Fk.d - Fm.d) - (Fn.d - Fy.d)) * dt;
                            const double du1 = (sc1 - (Fk.sr - Fm.sr) - (Fn.sr - Fy.sr)) * dt;
                            const double du2 = (sc2 - (Fk.sq - Fm.sq) - (Fn.sq - Fy.sq)) * dt;
                            const double du3 = (sc3 - (Fk.tau - Fm.tau) - (Fn.tau - Fy.tau)) * dt;
                            const double ds = ((Fk.s - Fm.s) + (Fn.s - Fy.s)) * -dt;
                            double v0 = ui.v[0] + du0;
                            double v1 = ui.v[1] + du1;
                            double v2 = ui.v[2] + du2;
                            double v3 = ui.v[3] + du3;
                            double v4 = ui.v[4] + ds;
                            if (avg) {
                                v0 = v0 * w1 + u0.v[0] * bw;
                                v1 = v1 * w1 + u0.v[1] * bw;
                                v2 = v2 * w1 + u0.v[2] * bw;
                                v3 = v3 * w1 + u0.v[3] * bw;
                                v4 = v4 * w1 + u0.v[4] * bw;
                            }
                            double* o = Uout + (int64_t)r * nq + j;
                            gst(o, 0, v0);
                            gst(o, np, v1);
                            gst(o, 2 * np, v2);
                            gst(o, 3 * np, v3);
                            gst(o, 4 * np, v4);
                        }
                    }
                    Fm = Fk;
                    Pm2 = Pm1;
                    Pm1 = Pk;
                }
            }
        } else {
            // ---- c2p waves: rows i0-2 .. i1+1, two columns per lane --------
            const int c0 = j0 + cl, c1 = c0 + 64;
            const bool act0 = c0 < j1, act1 = c1 < j1;
            const int jc0 = act0 ? c0 : j1 - 1, jc1 = act1 ? c1 : j1 - 1;
            RowSrc src = resolve_row<false>(a, B, it.i0 - 2);
            Raw5 cur0 = load_raw<GUESS, false>(src, jc0), cur1 = load_raw<GUESS, false>(src, jc1);
            for (int k = it.i0 - 2; k <= it.i1 + 2; ++k) {
                if (k <= it.i1 + 1) {
                    RowSrc nsrc = src;
                    Raw5 nxt0, nxt1;
                    if (k + 1 <= it.i1 + 1) {
                        nsrc = resolve_row<false>(a, B, k + 1);
                        nxt0 = load_raw<GUESS, false>(nsrc, jc0);
                        nxt1 = load_raw<GUESS, false>(nsrc, jc1);
                    }
                    const PS P0 = raw_primitive<SYS, STATS, GUESS>(a, src, cur0, jc0, lds_hist);
                    const PS P1 = raw_primitive<SYS, STATS, GUESS>(a, src, cur1, jc1, lds_hist);
                    if (GUESS && a.write_guess && k >= it.i0 && k < it.i1) {
                        float* gp = guess_plane(B, a.guess_out) + (int64_t)k * nq;
                        if (act0) gstf(gp, jc0, guess_log2(P0.P.p, cur0.v[0] * cur0.rdv));
                        if (act1) gstf(gp, jc1, guess_log2(P1.P.p, cur1.v[0] * cur1.rdv));
                    }
                    double* L = ldsP[k & (WS_RING - 1)];
                    if (act0) lds_put(L, WP, cl + 2, P0);
                    if (act1) lds_put(L, WP, cl + 66, P1);
                    cur0 = nxt0;
                    cur1 = nxt1;
                    src = nsrc;
                }
                __syncthreads();   // step k, as the flux waves
            }
        }
    }
    if (STATS) {
        __syncthreads();
        for (int b = tid; b < HIST_BINS; b += NTH)
            if (lds_hist[b]) atomicAdd(&a.stats->hist[b], (unsigned long long)lds_hist[b]);
    }
}

// Minimum over the wavefront by cross-lane exchanges.
__device__ inline double wave_min(double v)
{
    for (int off = 32; off >= 1; off >>= 1) {
        const double o = __shfl_xor(v, off, 64);
        v = dmin(v, o);
    }
    return v;
}

// ---------------------------------------------------------------------------
// c2p pass of the two-pass stage: the primitives and scalar concentration of
// every zone of the stage's input into the blocks' primitive planes, each
// zone recovered once.  One (work item, row) unit per workgroup iteration,
// NT columns at a time.  With GUESS, Newton starts from the guess plane
// written by k_guess.  With DT the pass is also the CFL reduction of the
// run's first stage (State::time_step, src/state.rs:143-162, on the same
// state): each zone's dl / (max signal speed x c), as k_dt forms it, is
// folded into a workgroup minimum in partial[blockIdx.x] for k_dt_finish,
// so the run needs no separate dt pass over the state.
// ---------------------------------------------------------------------------
template <int SYS, bool STATS, bool GUESS, int NT, bool DT>
__global__ void __launch_bounds__(NT) k_prim(StageArgs a, double* partial)
{
    __shared__ unsigned int lds_hist[HIST_BINS];
    __shared__ double wmin[NT / 64];
    const int tid = threadIdx.x;
    if (STATS) {
        for (int b = tid; b < HIST_BINS; b += NT) lds_hist[b] = 0u;
        __syncthreads();
    }
    double m = DBL_MAX;
    const int nunits = (a.nitems_ptr ? *a.nitems_ptr : a.nitems) * a.rows;
    for (int u = blockIdx.x; u < nunits; u += gridDim.x) {
        const WorkItem it = a.items[u / a.rows];
        const int i = it.i0 + u % a.rows;
        if (i >= it.i1) continue;
        const BlockDesc B = a.blocks[it.block];
        const int nq = B.nq;
        const int64_t np = (int64_t)B.nr * nq;
        const double* U = slot_of(B, a.in_slot) + (int64_t)i * nq;
        double* P = B.prim + (int64_t)i * nq;
        const double volA = gld(rg_volA(B), i);
        const double* rdvrow = zg_rdv(B) + (int64_t)i * nq;
        for (int j = it.j0 + tid; j < it.j1; j += NT) {
            const double rdv = gld(rdvrow, j);
            const double dv = volA * gld(a.th.dcost, j);
            const double u0 = gld(U, j), u1 = gld(U, np + j), u2 = gld(U, 2 * np + j), u3 = gld(U, 3 * np + j);
            const double u4 = gld(U, 4 * np + j);
            const double guess =
                GUESS ? mlp_guess_pressure(u0 * rdv, gldf(guess_plane(B, a.guess_in), (int64_t)i * nq + j)) : 0.0;
            {
                const PS q = zone_primitive<SYS, STATS, GUESS>(a, u0, u1, u2, u3, u4, dv, rdv, (B.order << 16) | i,
                                                               j, lds_hist, guess);
                gst(P, j, q.P.rho);
                gst(P, np + j, q.P.ur);
                gst(P, 2 * np + j, q.P.uq);
                gst(P, 3 * np + j, q.P.p);
                gst(P, 4 * np + j, q.s);
                if (GUESS && a.write_guess)
                    gstf(guess_plane(B, a.guess_out), (int64_t)i * nq + j, guess_log2(q.P.p, u0 * rdv));
                if (DT) {
                    const double dl = dmin(gld(rg_dr(B), i), gld(a.th.dq, j) * gld(rg_rv(B), i));
                    m = dmin(m, dl / (max_signal_speed<SYS>(q.P, a.zp) * a.zp.c));
                }
            }
        }
    }
    if (DT) {
        m = wave_min(m);
        if ((tid & 63) == 0) wmin[tid >> 6] = m;
        __syncthreads();
        if (tid == 0) {
            double r = wmin[0];
            for (int k = 1; k < NT / 64; ++k) r = dmin(r, wmin[k]);
            partial[blockIdx.x] = r;
        }
    }
    if (STATS) {
        __syncthreads();
        for (int b = tid; b < HIST_BINS; b += NT)
            if (lds_hist[b]) atomicAdd(&a.stats->hist[b], (unsigned long long)lds_hist[b]);
    }
}

// ---------------------------------------------------------------------------
// Learned c2p guess pass: before a learned-guess stage, the network's
// log2(p/D) for every zone of the stage's input into the blocks' guess
// planes.  A workgroup walks its (work item row, NT-column pass) steps in
// grid-stride order and issues the loads of the next step before the
// inference of the current one, so the wavefront's HBM latency overlaps
// the MFMA/LDS work.  The inference is wavefront-wide, so every lane runs
// it and lanes past the item's last column evaluate that column without
// storing.
// ---------------------------------------------------------------------------
// One step of a workgroup: unit u = (item, row), pass columns c0.., and
// the item's bounds.  All fields are uniform over the workgroup.
struct GuessStep {
    int u, c0, block, i, j1;
};

// The workgroup's next step after p (c0 < 0: the first step at or after
// unit p.u); false when none is left.
__device__ inline bool guess_advance(const StageArgs& a, int nunits, int nt, GuessStep& p)
{
    int u = p.u, c0 = p.c0;
    if (c0 >= 0) {
        c0 += nt;
        if (c0 < p.j1) {
            p.c0 = c0;
            return true;
        }
        u += gridDim.x;
    }
    for (; u < nunits; u += gridDim.x) {
        const WorkItem it = a.items[u / a.rows];
        const int i = it.i0 + u % a.rows;
        if (i < it.i1) {
            p = GuessStep{u, it.j0, it.block, i, it.j1};
            return true;
        }
    }
    return false;
}

// The raw inputs of one lane for one step: four conserved planes and the
// inverse zone volume.
struct GuessIn {
    double u[4];
    double rdv;
};

__device__ inline GuessIn guess_load(const StageArgs& a, const GuessStep& p)
{
    const BlockDesc B = a.blocks[p.block];
    const int j = p.c0 + (int)threadIdx.x;
    const int jj = j < p.j1 ? j : p.j1 - 1;
    const int64_t np = (int64_t)B.nr * B.nq;
    const double* U = slot_of(B, a.in_slot) + (int64_t)p.i * B.nq + jj;
    GuessIn x;
#pragma unroll
    for (int f = 0; f < 4; ++f) x.u[f] = gld(U, f * np);
    x.rdv = gld(zg_rdv(B), (int64_t)p.i * B.nq + jj);
    return x;
}

// At least four wavefronts per SIMD (at most 128 VGPRs) from 128-thread
// workgroups up: a fourth wavefront hides more of the MFMA -> tanh -> MFMA
// dependency chain than three (122 vs 146 us per pass on bomb_jet
// 4096x1024).  The pass now needs 90 VGPRs and runs five.  A 64-thread
// workgroup is held to two per SIMD by its 20 KiB of LDS anyway.
template <int NT>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(NT >= 128 ? 4 : 1))) k_guess(StageArgs a)
{
    __shared__ MlpShared mlp;
    mlp_shared_load(&mlp, a.mlp);
    const int nunits = (a.nitems_ptr ? *a.nitems_ptr : a.nitems) * a.rows;
    GuessStep p{(int)blockIdx.x, -1, 0, 0, 0};
    if (!guess_advance(a, nunits, NT, p)) return;
    GuessIn cur = guess_load(a, p);
    for (;;) {
        GuessStep q = p;
        const bool more = guess_advance(a, nunits, NT, q);
        GuessIn nxt{};
        if (more) nxt = guess_load(a, q);
        const double D = cur.u[0] * cur.rdv, sr = cur.u[1] * cur.rdv, sq = cur.u[2] * cur.rdv;
        const double tau = cur.u[3] * cur.rdv;
        const float l2 = mlp_log2_guess_wave(a.mlp, &mlp, D, sr * sr + sq * sq, tau, a.zp.gamma);
        const int j = p.c0 + (int)threadIdx.x;
        if (j < p.j1) {
            const BlockDesc& B = a.blocks[p.block];
            gstf(guess_plane(B, a.guess_in), (int64_t)p.i * B.nq + j, l2);
        }
        if (!more) break;
        p = q;
        cur = nxt;
    }
}

// ---------------------------------------------------------------------------
// CFL reduction.
// ---------------------------------------------------------------------------

template <int SYS>
__global__ void __launch_bounds__(256) k_dt(StageArgs a, double* partial)
{
    __shared__ double wmin[4];
    const int nunits = (a.nitems_ptr ? *a.nitems_ptr : a.nitems) * a.rows;
    double m = DBL_MAX;
    for (int u = blockIdx.x; u < nunits; u += gridDim.x) {
        const WorkItem it = a.items[u / a.rows];
        const int i = it.i0 + u % a.rows;
        if (i >= it.i1) continue;
        const BlockDesc B = a.blocks[it.block];
        const int nq = B.nq;
        const int64_t np = (int64_t)B.nr * nq;
        const double volA = gld(rg_volA(B), i);
        const double dr = gld(rg_dr(B), i);
        const double r0 = gld(rg_rv(B), i);
        for (int j = it.j0 + (int)threadIdx.x; j < it.j1; j += blockDim.x) {
            // the four conserved planes only: the scalar plays no part in dt.
            // 1/dV by division here: the pass needs no per-zone table.
            const double* s = slot_of(B, a.in_slot) + (int64_t)i * nq + j;
            const double dv = volA * gld(a.th.dcost, j);
            const PS q = zone_primitive<SYS, false, false>(a, gld(s, 0), gld(s, np), gld(s, 2 * np), gld(s, 3 * np),
                                                           0.0, dv, 1.0 / dv, (B.order << 16) | i, j, nullptr, 0.0);
            const double dl = dmin(dr, gld(a.th.dq, j) * r0);
            m = dmin(m, dl / (max_signal_speed<SYS>(q.P, a.zp) * a.zp.c));
        }
    }
    m = wave_min(m);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) wmin[wid] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = wmin[0];
        for (int k = 1; k < (int)(blockDim.x + 63) / 64; ++k) r = dmin(r, wmin[k]);
        partial[blockIdx.x] = r;
    }
}

// dt = min * cfl (state.rs:160) over the workgroup minima of k_dt; the
// minimum is exact, so the reduction order does not matter.
__global__ void __launch_bounds__(256) k_dt_finish(const double* partial, int n, double cfl, double* dt_out)
{
    __shared__ double wmin[4];
    double m = DBL_MAX;
    for (int k = threadIdx.x; k < n; k += blockDim.x) m = dmin(m, partial[k]);
    m = wave_min(m);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) wmin[wid] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = wmin[0];
        for (int k = 1; k < (int)(blockDim.x + 63) / 64; ++k) r = dmin(r, wmin[k]);
        *dt_out = r * cfl;
    }
}

// ---------------------------------------------------------------------------
// Per-zone trait methods over arrays (AoS, 4 doubles per state), used by the
// C-ABI's batched entry points.
// ---------------------------------------------------------------------------
template <int SYS>
__global__ void k_try_to_primitive(ZoneParams zp, int64_t n, const double* cons, const double* guess, double* prim,
                                   int32_t* code, int32_t* iters)
{
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
        const Cons u{cons[4 * k], cons[4 * k + 1], cons[4 * k + 2], cons[4 * k + 3]};
        Prim P{0.0, 0.0, 0.0, 0.0};
        int rc, it = 0;
        if (SYS == SYS_EULER) {
            rc = c2p_euler(u, zp, &P);
        } else if (u.d < 0.0) {
            rc = E_NEGATIVE_DENSITY;
        } else if (u.tau < 0.0) {
            rc = E_NEGATIVE_ENERGY_DENSITY;
        } else {
            const double S2 = u.sr * u.sr + u.sq * u.sq;
            double p = 0.0, rW = 1.0;
            rc = c2p_srhd_newton(u.d, S2, u.tau, zp, guess ? guess[k] : 0.0, &p, &it);
            if (rc == E_OK) rc = c2p_srhd_finish(u, S2, p, &P, &rW);
        }
        prim[4 * k] = P.rho; prim[4 * k + 1] = P.ur; prim[4 * k + 2] = P.uq; prim[4 * k + 3] = P.p;
        code[k] = rc;
        if (iters) iters[k] = it;
    }
}

template <int SYS>
__global__ void k_to_conserved(ZoneParams zp, int64_t n, const double* prim, double* cons)
{
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
        const Cons u = p2c<SYS>(Prim{prim[4 * k], prim[4 * k + 1], prim[4 * k + 2], prim[4 * k + 3]}, zp);
        cons[4 * k] = u.d; cons[4 * k + 1] = u.sr; cons[4 * k + 2] = u.sq; cons[4 * k + 3] = u.tau;
    }
}

template <int SYS, int RS, int AXIS>
__global__ void k_intercell_flux(ZoneParams zp, int64_t n, const double* pl, const double* pr, const double* sl,
                                 const double* sr, double* flux, double* sflux)
{
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
        PS L, R;
        L.P = Prim{pl[4 * k], pl[4 * k + 1], pl[4 * k + 2], pl[4 * k + 3]};
        R.P = Prim{pr[4 * k], pr[4 * k + 1], pr[4 * k + 2], pr[4 * k + 3]};
        L.s = sl[k];
        R.s = sr[k];
        Cons F;
        double g;
        riemann<SYS, RS, AXIS>(L, R, zp, &F, &g);
        const double c = zp.c;
        flux[4 * k] = F.d * c; flux[4 * k + 1] = F.sr * c; flux[4 * k + 2] = F.sq * c; flux[4 * k + 3] = F.tau * c;
        sflux[k] = g * c;
    }
}

template <int SYS>
__global__ void k_sources(ZoneParams zp, int64_t n, const double* prim, const double* rinv, const double* cot,
                          double* out)
{
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
        double a, b;
        sources<SYS>(Prim{prim[4 * k], prim[4 * k + 1], prim[4 * k + 2], prim[4 * k + 3]}, rinv[k], cot[k], zp, &a, &b);
        const double c = zp.c;
        out[4 * k] = 0.0 * c; out[4 * k + 1] = a * c; out[4 * k + 2] = b * c; out[4 * k + 3] = 0.0 * c;
    }
}

template <int SYS>
__global__ void k_max_signal_speed(ZoneParams zp, int64_t n, const double* prim, double* out)
{
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x)
        out[k] = max_signal_speed<SYS>(Prim{prim[4 * k], prim[4 * k + 1], prim[4 * k + 2], prim[4 * k + 3]}, zp) * zp.c;
}

__global__ void k_plm(double th, int64_t n, const double* a, const double* b, const double* c, double* out)
{
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x)
        out[k] = plm(th, a[k], b[k], c[k]);
}

// Exact FP64 arithmetic probe: a/b, sqrt(|a|) as the stage computes it
// (sqrt_rn_ge1 from 1 up, sqrt_rn below), a*b+c evaluated as two roundings,
// and a/|b| as the stage's U/dV computes it (div_rcp from 1/|b|).
__global__ void k_fp64_probe(int64_t n, const double* a, const double* b, const double* c, double* div,
                             double* sq, double* mad, double* quot)
{
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
        div[k] = a[k] / b[k];
        const double x = fabs(a[k]);
        sq[k] = x >= 1.0 ? sqrt_rn_ge1(x) : sqrt_rn(x);
        mad[k] = a[k] * b[k] + c[k];
        const double d = fabs(b[k]);
        quot[k] = d >= DIV_RCP_DMIN && d <= DIV_RCP_DMAX ? div_rcp(a[k], d, 1.0 / d) : a[k] / d;
    }
}

// FP64 FMA issue-rate probe: every lane runs CH independent chains
// x <- fma(x, b, c) for iters steps (one v_fma_f64 each); the chains' sum
// is stored so that no step is dead.  With one chain a wave issues only
// dependent FMAs, so the rate shows what latency leaves of the issue peak at
// a given number of waves per SIMD.
template <int CH>
__global__ void __launch_bounds__(256) k_fma_probe(int iters, double b, double c, double* out)
{
    double x[CH];
#pragma unroll
    for (int i = 0; i < CH; ++i) x[i] = (double)(threadIdx.x + 64 * i) * 1e-3;
#pragma unroll 4
    for (int n = 0; n < iters; ++n) {
#pragma unroll
        for (int i = 0; i < CH; ++i) x[i] = fma(x[i], b, c);
    }
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < CH; ++i) s += x[i];
    out[(int64_t)blockIdx.x * 256 + threadIdx.x] = s;
}

// ---------------------------------------------------------------------------
// Launch wrappers.
// ---------------------------------------------------------------------------
static int grid_for(int64_t n, int block)
{
    int64_t g = (n + block - 1) / block;
    if (g > 65536) g = 65536;
    if (g < 1) g = 1;
    return (int)g;
}

// Stage kernel variants by workgroup size.  The flux pass of the two-pass
// stage does no c2p, so it has no statistics or guess variants.
template <int SYS, int RS, bool ST, bool G, bool P>
static hipError_t launch_stage_p(const StageArgs& a, int grid, int threads, hipStream_t s)
{
    switch (threads) {
    case 64: hipLaunchKernelGGL((k_stage<SYS, RS, ST, 64, G, P>), dim3(grid), dim3(64), 0, s, a); break;
    case 128: hipLaunchKernelGGL((k_stage<SYS, RS, ST, 128, G, P>), dim3(grid), dim3(128), 0, s, a); break;
    case 256: hipLaunchKernelGGL((k_stage<SYS, RS, ST, 256, G, P>), dim3(grid), dim3(256), 0, s, a); break;
    case 512: hipLaunchKernelGGL((k_stage<SYS, RS, ST, 512, G, P>), dim3(grid), dim3(512), 0, s, a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <int SYS, int RS, bool ST, bool G, bool P>
static int occupancy_p(int threads)
{
    int n = 0;
    hipError_t e = hipErrorInvalidValue;
    switch (threads) {
    case 64: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_stage<SYS, RS, ST, 64, G, P>, 64, 0); break;
    case 128: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_stage<SYS, RS, ST, 128, G, P>, 128, 0); break;
    case 256: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_stage<SYS, RS, ST, 256, G, P>, 256, 0); break;
    case 512: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_stage<SYS, RS, ST, 512, G, P>, 512, 0); break;
    default: break;
    }
    return e == hipSuccess ? n : 0;
}

// The warp-specialised stage by polar columns per work item (NT flux
// threads plus NT/2 c2p threads per workgroup).
template <int SYS, int RS, bool ST, bool G>
static hipError_t launch_stage_ws(const StageArgs& a, int grid, int threads, hipStream_t s)
{
    switch (threads) {
    case 128: hipLaunchKernelGGL((k_stage_ws<SYS, RS, ST, 128, G>), dim3(grid), dim3(192), 0, s, a); break;
    case 256: hipLaunchKernelGGL((k_stage_ws<SYS, RS, ST, 256, G>), dim3(grid), dim3(384), 0, s, a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <int SYS, int RS, bool ST, bool G>
static int occupancy_ws(int threads)
{
    int n = 0;
    hipError_t e = hipErrorInvalidValue;
    switch (threads) {
    case 128: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_stage_ws<SYS, RS, ST, 128, G>, 192, 0); break;
    case 256: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_stage_ws<SYS, RS, ST, 256, G>, 384, 0); break;
    default: break;
    }
    return e == hipSuccess ? n : 0;
}

template <int SYS, int RS>
static hipError_t launch_stage_rows(const StageArgs& a, int grid, int threads, hipStream_t s)
{
    switch (threads) {
    case 64: hipLaunchKernelGGL((k_stage_rows<SYS, RS, 64>), dim3(grid), dim3(64), 0, s, a); break;
    case 128: hipLaunchKernelGGL((k_stage_rows<SYS, RS, 128>), dim3(grid), dim3(128), 0, s, a); break;
    case 256: hipLaunchKernelGGL((k_stage_rows<SYS, RS, 256>), dim3(grid), dim3(256), 0, s, a); break;
    case 512: hipLaunchKernelGGL((k_stage_rows<SYS, RS, 512>), dim3(grid), dim3(512), 0, s, a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <int SYS, int RS>
static int occupancy_rows(int threads)
{
    int n = 0;
    hipError_t e = hipErrorInvalidValue;
    switch (threads) {
    case 64: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_stage_rows<SYS, RS, 64>, 64, 0); break;
    case 128: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_stage_rows<SYS, RS, 128>, 128, 0); break;
    case 256: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_stage_rows<SYS, RS, 256>, 256, 0); break;
    case 512: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_stage_rows<SYS, RS, 512>, 512, 0); break;
    default: break;
    }
    return e == hipSuccess ? n : 0;
}

// Stage forms: 1 fused, 2 the flux pass of the two-pass stage, 3 the
// warp-specialised fused stage, 4 the row-unit flux pass of the two-pass
// stage.
template <int SYS, int RS, bool ST, bool G>
static hipError_t launch_stage_v(const StageArgs& a, int grid, int threads, int form, hipStream_t s)
{
    if (form == STAGE_ROWS) return launch_stage_rows<SYS, RS>(a, grid, threads, s);
    if (form == STAGE_TWO_PASS) return launch_stage_p<SYS, RS, false, false, true>(a, grid, threads, s);
    if (form == STAGE_SPECIALISED) return launch_stage_ws<SYS, RS, ST, G>(a, grid, threads, s);
    return launch_stage_p<SYS, RS, ST, G, false>(a, grid, threads, s);
}

template <int SYS, int RS, bool ST, bool G>
static int occupancy_v(int threads, int form)
{
    if (form == STAGE_ROWS) return occupancy_rows<SYS, RS>(threads);
    if (form == STAGE_TWO_PASS) return occupancy_p<SYS, RS, false, false, true>(threads);
    if (form == STAGE_SPECIALISED) return occupancy_ws<SYS, RS, ST, G>(threads);
    return occupancy_p<SYS, RS, ST, G, false>(threads);
}

// c2p pass variants by workgroup size.
template <int SYS, int RS, bool ST, bool G>
static hipError_t launch_prim_v(const StageArgs& a, double* partial, int grid, int threads, hipStream_t s)
{
    const dim3 g(grid);
    if (partial) {
        switch (threads) {
        case 64: hipLaunchKernelGGL((k_prim<SYS, ST, G, 64, true>), g, dim3(64), 0, s, a, partial); break;
        case 128: hipLaunchKernelGGL((k_prim<SYS, ST, G, 128, true>), g, dim3(128), 0, s, a, partial); break;
        case 256: hipLaunchKernelGGL((k_prim<SYS, ST, G, 256, true>), g, dim3(256), 0, s, a, partial); break;
        case 512: hipLaunchKernelGGL((k_prim<SYS, ST, G, 512, true>), g, dim3(512), 0, s, a, partial); break;
        default: return hipErrorInvalidValue;
        }
        return hipGetLastError();
    }
    switch (threads) {
    case 64: hipLaunchKernelGGL((k_prim<SYS, ST, G, 64, false>), g, dim3(64), 0, s, a, nullptr); break;
    case 128: hipLaunchKernelGGL((k_prim<SYS, ST, G, 128, false>), g, dim3(128), 0, s, a, nullptr); break;
    case 256: hipLaunchKernelGGL((k_prim<SYS, ST, G, 256, false>), g, dim3(256), 0, s, a, nullptr); break;
    case 512: hipLaunchKernelGGL((k_prim<SYS, ST, G, 512, false>), g, dim3(512), 0, s, a, nullptr); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

bool stage_variant_supported(int threads, int form)
{
    if (form == STAGE_SPECIALISED) return threads == 128 || threads == 256;
    return threads == 64 || threads == 128 || threads == 256 || threads == 512;
}

int stage_max_rows(int form) { return form == STAGE_SPECIALISED ? WS_MAX_ROWS : MAX_ROWS; }

int stage_block_threads(int threads, int form) { return form == STAGE_SPECIALISED ? threads + threads / 2 : threads; }

// Dispatch on (system, solver, stats, learned); the learned guess exists for
// SRHD only (the Euler c2p is closed form).
#define KNC_DISPATCH(FN, ...)                                                                              \
    (system == SYS_SRHD                                                                                    \
         ? (learned ? (rs == RS_HLLC ? (stats ? FN<SYS_SRHD, RS_HLLC, true, true>(__VA_ARGS__)             \
                                              : FN<SYS_SRHD, RS_HLLC, false, true>(__VA_ARGS__))           \
                                     : (stats ? FN<SYS_SRHD, RS_HLLE, true, true>(__VA_ARGS__)             \
                                              : FN<SYS_SRHD, RS_HLLE, false, true>(__VA_ARGS__)))          \
                    : (rs == RS_HLLC ? (stats ? FN<SYS_SRHD, RS_HLLC, true, false>(__VA_ARGS__)            \
                                              : FN<SYS_SRHD, RS_HLLC, false, false>(__VA_ARGS__))          \
                                     : (stats ? FN<SYS_SRHD, RS_HLLE, true, false>(__VA_ARGS__)            \
                                              : FN<SYS_SRHD, RS_HLLE, false, false>(__VA_ARGS__))))        \
         : (rs == RS_HLLC ? (stats ? FN<SYS_EULER, RS_HLLC, true, false>(__VA_ARGS__)                      \
                                   : FN<SYS_EULER, RS_HLLC, false, false>(__VA_ARGS__))                    \
                          : (stats ? FN<SYS_EULER, RS_HLLE, true, false>(__VA_ARGS__)                      \
                                   : FN<SYS_EULER, RS_HLLE, false, false>(__VA_ARGS__))))

int stage_occupancy(int system, int rs, bool stats, bool learned, int form, int threads)
{
    return KNC_DISPATCH(occupancy_v, threads, form);
}

hipError_t launch_stage(int system, int rs, bool stats, bool learned, int form, const StageArgs& a, int grid,
                        int threads, hipStream_t s)
{
    return KNC_DISPATCH(launch_stage_v, a, grid, threads, form, s);
}

hipError_t launch_prim(int system, bool stats, bool learned, const StageArgs& a, int grid, int threads,
                       hipStream_t s, double* dt_partial)
{
    if (dt_partial && (grid < 1 || grid > DT_MAX_GRID)) return hipErrorInvalidValue;
    const int rs = RS_HLLE;
    return KNC_DISPATCH(launch_prim_v, a, dt_partial, grid, threads, s);
}

hipError_t launch_dt(int system, const StageArgs& a, double* partial, int grid, hipStream_t s)
{
    if (grid < 1 || grid > DT_MAX_GRID) return hipErrorInvalidValue;
    if (system == SYS_SRHD) hipLaunchKernelGGL((k_dt<SYS_SRHD>), dim3(grid), dim3(256), 0, s, a, partial);
    else hipLaunchKernelGGL((k_dt<SYS_EULER>), dim3(grid), dim3(256), 0, s, a, partial);
    return hipGetLastError();
}

hipError_t launch_guess(const StageArgs& a, int grid, int threads, hipStream_t s)
{
    const dim3 g(grid);
    switch (threads) {
    case 64: hipLaunchKernelGGL(k_guess<64>, g, dim3(64), 0, s, a); break;
    case 128: hipLaunchKernelGGL(k_guess<128>, g, dim3(128), 0, s, a); break;
    case 256: hipLaunchKernelGGL(k_guess<256>, g, dim3(256), 0, s, a); break;
    case 512: hipLaunchKernelGGL(k_guess<512>, g, dim3(512), 0, s, a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_fma_probe(int chains, int iters, int grid, double* out, hipStream_t s)
{
    const double b = 0.999999, c = 1e-6;
    switch (chains) {
    case 1: hipLaunchKernelGGL(k_fma_probe<1>, dim3(grid), dim3(256), 0, s, iters, b, c, out); break;
    case 2: hipLaunchKernelGGL(k_fma_probe<2>, dim3(grid), dim3(256), 0, s, iters, b, c, out); break;
    case 4: hipLaunchKernelGGL(k_fma_probe<4>, dim3(grid), dim3(256), 0, s, iters, b, c, out); break;
    case 8: hipLaunchKernelGGL(k_fma_probe<8>, dim3(grid), dim3(256), 0, s, iters, b, c, out); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_dt_finish(const double* partial, int n, double cfl, double* dt_out, hipStream_t s)
{
    hipLaunchKernelGGL(k_dt_finish, dim3(1), dim3(256), 0, s, partial, n, cfl, dt_out);
    return hipGetLastError();
}

hipError_t launch_try_to_primitive(int system, const ZoneParams& zp, int64_t n, const double* cons,
                                   const double* guess, double* prim, int32_t* code, int32_t* iters, hipStream_t s)
{
    const int g = grid_for(n, 256);
    if (system == SYS_SRHD)
        hipLaunchKernelGGL((k_try_to_primitive<SYS_SRHD>), dim3(g), dim3(256), 0, s, zp, n, cons, guess, prim, code, iters);
    else
        hipLaunchKernelGGL((k_try_to_primitive<SYS_EULER>), dim3(g), dim3(256), 0, s, zp, n, cons, guess, prim, code, iters);
    return hipGetLastError();
}

hipError_t launch_to_conserved(int system, const ZoneParams& zp, int64_t n, const double* prim, double* cons,
                               hipStream_t s)
{
    const int g = grid_for(n, 256);
    if (system == SYS_SRHD) hipLaunchKernelGGL((k_to_conserved<SYS_SRHD>), dim3(g), dim3(256), 0, s, zp, n, prim, cons);
    else hipLaunchKernelGGL((k_to_conserved<SYS_EULER>), dim3(g), dim3(256), 0, s, zp, n, prim, cons);
    return hipGetLastError();
}

hipError_t launch_intercell_flux(int system, int rs, int axis, const ZoneParams& zp, int64_t n, const double* pl,
                                 const double* pr, const double* sl, const double* sr, double* flux, double* sflux,
                                 hipStream_t s)
{
    const int g = grid_for(n, 256);
#define KNC_FLUX(SYS, RS, AX) \
    hipLaunchKernelGGL((k_intercell_flux<SYS, RS, AX>), dim3(g), dim3(256), 0, s, zp, n, pl, pr, sl, sr, flux, sflux)
    if (system == SYS_SRHD) {
        if (rs == RS_HLLC) { if (axis == 0) KNC_FLUX(SYS_SRHD, RS_HLLC, 0); else KNC_FLUX(SYS_SRHD, RS_HLLC, 1); }
        else               { if (axis == 0) KNC_FLUX(SYS_SRHD, RS_HLLE, 0); else KNC_FLUX(SYS_SRHD, RS_HLLE, 1); }
    } else {
        if (rs == RS_HLLC) { if (axis == 0) KNC_FLUX(SYS_EULER, RS_HLLC, 0); else KNC_FLUX(SYS_EULER, RS_HLLC, 1); }
        else               { if (axis == 0) KNC_FLUX(SYS_EULER, RS_HLLE, 0); else KNC_FLUX(SYS_EULER, RS_HLLE, 1); }
    }
#undef KNC_FLUX
    return hipGetLastError();
}

hipError_t launch_sources(int system, const ZoneParams& zp, int64_t n, const double* prim, const double* r,
                          const double* cot, double* out, hipStream_t s)
{
    const int g = grid_for(n, 256);
    if (system == SYS_SRHD) hipLaunchKernelGGL((k_sources<SYS_SRHD>), dim3(g), dim3(256), 0, s, zp, n, prim, r, cot, out);
    else hipLaunchKernelGGL((k_sources<SYS_EULER>), dim3(g), dim3(256), 0, s, zp, n, prim, r, cot, out);
    return hipGetLastError();
}

hipError_t launch_max_signal_speed(int system, const ZoneParams& zp, int64_t n, const double* prim, double* out,
                                   hipStream_t s)
{
    const int g = grid_for(n, 256);
    if (system == SYS_SRHD) hipLaunchKernelGGL((k_max_signal_speed<SYS_SRHD>), dim3(g), dim3(256), 0, s, zp, n, prim, out);
    else hipLaunchKernelGGL((k_max_signal_speed<SYS_EULER>), dim3(g), dim3(256), 0, s, zp, n, prim, out);
    return hipGetLastError();
}

hipError_t launch_plm(double th, int64_t n, const double* a, const double* b, const double* c, double* out,
                      hipStream_t s)
{
    hipLaunchKernelGGL(k_plm, dim3(grid_for(n, 256)), dim3(256), 0, s, th, n, a, b, c, out);
    return hipGetLastError();
}

hipError_t launch_fp64_probe(int64_t n, const double* a, const double* b, const double* c, double* div, double* sq,
                             double* mad, double* quot, hipStream_t s)
{
    hipLaunchKernelGGL(k_fp64_probe, dim3(grid_for(n, 256)), dim3(256), 0, s, n, a, b, c, div, sq, mad, quot);
    return hipGetLastError();
}

}  // namespace knc
