// knc_kernels.hip -- gfx950 kernels of the KNC per-block RK stage.
//
// k_stage: one RK stage (src/scheme.rs:16-173) fused end to end.  A
// workgroup owns a work item (rows [i0,i1) of one block) and one thread per
// polar column; it marches outward in r.  Each step it recovers the
// primitives of one new row (c2p), forms the radial PLM slope and the radial
// Riemann flux of the face behind it (everything radial stays in the
// thread's registers), and then finishes the zone row two behind: polar PLM
// and polar Riemann fluxes through LDS, geometric sources, flux divergence,
// the update U + dt L(U) and, on the last RK stage, the weighted average
// with U^n (src/state.rs:179-212).  Each zone's c2p runs once per work item
// plus two halo rows on either side; nothing but U is read from or written to
// HBM.
//
// k_dt: the CFL reduction of State::time_step (src/state.rs:143-162): c2p,
// dl / (max_signal_speed * c), wavefront min by cross-lane shuffles, LDS, one
// 64-bit atomicMin per workgroup on the bit pattern of a positive double.
#include <hip/hip_runtime.h>
#include <float.h>

#include "knc_device.hpp"
#include "knc_kernels.hpp"
#include "knc_zone.hpp"

namespace knc {

struct C2PLocal {
    int iters;
};

__device__ inline void record_error(ErrLog* log, int step, int order, int row, int col, int code,
                                    int block_index, double value, const Cons& u)
{
    const unsigned long long key = ((unsigned long long)step << 48) | ((unsigned long long)order << 32) |
                                   ((unsigned long long)row << 16) | (unsigned long long)col;
    atomicMin(&log->min_key, key);
    const unsigned int slot = atomicAdd(&log->count, 1u);
    if (slot < (unsigned int)ERR_LOG_CAP) {
        ErrEntry& e = log->entry[slot];
        e.key = key;
        e.code = code;
        e.block_index = block_index;
        e.value = value;
        e.cons[0] = u.d; e.cons[1] = u.sr; e.cons[2] = u.sq; e.cons[3] = u.tau;
    }
}

// c2p of one zone (BlockState::try_to_primitive, src/state.rs:76-94, and the
// scalar concentration s = M/dV/Gamma of src/scheme.rs:35).  u is the
// volume-integrated state; rdv the reciprocal volume factor.
template <int SYS, bool STATS>
__device__ inline PS zone_primitive(const StageArgs& a, double ud, double usr, double usq, double utau,
                                    double um, double rdv, int order, int row, int col, int block_index,
                                    unsigned int* lds_hist)
{
    const Cons u{ud * rdv, usr * rdv, usq * rdv, utau * rdv};
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
            code = c2p_srhd_newton(u.d, S2, u.tau, a.zp, 0.0, &p, &it);
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
        record_error(a.err, a.step, order, row, col, code, block_index, value, u);
        out.P = Prim{0.0, 0.0, 0.0, 0.0};
        out.s = 0.0;
        return out;
    }
    out.s = (um * rdv) * rW;
    return out;
}

// Primitive + scalar of row k (relative to block B; -2 <= k < nr+2) at
// column j: the block's own rows, the last/first two rows of the inner/outer
// neighbour (src/scheme.rs:73-77), or the boundary rows.  k is uniform
// across the workgroup, so every branch here is uniform.
template <int SYS, bool STATS>
__device__ inline PS fetch_row(const StageArgs& a, const BlockDesc& B, int k, int j, unsigned int* lds_hist)
{
    const int nq = B.nq;
    if (k < 0) {
        if (B.left < 0) {
            if (a.bnd.inner_mode == BND_FIXED) {
                const double* r = a.bnd.inner_rows;
                const int64_t o = (int64_t)(k + 2) * nq + j, n2 = 2 * (int64_t)nq;
                PS q;
                q.P = Prim{r[o], r[o + n2], r[o + 2 * n2], r[o + 3 * n2]};
                q.s = r[o + 4 * n2];
                return q;
            }
            k = 0;
        } else {
            const BlockDesc N = a.blocks[B.left];
            const int row = N.nr + k;
            const int64_t np = (int64_t)N.nr * nq;
            const double* s = slot_of(N, a.in_slot) + (int64_t)row * nq + j;
            const double rdv = rg_rvolA(N)[row] * a.th.rdcost[j];
            return zone_primitive<SYS, STATS>(a, s[0], s[np], s[2 * np], s[3 * np], s[4 * np], rdv, N.order, row, j,
                                              N.index, lds_hist);
        }
    } else if (k >= B.nr) {
        if (B.right < 0) {
            if (a.bnd.outer_mode == BND_FIXED) {
                const double* r = a.bnd.outer_rows;
                const int64_t o = (int64_t)(k - B.nr) * nq + j, n2 = 2 * (int64_t)nq;
                PS q;
                q.P = Prim{r[o], r[o + n2], r[o + 2 * n2], r[o + 3 * n2]};
                q.s = r[o + 4 * n2];
                return q;
            }
            k = B.nr - 1;
        } else {
            const BlockDesc N = a.blocks[B.right];
            const int row = k - B.nr;
            const int64_t np = (int64_t)N.nr * nq;
            const double* s = slot_of(N, a.in_slot) + (int64_t)row * nq + j;
            const double rdv = rg_rvolA(N)[row] * a.th.rdcost[j];
            return zone_primitive<SYS, STATS>(a, s[0], s[np], s[2 * np], s[3 * np], s[4 * np], rdv, N.order, row, j,
                                              N.index, lds_hist);
        }
    }
    const int64_t np = (int64_t)B.nr * nq;
    const double* s = slot_of(B, a.in_slot) + (int64_t)k * nq + j;
    const double rdv = rg_rvolA(B)[k] * a.th.rdcost[j];
    return zone_primitive<SYS, STATS>(a, s[0], s[np], s[2 * np], s[3 * np], s[4 * np], rdv, B.order, k, j, B.index,
                                      lds_hist);
}

__device__ inline void lds_put(double* L, int NT, int j, const PS& q)
{
    L[j] = q.P.rho;
    L[NT + j] = q.P.ur;
    L[2 * NT + j] = q.P.uq;
    L[3 * NT + j] = q.P.p;
    L[4 * NT + j] = q.s;
}

__device__ inline PS lds_get(const double* L, int NT, int j)
{
    PS q;
    q.P = Prim{L[j], L[NT + j], L[2 * NT + j], L[3 * NT + j]};
    q.s = L[4 * NT + j];
    return q;
}

struct Flux5 { double d, sr, sq, tau, s; };

template <int SYS, int RS, bool STATS, int MAXT>
__global__ void __launch_bounds__(MAXT) k_stage(StageArgs a)
{
    extern __shared__ double lds[];            // 5 * blockDim.x doubles
    __shared__ unsigned int lds_hist[HIST_BINS];
    const int NT = blockDim.x;
    const int tid = threadIdx.x;
    if (STATS) {
        for (int b = tid; b < HIST_BINS; b += NT) lds_hist[b] = 0u;
        __syncthreads();
    }
    const int nitems = a.nitems_ptr ? *a.nitems_ptr : a.nitems;
    const double dt = a.dt_dev ? *a.dt_dev : a.dt;
    const double th = a.zp.plm_theta;
    const double c = a.zp.c;

    for (int w = blockIdx.x; w < nitems; w += gridDim.x) {
        const WorkItem it = a.items[w];
        const BlockDesc B = a.blocks[it.block];
        const int nq = B.nq;
        const bool act = tid < nq;
        const int j = act ? tid : nq - 1;
        const double aqj = a.th.aq[j];
        const double cqj = a.th.cq[j];
        const double dcost = a.th.dcost[j];
        const double cotj = a.th.cot[j];
        const double* rr2 = rg_rr2(B);
        const int64_t np = (int64_t)B.nr * nq;

        PS Pm2, Pm1, gm2;
        Flux5 Fm;
        for (int k = it.i0 - 2; k < it.i1 + 2; ++k) {
            const PS Pk = fetch_row<SYS, STATS>(a, B, k, j, lds_hist);
            if (k >= it.i0) {
                const PS g = plm5(th, Pm2, Pm1, Pk);                       // gx[k-1]
                if (k >= it.i0 + 1) {
                    // Radial face k-1, between rows k-2 and k-1 (scheme.rs:81-97).
                    Cons F;
                    double Fs;
                    riemann<SYS, RS, 0>(add_half(Pm2, gm2), sub_half(Pm1, g), a.zp, &F, &Fs);
                    const double A = rr2[k - 1] * aqj;
                    const Flux5 Fk{F.d * c * A, F.sr * c * A, F.sq * c * A, F.tau * c * A, Fs * c * A};
                    if (k >= it.i0 + 2) {
                        // Zone row r = k-2: polar fluxes (scheme.rs:107-128) via LDS.
                        const int r = k - 2;
                        const PS& P = Pm2;
                        __syncthreads();
                        lds_put(lds, NT, tid, P);
                        __syncthreads();
                        PS gy{{0.0, 0.0, 0.0, 0.0}, 0.0};
                        if (j >= 1 && j + 1 < nq) gy = plm5(th, lds_get(lds, NT, j - 1), P, lds_get(lds, NT, j + 1));
                        const PS QR = sub_half(P, gy);
                        const PS QLn = add_half(P, gy);
                        __syncthreads();
                        lds_put(lds, NT, tid, QLn);
                        __syncthreads();
                        Flux5 Fyj{0.0, 0.0, 0.0, 0.0, 0.0};
                        if (act && j >= 1) {
                            Cons G;
                            double Gs;
                            riemann<SYS, RS, 1>(lds_get(lds, NT, j - 1), QR, a.zp, &G, &Gs);
                            const double Aq = rg_rb(B)[r] * cqj;
                            Fyj = Flux5{G.d * c * Aq, G.sr * c * Aq, G.sq * c * Aq, G.tau * c * Aq, Gs * c * Aq};
                        }
                        __syncthreads();
                        lds[tid] = Fyj.d;
                        lds[NT + tid] = Fyj.sr;
                        lds[2 * NT + tid] = Fyj.sq;
                        lds[3 * NT + tid] = Fyj.tau;
                        lds[4 * NT + tid] = Fyj.s;
                        __syncthreads();
                        Flux5 Fyn{0.0, 0.0, 0.0, 0.0, 0.0};
                        if (j + 1 < nq) {
                            Fyn = Flux5{lds[j + 1], lds[NT + j + 1], lds[2 * NT + j + 1], lds[3 * NT + j + 1],
                                        lds[4 * NT + j + 1]};
                        }
                        if (act) {
                            // Sources x dV and the update (scheme.rs:130-158).
                            const double dv = rg_volA(B)[r] * dcost;
                            double ssr, ssq;
                            sources<SYS>(P.P, rg_rinv(B)[r], cotj, a.zp, &ssr, &ssq);
                            const double sc0 = 0.0 * c * dv;
                            const double sc1 = ssr * c * dv;
                            const double sc2 = ssq * c * dv;
                            const double sc3 = 0.0 * c * dv;
                            const double du0 = (sc0 - (Fk.d - Fm.d) - (Fyn.d - Fyj.d)) * dt;
                            const double du1 = (sc1 - (Fk.sr - Fm.sr) - (Fyn.sr - Fyj.sr)) * dt;
                            const double du2 = (sc2 - (Fk.sq - Fm.sq) - (Fyn.sq - Fyj.sq)) * dt;
                            const double du3 = (sc3 - (Fk.tau - Fm.tau) - (Fyn.tau - Fyj.tau)) * dt;
                            const double ds = ((Fk.s - Fm.s) + (Fyn.s - Fyj.s)) * -dt;
                            const int64_t o = (int64_t)r * nq + j;
                            const double* Ui = slot_of(B, a.in_slot) + o;
                            double* Uo = slot_of(B, a.out_slot) + o;
                            double v0 = Ui[0] + du0;
                            double v1 = Ui[np] + du1;
                            double v2 = Ui[2 * np] + du2;
                            double v3 = Ui[3 * np] + du3;
                            double v4 = Ui[4 * np] + ds;
                            if (a.avg_slot >= 0) {
                                const double* U0 = slot_of(B, a.avg_slot) + o;
                                const double b = a.bweight;
                                const double w1 = -b + 1.0;
                                v0 = v0 * w1 + U0[0] * b;
                                v1 = v1 * w1 + U0[np] * b;
                                v2 = v2 * w1 + U0[2 * np] * b;
                                v3 = v3 * w1 + U0[3 * np] * b;
                                v4 = v4 * w1 + U0[4 * np] * b;
                            }
                            Uo[0] = v0;
                            Uo[np] = v1;
                            Uo[2 * np] = v2;
                            Uo[3 * np] = v3;
                            Uo[4 * np] = v4;
                        }
                    }
                    Fm = Fk;
                }
                gm2 = g;
            }
            Pm2 = Pm1;
            Pm1 = Pk;
        }
    }
    if (STATS) {
        __syncthreads();
        for (int b = tid; b < HIST_BINS; b += NT)
            if (lds_hist[b]) atomicAdd(&a.stats->hist[b], (unsigned long long)lds_hist[b]);
    }
}

// ---------------------------------------------------------------------------
// CFL reduction.
// ---------------------------------------------------------------------------
__device__ inline double wave_min(double v)
{
    for (int off = 32; off >= 1; off >>= 1) {
        const double o = __shfl_xor(v, off, 64);
        v = dmin(v, o);
    }
    return v;
}

template <int SYS>
__global__ void __launch_bounds__(256) k_dt(StageArgs a, unsigned long long* min_bits)
{
    __shared__ double wmin[8];
    const int nitems = a.nitems_ptr ? *a.nitems_ptr : a.nitems;
    double m = DBL_MAX;
    for (int w = blockIdx.x; w < nitems; w += gridDim.x) {
        const WorkItem it = a.items[w];
        const BlockDesc B = a.blocks[it.block];
        const int nq = B.nq;
        const int64_t np = (int64_t)B.nr * nq;
        const int nz = (it.i1 - it.i0) * nq;
        for (int z = threadIdx.x; z < nz; z += blockDim.x) {
            const int i = it.i0 + z / nq;
            const int j = z % nq;
            const double* s = slot_of(B, a.in_slot) + (int64_t)i * nq + j;
            const double rdv = rg_rvolA(B)[i] * a.th.rdcost[j];
            const PS q = zone_primitive<SYS, false>(a, s[0], s[np], s[2 * np], s[3 * np], s[4 * np], rdv, B.order, i, j,
                                                    B.index, nullptr);
            const double dr = rg_dr(B)[i];
            const double dl = dmin(dr, a.th.dq[j] * rg_rv(B)[i]);
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
        atomicMin(min_bits, (unsigned long long)__double_as_longlong(r));
    }
}

// dt = min * cfl (state.rs:160); resets the accumulator for the next pass.
__global__ void k_dt_finish(unsigned long long* min_bits, double cfl, double* dt_out)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        const double m = __longlong_as_double((long long)*min_bits);
        *dt_out = m * cfl;
        *min_bits = (unsigned long long)__double_as_longlong(DBL_MAX);
    }
}

__global__ void k_dt_reset(unsigned long long* min_bits)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *min_bits = (unsigned long long)__double_as_longlong(DBL_MAX);
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

// Exact FP64 arithmetic probe: a/b, sqrt(a), a*b+c evaluated as two roundings.
__global__ void k_fp64_probe(int64_t n, const double* a, const double* b, const double* c, double* div,
                             double* sq, double* mad)
{
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
        div[k] = a[k] / b[k];
        sq[k] = sqrt(fabs(a[k]));
        mad[k] = a[k] * b[k] + c[k];
    }
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

hipError_t launch_stage(int system, int rs, bool stats, const StageArgs& a, int grid, int threads, hipStream_t s)
{
    const size_t shmem = (size_t)5 * threads * sizeof(double);
#define KNC_STAGE(SYS, RS, ST)                                                                       \
    do {                                                                                             \
        if (threads <= 256)                                                                          \
            hipLaunchKernelGGL((k_stage<SYS, RS, ST, 256>), dim3(grid), dim3(threads), shmem, s, a); \
        else                                                                                         \
            hipLaunchKernelGGL((k_stage<SYS, RS, ST, 1024>), dim3(grid), dim3(threads), shmem, s, a); \
    } while (0)
    if (system == SYS_SRHD) {
        if (rs == RS_HLLC) { if (stats) KNC_STAGE(SYS_SRHD, RS_HLLC, true); else KNC_STAGE(SYS_SRHD, RS_HLLC, false); }
        else               { if (stats) KNC_STAGE(SYS_SRHD, RS_HLLE, true); else KNC_STAGE(SYS_SRHD, RS_HLLE, false); }
    } else {
        if (rs == RS_HLLC) { if (stats) KNC_STAGE(SYS_EULER, RS_HLLC, true); else KNC_STAGE(SYS_EULER, RS_HLLC, false); }
        else               { if (stats) KNC_STAGE(SYS_EULER, RS_HLLE, true); else KNC_STAGE(SYS_EULER, RS_HLLE, false); }
    }
#undef KNC_STAGE
    return hipGetLastError();
}

hipError_t launch_dt(int system, const StageArgs& a, unsigned long long* min_bits, int grid, hipStream_t s)
{
    if (system == SYS_SRHD) hipLaunchKernelGGL((k_dt<SYS_SRHD>), dim3(grid), dim3(256), 0, s, a, min_bits);
    else hipLaunchKernelGGL((k_dt<SYS_EULER>), dim3(grid), dim3(256), 0, s, a, min_bits);
    return hipGetLastError();
}

hipError_t launch_dt_finish(unsigned long long* min_bits, double cfl, double* dt_out, hipStream_t s)
{
    hipLaunchKernelGGL(k_dt_finish, dim3(1), dim3(64), 0, s, min_bits, cfl, dt_out);
    return hipGetLastError();
}

hipError_t launch_dt_reset(unsigned long long* min_bits, hipStream_t s)
{
    hipLaunchKernelGGL(k_dt_reset, dim3(1), dim3(64), 0, s, min_bits);
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
                             double* mad, hipStream_t s)
{
    hipLaunchKernelGGL(k_fp64_probe, dim3(grid_for(n, 256)), dim3(256), 0, s, n, a, b, c, div, sq, mad);
    return hipGetLastError();
}

}  // namespace knc
