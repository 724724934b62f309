// knc_models.hip -- device kernels of the initial models: point evaluation,
// BlockState::from_model for a whole block (src/state.rs:55-69) and the
// model ghost rows of a stage (src/scheme.rs:30-52).
#include <hip/hip_runtime.h>

#include "../../include/knc_gpu.h"
#include "knc_device.hpp"
#include "knc_kernels.hpp"
#include "knc_model.hpp"
#include "knc_zone.hpp"

namespace knc {

namespace {

// First failure of a model kernel: fail[0] = code, fail[1] = zone index.
__device__ inline void note_failure(int* fail, int code, int zone)
{
    if (atomicCAS(&fail[0], 0, code) == 0) fail[1] = zone;
}

__global__ void k_model_eval(ModelDev m, int64_t n, const double* r, const double* q, double t, double* prim,
                             double* scal, int* fail)
{
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
        Prim P{0.0, 0.0, 0.0, 0.0};
        double X = 0.0;
        if (!model_at(m, r[k], q[k], t, &P, &X)) note_failure(fail, KNC_ERR_MODEL, (int)k);
        prim[4 * k] = P.rho;
        prim[4 * k + 1] = P.ur;
        prim[4 * k + 2] = P.uq;
        prim[4 * k + 3] = P.p;
        scal[k] = X;
    }
}

// from_model: conserved = to_conserved(P) * dV, scalar_mass =
// lab_frame_mass(conserved) * X, written as the five planes of a block slot.
template <int SYS>
__global__ void k_block_from_model(ModelDev m, ZoneParams zp, int nr, int nq, const double* rc, const double* volA,
                                   const double* qc, const double* dcost, double t, double* planes, int* fail)
{
    const int64_t n = (int64_t)nr * nq;
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(k / nq), j = (int)(k % nq);
        Prim P{0.0, 0.0, 0.0, 0.0};
        double X = 0.0;
        if (!model_at(m, rc[i], qc[j], t, &P, &X)) note_failure(fail, KNC_ERR_MODEL, (int)k);
        const Cons u = p2c<SYS>(P, zp);
        const double dv = volA[i] * dcost[j];
        const double d = u.d * dv;
        planes[k] = d;
        planes[n + k] = u.sr * dv;
        planes[2 * n + k] = u.sq * dv;
        planes[3 * n + k] = u.tau * dv;
        planes[4 * n + k] = d * X;
    }
}

// Ghost rows of one side at one stage time: from_model on the two rows next
// to the mesh edge, then try_to_primitive of U/dV and s = M/dV/Gamma, each
// quotient divided as the reference divides, written as the boundary
// row set (5 planes x 2nq, radial order).  A c2p failure is logged like a
// stage failure, under the ghost block's tag.  model_row_zone is one zone
// of the row set k = i nq + j from the model's primitive P and scalar X
// (ok false: the model cannot be evaluated there).
template <int SYS>
__device__ inline void model_row_zone(bool ok, Prim P, double X, const ZoneParams& zp, int nq, const GhostRows& g,
                                      const double* dcost, int seq, int tag0, double* out, ErrLog* err, int k)
{
    const int n2 = 2 * nq;
    const int i = k / nq, j = k % nq;
    const Cons c = p2c<SYS>(P, zp);
    const double dv = g.volA[i] * dcost[j];
    const double ud = c.d * dv, usr = c.sr * dv, usq = c.sq * dv, utau = c.tau * dv;
    const double um = ud * X;
    const Cons u{ud / dv, usr / dv, usq / dv, utau / dv};
    Prim Q{0.0, 0.0, 0.0, 0.0};
    double rW = 1.0, value = u.d;
    int code = ok ? E_OK : KNC_ERR_MODEL;
    if (code == E_OK) {
        if (SYS == SYS_EULER) {
            code = c2p_euler(u, zp, &Q);
        } else if (u.d < 0.0) {
            code = E_NEGATIVE_DENSITY;
        } else if (u.tau < 0.0) {
            code = E_NEGATIVE_ENERGY_DENSITY;
            value = u.tau;
        } else {
            const double S2 = u.sr * u.sr + u.sq * u.sq;
            double p = 0.0;
            int it = 0;
            code = c2p_srhd_newton(u.d, S2, u.tau, zp, 0.0, &p, &it);
            if (code == E_OK) {
                code = c2p_srhd_finish(u, S2, p, &Q, &rW);
                if (code == E_NEGATIVE_PRESSURE) code = E_OK;
            } else {
                value = u.tau;
            }
        }
    }
    double s = 0.0;
    if (code != E_OK) {
        record_error(err, seq, tag0 | i, j, code, value, u);
        Q = Prim{0.0, 0.0, 0.0, 0.0};
    } else {
        s = um / dv;
        if (SYS == SYS_SRHD) s = s / sqrt_rn_ge1(1.0 + Q.ur * Q.ur + Q.uq * Q.uq);
    }
    out[k] = Q.rho;
    out[n2 + k] = Q.ur;
    out[2 * n2 + k] = Q.uq;
    out[3 * n2 + k] = Q.p;
    out[4 * n2 + k] = s;
}

template <int SYS>
__global__ void k_model_rows(ModelDev m, ZoneParams zp, int nq, GhostRows g, const double* qc, const double* dcost,
                             const RowSet* sets, int nsets, int stride, int tag0, double* base, ErrLog* err)
{
    const int64_t n2 = 2 * (int64_t)nq, n = nsets * n2;
    for (int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; x < n; x += (int64_t)gridDim.x * blockDim.x) {
        const RowSet rs = sets[(x / n2) * stride];
        const int k = (int)(x % n2);
        Prim P{0.0, 0.0, 0.0, 0.0};
        double X = 0.0;
        const bool ok = model_at(m, g.rc[k / nq], qc[k % nq], rs.t, &P, &X);
        model_row_zone<SYS>(ok, P, X, zp, nq, g, dcost, rs.seq, tag0, base + rs.offset, err, k);
    }
}

// The ghost rows of halo_kilonova, one workgroup per zone.  The zone's
// pressure is 1e-3 plus some thousand RK4 steps of the vertical gravity,
// each three evaluations of the disk's density and field (fractional
// powers, square roots and divisions of FP64): as one thread's loop
// (k_model_rows) the rows took milliseconds per stage.  The altitudes of the steps are accumulated
// in order by one thread (the reference's zz += dz), the steps evaluated
// by all threads, and their sum taken in order by one thread, so the
// pressure is the loop's bit for bit.  The grid is exactly nsets x 2 nq
// workgroups.
template <int SYS>
__global__ void __launch_bounds__(256) k_model_rows_halo(ModelDev m, ZoneParams zp, int nq, GhostRows g,
                                                        const double* qc, const double* dcost, const RowSet* sets,
                                                        int stride, int tag0, double* base, ErrLog* err)
{
    __shared__ double alt[HALO_MAX_STEPS];
    __shared__ double step[HALO_MAX_STEPS];
    __shared__ int nsteps;
    const int k = blockIdx.x % (2 * nq);
    const RowSet rs = sets[(blockIdx.x / (2 * nq)) * stride];
    const double t = rs.t;
    const double r = g.rc[k / nq];
    Prim P{0.0, 0.0, 0.0, 0.0};
    double X = 0.0;
    HaloDisk h;
    // the same inputs in every thread: c is uniform over the workgroup
    const int c = halo_kilonova_begin(m.p, r, qc[k % nq], t, &P, &X, &h);
    if (c == 1) {
        if (threadIdx.x == 0) {
            double zz = 0.0;
            int n = 0;
            for (; zz < h.z && n < HALO_MAX_STEPS; ++n) {
                alt[n] = zz;
                zz += h.dz;
            }
            nsteps = n;
        }
        __syncthreads();
        const int ns = nsteps;
        for (int n = threadIdx.x; n < ns; n += blockDim.x) step[n] = gal::pressure_difference_rk4(r, alt[n], h.dz);
        __syncthreads();
        if (threadIdx.x == 0) {
            double pr = 1e-3;
            for (int n = 0; n < ns; ++n) pr += step[n];
            P = Prim{h.d, 0.0, 0.0, pr};
            X = 0.0;
        }
    }
    if (threadIdx.x == 0) model_row_zone<SYS>(c >= 0, P, X, zp, nq, g, dcost, rs.seq, tag0, base + rs.offset, err, k);
}

int grid_of(int64_t n, int block)
{
    int64_t g = (n + block - 1) / block;
    if (g > 4096) g = 4096;
    return g < 1 ? 1 : (int)g;
}

}  // namespace

hipError_t launch_model_eval(const ModelDev& m, int64_t n, const double* r, const double* q, double t, double* prim,
                             double* scal, int* fail, hipStream_t s)
{
    hipLaunchKernelGGL(k_model_eval, dim3(grid_of(n, 256)), dim3(256), 0, s, m, n, r, q, t, prim, scal, fail);
    return hipGetLastError();
}

hipError_t launch_block_from_model(int system, const ModelDev& m, const ZoneParams& zp, int nr, int nq,
                                   const double* rc, const double* volA, const double* qc, const double* dcost,
                                   double t, double* planes, int* fail, hipStream_t s)
{
    const dim3 g(grid_of((int64_t)nr * nq, 256)), b(256);
    if (system == SYS_SRHD)
        hipLaunchKernelGGL(k_block_from_model<SYS_SRHD>, g, b, 0, s, m, zp, nr, nq, rc, volA, qc, dcost, t, planes,
                           fail);
    else
        hipLaunchKernelGGL(k_block_from_model<SYS_EULER>, g, b, 0, s, m, zp, nr, nq, rc, volA, qc, dcost, t, planes,
                           fail);
    return hipGetLastError();
}

hipError_t launch_model_rows(int system, const ModelDev& m, const ZoneParams& zp, int nq, const GhostRows& rows,
                             const double* qc, const double* dcost, const RowSet* sets, int nsets, int stride,
                             int tag0, double* base, ErrLog* err, hipStream_t s, bool one_thread_per_zone)
{
    if (nq < 1 || nsets < 1) return hipSuccess;
    if (m.kind == MODEL_HALO_KILONOVA && !one_thread_per_zone) {
        const dim3 g(nsets * 2 * nq), b(256);
        if (system == SYS_SRHD)
            hipLaunchKernelGGL(k_model_rows_halo<SYS_SRHD>, g, b, 0, s, m, zp, nq, rows, qc, dcost, sets, stride,
                               tag0, base, err);
        else
            hipLaunchKernelGGL(k_model_rows_halo<SYS_EULER>, g, b, 0, s, m, zp, nq, rows, qc, dcost, sets, stride,
                               tag0, base, err);
        return hipGetLastError();
    }
    const dim3 g(grid_of((int64_t)nsets * 2 * nq, 256)), b(256);
    if (system == SYS_SRHD)
        hipLaunchKernelGGL(k_model_rows<SYS_SRHD>, g, b, 0, s, m, zp, nq, rows, qc, dcost, sets, nsets, stride, tag0,
                           base, err);
    else
        hipLaunchKernelGGL(k_model_rows<SYS_EULER>, g, b, 0, s, m, zp, nq, rows, qc, dcost, sets, nsets, stride,
                           tag0, base, err);
    return hipGetLastError();
}

}  // namespace knc
