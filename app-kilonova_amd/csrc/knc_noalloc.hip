// knc_noalloc.hip -- knc_block_time_step_noalloc: the CFL time step of one
// block whose conserved planes live in the caller's device memory, for
// callers that cannot allocate device memory.
//
// The kernels are the context's (k_dt, k_dt_finish).  Everything they need
// beyond the caller's planes -- the block descriptor and work item, the
// radial and polar geometry tables, the per-workgroup minima, dt and the
// error log -- lives in one file-scope device array, `scratch`, of a fixed
// size.  Its device address is taken once per device with
// hipGetSymbolAddress; host code only ever uses that address, never the
// array's name.  The host-side tables are staged in static buffers, so a
// call allocates nothing on the device or the host.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>

#include "../../include/knc_gpu.h"
#include "knc_context.hpp"
#include "knc_device.hpp"
#include "knc_kernels.hpp"
#include "knc_zone.hpp"

using namespace knc;

namespace {

static_assert(KNC_NOALLOC_MAX_NQ <= MAX_NQ, "the no-allocation limit exceeds the stage kernels' polar limit");

constexpr size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// Layout of the scratch array (byte offsets, each part 256-byte aligned).
constexpr size_t OFF_ERR = 0;
constexpr size_t OFF_PART = align256(OFF_ERR + sizeof(ErrLog));
constexpr size_t OFF_DT = align256(OFF_PART + DT_MAX_GRID * sizeof(double));
constexpr size_t OFF_DESC = align256(OFF_DT + sizeof(double));
constexpr size_t OFF_ITEM = align256(OFF_DESC + sizeof(BlockDesc));
constexpr size_t OFF_THETA = align256(OFF_ITEM + sizeof(WorkItem));
constexpr size_t OFF_RGEO = align256(OFF_THETA + theta_table_size(KNC_NOALLOC_MAX_NQ) * sizeof(double));
constexpr size_t SCRATCH_BYTES = align256(OFF_RGEO + rgeo_size(KNC_NOALLOC_MAX_NR) * sizeof(double));
static_assert(SCRATCH_BYTES <= KNC_NOALLOC_SCRATCH_BYTES, "scratch layout exceeds its stated size");

// Held by every device the library runs on until the process ends.
__device__ __align__(256) unsigned char scratch[KNC_NOALLOC_SCRATCH_BYTES];

constexpr int MAX_DEVICES = 64;

std::mutex g_lock;                         // one call at a time in a process
unsigned char* g_scratch[MAX_DEVICES] = {};   // device address of `scratch` per device

// Host staging of the tables and the error log (static: nothing allocated).
double h_theta[theta_table_size(KNC_NOALLOC_MAX_NQ)];
double h_rgeo[rgeo_size(KNC_NOALLOC_MAX_NR)];
ErrLog h_log;

// Clears the error log with a kernel (no memset node for a capturing caller).
__global__ void k_noalloc_reset(ErrLog* e)
{
    if (threadIdx.x == 0) {
        e->min_key = ~0ull;
        e->count = 0u;
    }
}

// The device address of `scratch` on the current device, taken on the
// first call on that device.
int scratch_of_current(int* dev_out, unsigned char** out)
{
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= MAX_DEVICES) return fail(KNC_ERR_UNSUPPORTED, "device %d beyond the %d supported", dev, MAX_DEVICES);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return fail(KNC_ERR_HIP, "hipGetSymbolAddress returned a null address");
        g_scratch[dev] = (unsigned char*)p;
    }
    *dev_out = dev;
    *out = g_scratch[dev];
    return KNC_OK;
}

}  // namespace

extern "C" int knc_block_time_step_noalloc(const knc_hydro* hydro, int32_t index, int32_t nr, int32_t nq,
                                           const double* radial_vertices, const double* polar_vertices,
                                           const double* conserved, void* stream, double* dt, knc_error* err)
{
    if (err) std::memset(err, 0, sizeof *err);
    int rc;
    if ((rc = check_hydro(hydro))) return rc;
    if (!radial_vertices || !polar_vertices || !conserved || !dt) return fail(KNC_ERR_INVALID_ARGUMENT, "NULL argument");
    if (nr < 2 || nr > KNC_NOALLOC_MAX_NR)
        return fail(KNC_ERR_INVALID_ARGUMENT, "nr = %d outside [2, %d], the limit of the no-allocation scratch array",
                    nr, KNC_NOALLOC_MAX_NR);
    if (nq < 1 || nq > KNC_NOALLOC_MAX_NQ)
        return fail(KNC_ERR_INVALID_ARGUMENT, "nq = %d outside [1, %d], the limit of the no-allocation scratch array",
                    nq, KNC_NOALLOC_MAX_NQ);
    if ((rc = check_cell_volumes(nr, nq, radial_vertices, polar_vertices))) return rc;
    // State::time_step with a global signal speed (src/state.rs:146-149)
    if (hydro->system == KNC_SRHD && !hydro->adaptive_time_step) {
        *dt = hydro->cfl_number * (radial_vertices[1] - radial_vertices[0]) / hydro->light_speed;
        return KNC_OK;
    }

    std::lock_guard<std::mutex> lock(g_lock);
    int dev = 0;
    unsigned char* base = nullptr;
    if ((rc = scratch_of_current(&dev, &base))) return rc;
    ErrLog* d_err = (ErrLog*)(base + OFF_ERR);
    double* d_part = (double*)(base + OFF_PART);
    double* d_dt = (double*)(base + OFF_DT);
    BlockDesc* d_desc = (BlockDesc*)(base + OFF_DESC);
    WorkItem* d_item = (WorkItem*)(base + OFF_ITEM);
    double* d_theta = (double*)(base + OFF_THETA);
    double* d_rgeo = (double*)(base + OFF_RGEO);

    fill_theta_table(nq, polar_vertices, h_theta);
    fill_radial_table(nr, radial_vertices, h_rgeo);
    BlockDesc desc{};
    for (int s = 0; s < NSLOT; ++s) desc.slot[s] = const_cast<double*>(conserved);
    desc.rgeo = d_rgeo;
    desc.nr = nr;
    desc.nq = nq;
    desc.left = -1;
    desc.right = -1;
    desc.order = 0;
    desc.index = index;
    WorkItem item{};
    item.block = 0;
    item.i0 = 0;
    item.i1 = nr;
    item.j0 = 0;
    item.j1 = nq;

    hipStream_t s = (hipStream_t)stream;
    int ncu = 0;
    HIP_TRY(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    HIP_TRY(hipMemcpyAsync(d_theta, h_theta, (size_t)theta_table_size(nq) * sizeof(double), hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(d_rgeo, h_rgeo, (size_t)rgeo_size(nr) * sizeof(double), hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(d_desc, &desc, sizeof desc, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(d_item, &item, sizeof item, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_noalloc_reset, dim3(1), dim3(64), 0, s, d_err);
    HIP_TRY(hipGetLastError());

    // One work item over the whole block; k_dt takes one of its rows per
    // workgroup iteration.  dt is a minimum, so it does not depend on the
    // decomposition, and neither does the reported (lowest) failure.
    StageArgs a{};
    a.blocks = d_desc;
    a.items = d_item;
    a.nitems_ptr = nullptr;
    a.nitems = 1;
    a.in_slot = 0;
    a.out_slot = 0;
    a.avg_slot = -1;
    a.zp = zone_params(*hydro);
    a.th = theta_view(d_theta, nq);
    a.err = d_err;
    a.stats = nullptr;
    a.step = 0;
    a.rows = nr;
    const int grid = std::max(1, std::min(nr, std::min(DT_MAX_GRID, 8 * std::max(1, ncu))));
    HIP_TRY(launch_dt(hydro->system, a, d_part, grid, s));
    HIP_TRY(launch_dt_finish(d_part, grid, hydro->cfl_number, d_dt, s));

    double v = 0.0;
    unsigned long long key = 0;
    HIP_TRY(hipMemcpyAsync(&v, d_dt, sizeof v, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(&key, &d_err->min_key, sizeof key, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    if (key == ~0ull) {
        *dt = v;
        return KNC_OK;
    }
    // The lowest failing zone, reported as knc_time_step reports it.
    HIP_TRY(hipMemcpyAsync(&h_log, d_err, sizeof h_log, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    knc_error e{};
    e.code = KNC_HYDRO_ROOT_FINDER_FAILED;
    e.block = index;
    e.i = err_key_row(key);
    e.j = err_key_col(key);
    if (e.i < nr && e.j < nq) {
        e.r = std::sqrt(radial_vertices[e.i] * radial_vertices[e.i + 1]);
        e.theta = 0.5 * (polar_vertices[e.j] + polar_vertices[e.j + 1]);
    }
    const unsigned int n = std::min(h_log.count, (unsigned int)ERR_LOG_CAP);
    for (unsigned int k = 0; k < n; ++k)
        if (h_log.entry[k].key == key) {
            e.code = h_log.entry[k].code;
            e.value = h_log.entry[k].value;
            for (int q = 0; q < 4; ++q) e.conserved[q] = h_log.entry[k].cons[q];
            break;
        }
    if (err) *err = e;
    return fail(e.code, "hydro error %d at block %d zone (%d,%d) (r, theta) = (%.4e, %.4f)", e.code, e.block, e.i, e.j,
                e.r, e.theta);
}
