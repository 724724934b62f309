// knc_kernels.hpp -- launch interface of the kernels in knc_kernels.hip.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "knc_device.hpp"
#include "knc_mlp.hpp"
#include "knc_model.hpp"
#include "knc_zone.hpp"

namespace knc {

// Arguments of one stage (or dt) launch.  Passed by value; everything it
// points to is device memory.
struct StageArgs {
    const BlockDesc* blocks;
    const WorkItem* items;
    const int32_t* nitems_ptr;   // device-resident work count, or null
    int32_t nitems;              // used when nitems_ptr is null
    int32_t in_slot;
    int32_t out_slot;
    int32_t avg_slot;            // < 0: no RK average
    double bweight;              // RK weight b: out = (U + dU)(1-b) + U0 b
    const double* dt_dev;        // device-resident dt, or null
    double dt;                   // used when dt_dev is null
    ZoneParams zp;
    ThetaTable th;
    BoundaryDesc bnd;
    ErrLog* err;
    C2PStats* stats;
    int32_t step;                // error-key sequence (0 = dt pass, 2k = k-th RK stage of the run)
    int32_t rows;                // rows per work item (row-parallel passes: k_prim, k_guess, k_dt)
    MlpDev mlp;                  // learned c2p guess (learned-guess kernels only)
    int32_t write_guess;         // guess kernels also store each owned zone's recovered log2(p/D)
    int32_t guess_in;            // guess plane the c2p starts from (k_guess writes it)
    int32_t guess_out;           // guess plane write_guess stores into (never guess_in)
};

// One failed zone into the error log.  seq = 0 for the time-step pass, 2k
// for the k-th RK stage of the run of steps and 2k - 1 for the model ghost
// rows that stage reads; tag = (block order << 16) | row, or
// (GHOST_INNER_ORDER / GHOST_OUTER_ORDER << 16) | row for the model ghost
// rows of that stage.  The key orders failures by (seq, block, row, column)
// and the host maps the block order back to the block index.
__device__ inline void record_error(ErrLog* log, int seq, int tag, int col, int code, double value, const Cons& u)
{
    const unsigned long long key = err_key((unsigned)seq, tag, col);
    atomicMin(&log->min_key, key);
    const unsigned int slot = atomicAdd(&log->count, 1u);
    if (slot < (unsigned int)ERR_LOG_CAP) {
        ErrEntry& e = log->entry[slot];
        e.key = key;
        e.code = code;
        e.value = value;
        e.cons[0] = u.d; e.cons[1] = u.sr; e.cons[2] = u.sq; e.cons[3] = u.tau;
    }
}

constexpr int GHOST_INNER_ORDER = 0x7ffe;
constexpr int GHOST_OUTER_ORDER = 0x7fff;

// Geometry of the two ghost rows of a model boundary: centre radius
// sqrt(r0 r1) and the radial volume factor 2 pi (r1^3 - r0^3)/3.
struct GhostRows {
    double rc[2];
    double volA[2];
};

// One stage in one of its forms (knc_options::stage_passes): fused (c2p
// inside the stage kernel), the flux pass of the two-pass stage (which reads
// the primitive planes k_prim wrote), the warp-specialised fused stage, or
// the two-pass stage with a flux pass of one (work item, row) unit per
// workgroup pass (k_stage_rows; launched with the row-pass grid).
enum { STAGE_FUSED = 1, STAGE_TWO_PASS = 2, STAGE_SPECIALISED = 3, STAGE_ROWS = 4 };
hipError_t launch_stage(int system, int rs, bool stats, bool learned, int form, const StageArgs& a, int grid,
                        int threads, hipStream_t s);
// The c2p pass of the two-pass stage (one (work item, row) unit per
// workgroup iteration).  With dt_partial it also writes each workgroup's CFL
// minimum there, as the dt pass does (grid <= DT_MAX_GRID), for
// launch_dt_finish.
hipError_t launch_prim(int system, bool stats, bool learned, const StageArgs& a, int grid, int threads,
                       hipStream_t s, double* dt_partial = nullptr);
// Polar columns per work item (`threads`) a form supports, its largest
// radial rows per item, and the workgroup size it launches with.
bool stage_variant_supported(int threads, int form);
int stage_occupancy(int system, int rs, bool stats, bool learned, int form, int threads);
int stage_max_rows(int form);
int stage_block_threads(int threads, int form);
// The learned c2p guess of every zone of a.in_slot into the blocks' guess
// planes.  Row-parallel passes (k_guess, k_prim, k_dt) take one (work item,
// row) unit per workgroup iteration, a.rows rows per item, with a 1-D grid
// striding over the units.
hipError_t launch_guess(const StageArgs& a, int grid, int threads, hipStream_t s);
// CFL reduction: each workgroup writes the minimum of dl / (max signal
// speed x c) over its zones to partial[blockIdx.x] (grid <= DT_MAX_GRID);
// k_dt_finish reduces the partials to dt = cfl x min.
constexpr int DT_MAX_GRID = 4096;
hipError_t launch_dt(int system, const StageArgs& a, double* partial, int grid, hipStream_t s);
hipError_t launch_dt_finish(const double* partial, int n, double cfl, double* dt_out, hipStream_t s);

hipError_t launch_try_to_primitive(int system, const ZoneParams& zp, int64_t n, const double* cons,
                                   const double* guess, double* prim, int32_t* code, int32_t* iters, hipStream_t s);
hipError_t launch_to_conserved(int system, const ZoneParams& zp, int64_t n, const double* prim, double* cons,
                               hipStream_t s);
hipError_t launch_intercell_flux(int system, int rs, int axis, const ZoneParams& zp, int64_t n, const double* pl,
                                 const double* pr, const double* sl, const double* sr, double* flux, double* sflux,
                                 hipStream_t s);
hipError_t launch_sources(int system, const ZoneParams& zp, int64_t n, const double* prim, const double* r,
                          const double* cot, double* out, hipStream_t s);
hipError_t launch_max_signal_speed(int system, const ZoneParams& zp, int64_t n, const double* prim, double* out,
                                   hipStream_t s);
hipError_t launch_plm(double th, int64_t n, const double* a, const double* b, const double* c, double* out,
                      hipStream_t s);
// Initial models (knc_models.hip).  fail[0..1] receives the first failure
// (KNC_ERR_MODEL, zone index) of a point or block evaluation.
hipError_t launch_model_eval(const ModelDev& m, int64_t n, const double* r, const double* q, double t, double* prim,
                             double* scal, int* fail, hipStream_t s);
hipError_t launch_block_from_model(int system, const ModelDev& m, const ZoneParams& zp, int nr, int nq,
                                   const double* rc, const double* volA, const double* qc, const double* dcost,
                                   double t, double* planes, int* fail, hipStream_t s);
// One ghost-row set of a side: the stage time, the error-log sequence of
// its stage and where its 5 planes x 2 nq go (in doubles from the base).
struct RowSet {
    double t;
    int32_t seq;
    int32_t pad;
    int64_t offset;
};

// The ghost rows of one side from a device model, for nsets row sets at
// sets[0], sets[stride], ... (device memory), in one launch.  With
// one_thread_per_zone the zone-per-thread kernel runs for every model (the
// library otherwise takes the workgroup-per-zone kernel for halo_kilonova);
// knc_model_rows_check compares the two.
hipError_t launch_model_rows(int system, const ModelDev& m, const ZoneParams& zp, int nq, const GhostRows& rows,
                             const double* qc, const double* dcost, const RowSet* sets, int nsets, int stride,
                             int tag0, double* base, ErrLog* err, hipStream_t s, bool one_thread_per_zone = false);
hipError_t launch_fma_probe(int chains, int iters, int grid, double* out, hipStream_t s);
hipError_t launch_fp64_probe(int64_t n, const double* a, const double* b, const double* c, double* div, double* sq,
                             double* mad, double* quot, hipStream_t s);

}  // namespace knc
