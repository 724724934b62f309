// knc_kernels.hpp -- launch interface of the kernels in knc_kernels.hip.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "knc_device.hpp"
#include "knc_mlp.hpp"
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
    int32_t step;                // error-key step number (0 = dt pass)
    int32_t rows;                // rows per work item (row-parallel passes: k_prim, k_guess, k_dt)
    MlpDev mlp;                  // learned c2p guess (learned-guess kernels only)
};

// One stage: fused (c2p inside the stage kernel) or, with prims, the flux
// pass of the two-pass stage, which reads the primitive planes k_prim wrote.
hipError_t launch_stage(int system, int rs, bool stats, bool learned, bool prims, const StageArgs& a, int grid,
                        int threads, hipStream_t s);
// The c2p pass of the two-pass stage (grid x over work items, y over rows).
hipError_t launch_prim(int system, bool stats, bool learned, const StageArgs& a, int grid, int threads,
                       hipStream_t s);
bool stage_variant_supported(int threads);
int stage_occupancy(int system, int rs, bool stats, bool learned, bool prims, int threads);
int stage_max_rows();
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
hipError_t launch_fp64_probe(int64_t n, const double* a, const double* b, const double* c, double* div, double* sq,
                             double* mad, hipStream_t s);

}  // namespace knc
