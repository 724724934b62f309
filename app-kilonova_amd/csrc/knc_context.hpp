// knc_context.hpp -- the context behind the C ABI (knc_context), shared by
// the translation units of the library's host side: the block set resident
// in HBM, the polar table, boundaries, launch tables, scratch, graphs and
// the models of the device boundaries and births.
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <set>
#include <tuple>
#include <utility>
#include <vector>

#include "../../include/knc_gpu.h"
#include "knc_device.hpp"
#include "knc_kernels.hpp"
#include "knc_mlp.hpp"
#include "knc_model.hpp"
#include "knc_zone.hpp"

namespace knc {

// Sets the message of knc_last_error_message and returns code.
int fail(int code, const char* fmt, ...);

#define HIP_TRY(expr)                                                                              \
    do {                                                                                           \
        hipError_t e_ = (expr);                                                                    \
        if (e_ != hipSuccess)                                                                      \
            return knc::fail(KNC_ERR_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, \
                             __LINE__);                                                            \
    } while (0)

// A device copy of an initial model (knc_model) and the mesh its ghost
// blocks and births come from.
struct DevModel {
    bool on = false;
    ModelDev dev{};
    double* d_table = nullptr;
    knc_mesh mesh{};
};

struct HostBlock {
    int32_t index = 0;
    int nr = 0, nq = 0;
    double* mem = nullptr;      // NSLOT * 5 * nr * nq doubles
    double* rgeo = nullptr;     // device radial table
    double* zgeo = nullptr;     // device per-zone geometry (face areas, 1/dV)
    float* guess = nullptr;     // guess planes, NSLOT * nr * nq floats
    double* prim = nullptr;     // primitive planes of the two-pass stage, 5 * nr * nq
    std::vector<double> rv;
};

}  // namespace knc

struct knc_context {
    knc_hydro h{};
    knc_options o{};
    knc::ZoneParams zp{};
    hipStream_t stream = nullptr;
    int num_cu = 0;

    std::map<int32_t, knc::HostBlock> blocks;  // ordered by radial index
    int cur = 0;                          // slot holding U^n
    int segment_slot = 0;                 // input slot of the current run of steps (KNC_GUESS_LEARNED_FOLD)
    double time = 0.0;
    int64_t iteration = 0;

    // polar table
    int nq = 0;
    std::vector<double> qv;
    double* d_theta = nullptr;
    knc::ThetaTable th{};

    // boundary rows
    int inner_mode = knc::BND_OUTFLOW, outer_mode = knc::BND_OUTFLOW;
    double* d_bnd = nullptr;   // 2 x (5 planes x 2nq)

    // launch tables
    bool dirty = true;
    int rows_per_item = 16;
    knc::BlockDesc* d_blocks = nullptr;
    int cap_blocks = 0;
    knc::WorkItem* d_items = nullptr;
    int cap_items = 0;
    int32_t* d_nitems = nullptr;
    int nitems = 0;
    std::vector<std::pair<int, int>> block_items;  // (offset, count) per table entry
    int threads = 64;
    int persistent_grid = 0;
    int persistent_grid_two_pass = 0;     // the two-pass form's flux kernel (first stage with the dt reduction)
    bool dt_in_first_stage = false;       // the next run's first stage computes dt (its c2p pass reduces the CFL minimum)

    // per-call device state
    knc::ErrLog* d_err = nullptr;
    knc::C2PStats* d_stats = nullptr;
    double* d_dt_partial = nullptr;       // per-workgroup minima of the CFL pass
    double* d_dt = nullptr;

    // graphs keyed by (starting slot, steps, first fold step for model
    // boundaries, whether the run's first stage carries the CFL reduction)
    std::map<std::tuple<int, int, int, int>, hipGraphExec_t> graphs;
    // keys run once, directly, since the graphs were last dropped
    std::set<std::tuple<int, int, int, int>> graph_keys_seen;

    // learned c2p guess: device copy of the network's operand fragments
    void* d_mlp = nullptr;
    knc::MlpDev mlp{};
    bool has_mlp = false;

    // model boundaries: ghost rows per stage of a fold, side-major within a
    // stage (5 planes x 2nq each), filled on the host and uploaded per segment
    knc_boundary_fn bsrc[2] = {nullptr, nullptr};
    void* buser[2] = {nullptr, nullptr};
    double* h_bseq = nullptr;   // pinned
    double* d_bseq = nullptr;
    size_t cap_bseq = 0;        // doubles
    // the device-model sides' row sets of a fold, one per (step, stage,
    // side) in the order of the sets in d_bseq
    knc::RowSet* h_rowsets = nullptr;   // pinned
    knc::RowSet* d_rowsets = nullptr;
    size_t cap_rowsets = 0;
    bool trial = false;         // autotuner trial steps: fixed rows only

    // initial models evaluated on the device: per side for model boundaries,
    // one for excision births; stage times of the fold's ghost-row sets
    knc::DevModel bmodel[2];
    knc::DevModel birth;
    double* d_qc = nullptr;          // polar zone centres 0.5 (q0 + q1)
    std::vector<double> stage_times; // per (fold step, stage)

    // moving excision surfaces
    bool exc_on = false;
    knc_excision exc{};
    int64_t exc_removed = 0, exc_added = 0;

    // timing
    bool timing = false;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double last_stage_ms = 0.0;
    int last_launches = 0;
};


namespace knc {

// Hydrodynamics::validate of a configuration (and its enum fields), and the
// per-zone parameters derived from it.
int check_hydro(const knc_hydro* h);
ZoneParams zone_params(const knc_hydro& h);

// Geometry tables from the vertices (layouts in knc_device.hpp), with the
// formulas of src/mesh.rs:119-134,157-161: the radial table of a block
// (rgeo_size(nr) doubles) and the polar table (theta_table_size(nq) doubles,
// six arrays of nq + 1), and a ThetaTable over a device copy of the latter.
void fill_radial_table(int nr, const double* rv, double* out);
void fill_theta_table(int nq, const double* qv, double* out);
ThetaTable theta_view(const double* d_table, int nq);
// Per-zone geometry of a block (layout in knc_device.hpp, zgeo_size(nr, nq)
// doubles): face areas and 1/dV from the vertices, as src/mesh.rs:119-134
// writes them.
void fill_zone_geometry(int nr, int nq, const double* rv, const double* qv, double* out);
// U/dV divides through div_rcp, whose three-operation path is exact for
// volumes in [DIV_RCP_DMIN, DIV_RCP_DMAX]: fails unless every cell volume of
// the block lies within a factor of two of that range.
int check_cell_volumes(int nr, int nq, const double* rv, const double* qv);

// Inserts a block from host data (conserved may be NULL: the slot is left for
// the caller to fill on the device).
int insert_block(knc_context* c, int32_t index, int32_t nr, int32_t nq, const double* rv, const double* qv,
                 const double* conserved, const double* scalar_mass);
// Inserts block `index` of the model's mesh built on the device by from_model.
int insert_block_model(knc_context* c, int32_t index, const DevModel& dm, double t);
// Clears the context's error log on its stream.
int reset_errors(knc_context* c);
// Reads the error log after a synchronised call; fills err and returns the
// lowest failure's code (KNC_OK when there is none).
int collect_error(knc_context* c, knc_error* err);

}  // namespace knc
