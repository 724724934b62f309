// knc_capi.hip -- host side of the library: the C ABI declared in
// include/knc_gpu.h and the context that keeps a KNC State resident in HBM.
//
// The context mirrors the pieces of the reference's scheme that surround
// the stage kernel: the block set and its radial chain (State,
// src/state.rs:26-43, with neighbours at index +-1 as in src/scheme.rs:70-75),
// the geometry cache (src/mesh.rs:246-264, kept as 1-D factors), the RK
// driver (godunov-core RungeKuttaOrder, weights of src/state.rs:179-212), the
// CFL time step (src/state.rs:143-162) and scheme::advance's fold loop
// (src/scheme.rs:212-240).  Errors follow HydroError (src/physics/mod.rs:
// 82-117) with the lowest (step, block, i, j) reported.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
#include <string>
#include <vector>

#include "../../include/knc_gpu.h"
#include "knc_context.hpp"
#include "knc_device.hpp"
#include "knc_kernels.hpp"
#include "knc_mlp_host.hpp"
#include "knc_zone.hpp"

using namespace knc;

namespace {

thread_local std::string g_last_error;

}  // namespace

int knc::fail(int code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

void knc::fill_radial_table(int nr, const double* rv, double* v)
{
    double* volA = v + (nr + 1);
    double* rinv = volA + nr;
    double* dr = rinv + nr;
    for (int i = 0; i <= nr; ++i) v[i] = rv[i];
    for (int i = 0; i < nr; ++i) {
        const double r0 = rv[i], r1 = rv[i + 1];
        volA[i] = 2.0 * KNC_PI * ((r1 * r1) * r1 - (r0 * r0) * r0) / 3.0;
        rinv[i] = 1.0 / std::sqrt(r0 * r1);
        dr[i] = r1 - r0;
    }
}

void knc::fill_theta_table(int nq, const double* qv, double* t)
{
    const size_t m = (size_t)nq + 1;
    double* dcost = t;
    double* cot = dcost + m;
    double* dq = cot + m;
    for (int j = 0; j <= nq; ++j) {
        const double q0 = qv[j], q1 = j < nq ? qv[j + 1] : qv[j];
        dcost[j] = -(std::cos(q1) - std::cos(q0));
        const double q = 0.5 * (q0 + q1);
        cot[j] = std::cos(q) / std::sin(q);
        dq[j] = q1 - q0;
    }
}

// face_area of src/mesh.rs:125-134 between the corners (r0, q0) and (r1, q1),
// written as there: pi (s0 + s1) sqrt(ds^2 + dz^2), s = r sin q, z = r cos q.
static double face_area(double r0, double sin0, double cos0, double r1, double sin1, double cos1)
{
    const double s0 = r0 * sin0;
    const double s1 = r1 * sin1;
    const double z0 = r0 * cos0;
    const double z1 = r1 * cos1;
    const double ds = s1 - s0;
    const double dz = z1 - z0;
    return KNC_PI * (s0 + s1) * std::sqrt(ds * ds + dz * dz);
}

void knc::fill_zone_geometry(int nr, int nq, const double* rv, const double* qv, double* out)
{
    std::vector<double> sn((size_t)nq + 1), cs((size_t)nq + 1), dcost((size_t)nq);
    for (int j = 0; j <= nq; ++j) {
        sn[j] = std::sin(qv[j]);
        cs[j] = std::cos(qv[j]);
    }
    for (int j = 0; j < nq; ++j) dcost[j] = -(cs[j + 1] - cs[j]);
    double* ar = out;
    double* aq = ar + (size_t)(nr + 1) * nq;
    double* rdv = aq + (size_t)nr * (nq + 1);
    for (int i = 0; i <= nr; ++i)   // zone(i, j).face_area_r(): corners (r_i, q_j), (r_i, q_j+1)
        for (int j = 0; j < nq; ++j)
            ar[(size_t)i * nq + j] = face_area(rv[i], sn[j], cs[j], rv[i], sn[j + 1], cs[j + 1]);
    for (int i = 0; i < nr; ++i)    // zone(i, j).face_area_q(): corners (r_i, q_j), (r_i+1, q_j)
        for (int j = 0; j <= nq; ++j)
            aq[(size_t)i * (nq + 1) + j] = face_area(rv[i], sn[j], cs[j], rv[i + 1], sn[j], cs[j]);
    for (int i = 0; i < nr; ++i) {  // cell_volume (mesh.rs:119-123): 2 pi (r1^3 - r0^3) / 3 * dcost
        const double r0 = rv[i], r1 = rv[i + 1];
        const double volA = 2.0 * KNC_PI * ((r1 * r1) * r1 - (r0 * r0) * r0) / 3.0;
        for (int j = 0; j < nq; ++j) rdv[(size_t)i * nq + j] = 1.0 / (volA * dcost[j]);
    }
}

int knc::check_cell_volumes(int nr, int nq, const double* rv, const double* qv)
{
    double vlo = HUGE_VAL, vhi = 0.0, clo = HUGE_VAL, chi = 0.0;
    for (int i = 0; i < nr; ++i) {
        const double r0 = rv[i], r1 = rv[i + 1];
        const double volA = 2.0 * KNC_PI * ((r1 * r1) * r1 - (r0 * r0) * r0) / 3.0;
        vlo = std::min(vlo, volA);
        vhi = std::max(vhi, volA);
    }
    for (int j = 0; j < nq; ++j) {
        const double dcost = -(std::cos(qv[j + 1]) - std::cos(qv[j]));
        clo = std::min(clo, dcost);
        chi = std::max(chi, dcost);
    }
    if (!(vlo * clo >= 2.0 * DIV_RCP_DMIN && vhi * chi <= 0.5 * DIV_RCP_DMAX))
        return fail(KNC_ERR_INVALID_ARGUMENT, "cell volumes outside [2^-399, 2^399] (or not positive): %.3e .. %.3e",
                    vlo * clo, vhi * chi);
    return KNC_OK;
}

knc::ThetaTable knc::theta_view(const double* d, int nq)
{
    const size_t m = (size_t)nq + 1;
    ThetaTable th;
    th.dcost = d;
    th.cot = d + m;
    th.dq = d + 2 * m;
    th.nq = nq;
    return th;
}

int knc::check_hydro(const knc_hydro* h)
{
    if (!h) return fail(KNC_ERR_INVALID_ARGUMENT, "hydro is NULL");
    if (h->system != KNC_EULER && h->system != KNC_SRHD) return fail(KNC_ERR_INVALID_ARGUMENT, "unknown system");
    if (h->riemann_solver != KNC_HLLE && h->riemann_solver != KNC_HLLC)
        return fail(KNC_ERR_INVALID_ARGUMENT, "unknown riemann solver");
    if (h->runge_kutta_order < 1 || h->runge_kutta_order > 3)
        return fail(KNC_ERR_INVALID_ARGUMENT, "runge_kutta_order must be 1, 2 or 3");
    // Hydrodynamics::validate (relativistic_hydro.rs:46-54, newtonian_hydro.rs:38-46)
    if (h->plm_theta < 1.0 || h->plm_theta > 2.0)
        return fail(KNC_ERR_INVALID_ARGUMENT, "plm_theta must be in the range [1, 2]");
    if (h->cfl_number < 0.0 || h->cfl_number > 0.7)
        return fail(KNC_ERR_INVALID_ARGUMENT, "cfl_number must be in the range [0.0, 0.7]");
    if (!(h->gamma_law_index > 1.0)) return fail(KNC_ERR_INVALID_ARGUMENT, "gamma_law_index must exceed 1");
    if (!(h->light_speed > 0.0)) return fail(KNC_ERR_INVALID_ARGUMENT, "light_speed must be positive");
    return KNC_OK;
}

ZoneParams knc::zone_params(const knc_hydro& h)
{
    ZoneParams zp;
    zp.gamma = h.gamma_law_index;
    zp.gm1 = h.gamma_law_index - 1.0;
    zp.gfac = h.gamma_law_index / (h.gamma_law_index - 1.0);
    zp.plm_theta = h.plm_theta;
    zp.c = h.light_speed;
    return zp;
}

namespace {

// Radial table of one block (layout in knc_device.hpp), from the vertices
// with the formulas of src/mesh.rs:119-134,157-161.
std::vector<double> radial_table(int nr, const double* rv)
{
    std::vector<double> t((size_t)rgeo_size(nr));
    fill_radial_table(nr, rv, t.data());
    return t;
}

}  // namespace

namespace {

// The network supplies guesses (k_guess passes, a network must be attached).
bool learned(const knc_context* c)
{
    return (c->o.c2p_guess == KNC_GUESS_LEARNED || c->o.c2p_guess == KNC_GUESS_LEARNED_STEP ||
            c->o.c2p_guess == KNC_GUESS_LEARNED_FOLD) &&
           c->h.system == KNC_SRHD;
}
// The stage's c2p starts Newton from the guess planes (learned or previous).
bool guess_read(const knc_context* c)
{
    return learned(c) || (c->o.c2p_guess == KNC_GUESS_PREVIOUS && c->h.system == KNC_SRHD);
}
// The stage form (STAGE_FUSED, STAGE_TWO_PASS, STAGE_SPECIALISED, STAGE_ROWS).
int stage_form(const knc_context* c) { return c->o.stage_passes == 0 ? STAGE_FUSED : c->o.stage_passes; }
// Forms whose flux pass reads the primitive planes of a c2p pass.
bool two_pass(const knc_context* c) { return stage_form(c) == STAGE_TWO_PASS || stage_form(c) == STAGE_ROWS; }
// A side whose ghost rows are refreshed at every stage time (host callback
// or device model).
bool model_side(const knc_context* c, int side) { return c->bsrc[side] || c->bmodel[side].on; }
bool any_model_side(const knc_context* c) { return model_side(c, 0) || model_side(c, 1); }

// Mesh::zone_dlogr and the vertices of block `index`
// (Mesh::subgrid_extent, SphericalPolarGrid::vertex_coordinate_signed;
// src/mesh.rs:341-375, 202-213).  f64::log(10.0) is ln(x)/ln(10).
double mesh_zone_dlogr(const knc_mesh& m)
{
    return m.num_radial_zones > 0 ? 1.0 / (double)m.num_radial_zones : KNC_PI / (double)m.num_polar_zones;
}

void mesh_vertices(const knc_mesh& m, int32_t index, double* rv, double* qv)
{
    const double dlogr = mesh_zone_dlogr(m);
    const double b = (double)m.block_size * dlogr;
    const double r0 = m.reference_radius * std::pow(1.0 + b, (double)index);
    const double r1 = m.reference_radius * std::pow(1.0 + b, (double)index + 1.0);
    const double q0 = m.num_polar_zones == 1 ? KNC_PI * 0.5 - dlogr : 0.0;
    const double q1 = m.num_polar_zones == 1 ? KNC_PI * 0.5 + dlogr : KNC_PI;
    const double y0 = std::log(r0) / std::log(10.0), y1 = std::log(r1) / std::log(10.0);
    const double dy = (y1 - y0) / (double)m.block_size;
    const double dq = (q1 - q0) / (double)m.num_polar_zones;
    if (rv)
        for (int i = 0; i <= m.block_size; ++i) rv[i] = std::pow(10.0, y0 + dy * (double)i);
    if (qv)
        for (int j = 0; j <= m.num_polar_zones; ++j) qv[j] = q0 + dq * (double)j;
}

int check_mesh(const knc_mesh* m)
{
    if (!m) return fail(KNC_ERR_INVALID_ARGUMENT, "mesh is NULL");
    if (!(m->reference_radius > 0.0)) return fail(KNC_ERR_INVALID_ARGUMENT, "reference_radius must be positive");
    if (m->block_size < 2) return fail(KNC_ERR_INVALID_ARGUMENT, "must have at least 2 radial zones per block");
    if (m->block_size >= 65536) return fail(KNC_ERR_INVALID_ARGUMENT, "block_size must be below 65536");
    if (m->num_polar_zones < 1 || m->num_polar_zones > MAX_NQ)
        return fail(KNC_ERR_INVALID_ARGUMENT, "num_polar_zones must be in [1, %d]", MAX_NQ);
    if (m->num_radial_zones < 0) return fail(KNC_ERR_INVALID_ARGUMENT, "num_radial_zones must be >= 0");
    if (m->num_polar_zones == 1 && m->num_radial_zones == 0)
        return fail(KNC_ERR_INVALID_ARGUMENT, "num_radial_zones is not optional when num_polar_zones=1");
    return KNC_OK;
}

// The two ghost rows next to the live mesh: rows nr-2, nr-1 of block
// inner-1 (side 0) or rows 0, 1 of block outer+1 (side 1), with the
// radial factors of the block tables.
GhostRows ghost_rows(const knc_mesh& m, int side, int32_t inner, int32_t outer)
{
    std::vector<double> rv((size_t)m.block_size + 1);
    mesh_vertices(m, side ? outer + 1 : inner - 1, rv.data(), nullptr);
    const int i0 = side ? 0 : m.block_size - 2;
    GhostRows g;
    for (int i = 0; i < 2; ++i) {
        const double r0 = rv[i0 + i], r1 = rv[i0 + i + 1];
        g.rc[i] = std::sqrt(r0 * r1);
        g.volA[i] = 2.0 * KNC_PI * ((r1 * r1) * r1 - (r0 * r0) * r0) / 3.0;
    }
    return g;
}

void free_model(DevModel& dm)
{
    if (dm.d_table) (void)hipFree(dm.d_table);
    dm = DevModel{};
}

// Validates and uploads a model; the table rows must strictly increase in
// their first column (LookupTable::from_rows, src/lookup_table_v2.rs:35-48).
int load_model(DevModel& dm, const knc_mesh* mesh, const knc_model* m)
{
    if (!m) return fail(KNC_ERR_INVALID_ARGUMENT, "model is NULL");
    if (m->kind < KNC_MODEL_KINETIC_BOMB || m->kind > KNC_MODEL_WIND_SHOCK)
        return fail(KNC_ERR_INVALID_ARGUMENT, "unknown model kind %d", m->kind);
    int rows = m->kind == KNC_MODEL_WIND_SHOCK ? m->table_rows : 0;
    if (rows < 0) return fail(KNC_ERR_INVALID_ARGUMENT, "table_rows must be >= 0");
    if (rows > 0) {
        if (!m->table) return fail(KNC_ERR_INVALID_ARGUMENT, "table is NULL");
        if (rows < 2) return fail(KNC_ERR_INVALID_ARGUMENT, "the table must have at least two rows");
        for (int k = 1; k < rows; ++k)
            if (!(m->table[4 * k] > m->table[4 * (k - 1)]))
                return fail(KNC_ERR_INVALID_ARGUMENT, "the left-most column of the table must increase monotonically");
    }
    DevModel nd;
    nd.on = true;
    nd.dev.kind = m->kind;
    nd.dev.table_rows = rows;
    for (int k = 0; k < MODEL_NPARAM; ++k) nd.dev.p[k] = m->params[k];
    if (mesh) nd.mesh = *mesh;
    if (rows > 0) {
        HIP_TRY(hipMalloc(&nd.d_table, (size_t)4 * rows * sizeof(double)));
        if (hipMemcpy(nd.d_table, m->table, (size_t)4 * rows * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
            (void)hipFree(nd.d_table);
            return fail(KNC_ERR_HIP, "hipMemcpy failed");
        }
    }
    nd.dev.table = nd.d_table;
    free_model(dm);
    dm = nd;
    return KNC_OK;
}

void drop_graphs(knc_context* c)
{
    for (auto& kv : c->graphs) (void)hipGraphExecDestroy(kv.second);
    c->graphs.clear();
    c->graph_keys_seen.clear();
}

// Polar table (layout in knc_device.hpp), with separate sin and cos calls.
// Stage workgroup size: one thread per polar column for meshes up to 256
// columns; wider meshes are cut into 256-column tiles.
int choose_stage_shape(knc_context* c)
{
    const int nq = c->nq;
    int t = c->o.stage_threads;
    if (t <= 0) t = nq <= 64 && stage_form(c) != STAGE_SPECIALISED ? 64 : nq <= 128 ? 128 : 256;
    if (!stage_variant_supported(t, stage_form(c)))
        return fail(KNC_ERR_INVALID_ARGUMENT, "unsupported stage shape: %d columns in form %d", t, stage_form(c));
    c->threads = t;
    return KNC_OK;
}

int upload_theta(knc_context* c, int nq, const double* qv)
{
    c->nq = nq;
    c->qv.assign(qv, qv + nq + 1);
    std::vector<double> t((size_t)theta_table_size(nq), 0.0);
    fill_theta_table(nq, qv, t.data());
    if (c->d_theta) HIP_TRY(hipFree(c->d_theta));
    HIP_TRY(hipMalloc(&c->d_theta, t.size() * sizeof(double)));
    HIP_TRY(hipMemcpy(c->d_theta, t.data(), t.size() * sizeof(double), hipMemcpyHostToDevice));
    c->th = theta_view(c->d_theta, nq);
    std::vector<double> qc(nq);
    for (int j = 0; j < nq; ++j) qc[j] = 0.5 * (qv[j] + qv[j + 1]);
    if (c->d_qc) HIP_TRY(hipFree(c->d_qc));
    HIP_TRY(hipMalloc(&c->d_qc, nq * sizeof(double)));
    HIP_TRY(hipMemcpy(c->d_qc, qc.data(), nq * sizeof(double), hipMemcpyHostToDevice));
    if (c->d_bnd) HIP_TRY(hipFree(c->d_bnd));
    HIP_TRY(hipMalloc(&c->d_bnd, (size_t)2 * 5 * 2 * nq * sizeof(double)));
    HIP_TRY(hipMemset(c->d_bnd, 0, (size_t)2 * 5 * 2 * nq * sizeof(double)));
    return choose_stage_shape(c);
}

int choose_rows(knc_context* c)
{
    if (c->o.rows_per_item > 0) return std::min(c->o.rows_per_item, stage_max_rows(stage_form(c)));
    // Enough work items for one full round of resident workgroups, while
    // keeping the 4 recomputed halo rows per item a small fraction.
    int64_t rows = 0;
    for (auto& kv : c->blocks) rows += kv.second.nr;
    const int tiles = (c->nq + c->threads - 1) / c->threads;
    const int occ = std::max(1, stage_occupancy(c->h.system, c->h.riemann_solver, c->o.collect_stats != 0, guess_read(c),
                                                stage_form(c), c->threads));
    const int64_t want = (int64_t)c->num_cu * occ;
    int64_t r = rows * tiles / std::max<int64_t>(1, want);
    return (int)std::max<int64_t>(4, std::min<int64_t>(32, r));
}

// Rebuilds the device block table and work list after the block set changed.
int rebuild_tables(knc_context* c)
{
    if (!c->dirty) return KNC_OK;
    // captured launches hold the grid size and work count of the old tables
    drop_graphs(c);
    const int nb = (int)c->blocks.size();
    c->rows_per_item = choose_rows(c);
    std::vector<BlockDesc> desc;
    std::vector<WorkItem> items;
    c->block_items.clear();
    int order = 0;
    int32_t prev_index = 0;
    for (auto& kv : c->blocks) {
        HostBlock& b = kv.second;
        if (order > 0 && b.index != prev_index + 1)
            return fail(KNC_ERR_INVALID_ARGUMENT, "block indices must be contiguous (gap after %d)", prev_index);
        prev_index = b.index;
        BlockDesc d{};
        const size_t plane = (size_t)5 * b.nr * b.nq;
        for (int s = 0; s < NSLOT; ++s) d.slot[s] = b.mem + s * plane;
        d.rgeo = b.rgeo;
        d.zgeo = b.zgeo;
        d.guess = b.guess;
        d.prim = b.prim;
        d.nr = b.nr;
        d.nq = b.nq;
        d.left = order > 0 ? order - 1 : -1;
        d.right = order + 1 < nb ? order + 1 : -1;
        d.order = order;
        d.index = b.index;
        desc.push_back(d);
        const int off = (int)items.size();
        for (int i0 = 0; i0 < b.nr; i0 += c->rows_per_item)
            for (int j0 = 0; j0 < b.nq; j0 += c->threads) {
                WorkItem w{};
                w.block = order;
                w.i0 = i0;
                w.i1 = std::min(b.nr, i0 + c->rows_per_item);
                w.j0 = j0;
                w.j1 = std::min(b.nq, j0 + c->threads);
                items.push_back(w);
            }
        c->block_items.push_back({off, (int)items.size() - off});
        ++order;
    }
    if (nb > c->cap_blocks) {
        drop_graphs(c);
        if (c->d_blocks) HIP_TRY(hipFree(c->d_blocks));
        c->cap_blocks = std::max(nb, 2 * c->cap_blocks);
        HIP_TRY(hipMalloc(&c->d_blocks, (size_t)c->cap_blocks * sizeof(BlockDesc)));
    }
    if ((int)items.size() > c->cap_items) {
        drop_graphs(c);
        if (c->d_items) HIP_TRY(hipFree(c->d_items));
        c->cap_items = std::max((int)items.size(), 2 * c->cap_items);
        HIP_TRY(hipMalloc(&c->d_items, (size_t)c->cap_items * sizeof(WorkItem)));
    }
    if (nb) HIP_TRY(hipMemcpyAsync(c->d_blocks, desc.data(), desc.size() * sizeof(BlockDesc), hipMemcpyHostToDevice,
                                   c->stream));
    if (!items.empty())
        HIP_TRY(hipMemcpyAsync(c->d_items, items.data(), items.size() * sizeof(WorkItem), hipMemcpyHostToDevice,
                               c->stream));
    c->nitems = (int)items.size();
    int32_t n = c->nitems;
    HIP_TRY(hipMemcpyAsync(c->d_nitems, &n, sizeof n, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    // Persistent grid: every CU holds as many stage workgroups as fit.
    const int per_cu = std::max(1, stage_occupancy(c->h.system, c->h.riemann_solver, c->o.collect_stats != 0,
                                                   guess_read(c), stage_form(c), c->threads));
    c->persistent_grid = std::min(c->nitems, c->num_cu * per_cu);
    const int per_cu_tp = std::max(1, stage_occupancy(c->h.system, c->h.riemann_solver, c->o.collect_stats != 0,
                                                      guess_read(c), STAGE_TWO_PASS, c->threads));
    c->persistent_grid_two_pass = std::min(c->nitems, c->num_cu * per_cu_tp);
    c->dirty = false;
    return KNC_OK;
}

StageArgs base_args(knc_context* c)
{
    StageArgs a{};
    a.blocks = c->d_blocks;
    a.items = c->d_items;
    a.nitems_ptr = nullptr;
    a.nitems = c->nitems;
    a.in_slot = c->cur;
    a.out_slot = c->cur;
    a.avg_slot = -1;
    a.bweight = 0.0;
    a.dt_dev = c->d_dt;
    a.dt = 0.0;
    a.zp = c->zp;
    a.th = c->th;
    a.bnd.inner_mode = c->inner_mode;
    a.bnd.outer_mode = c->outer_mode;
    a.bnd.inner_rows = c->d_bnd;
    a.bnd.outer_rows = c->d_bnd + (size_t)5 * 2 * c->nq;
    a.err = c->d_err;
    a.stats = c->d_stats;
    a.step = 0;
    a.rows = c->rows_per_item;
    a.mlp = c->mlp;
    a.write_guess = c->o.c2p_guess == KNC_GUESS_PREVIOUS;
    a.guess_in = c->cur;
    a.guess_out = c->cur;
    return a;
}

__global__ void k_reset_err(ErrLog* e)
{
    if (threadIdx.x == 0) {
        e->min_key = ~0ull;
        e->count = 0u;
    }
}

__global__ void k_set_double(double* p, double v)
{
    if (threadIdx.x == 0) *p = v;
}

// Grid of a row-parallel pass over `nitems` work items: one workgroup per
// (item, row) unit up to eight per CU, which then stride over the units.
int row_pass_grid(const knc_context* c, int nitems)
{
    const int64_t units = (int64_t)nitems * c->rows_per_item;
    return (int)std::max<int64_t>(1, std::min<int64_t>(units, std::min(DT_MAX_GRID, 8 * c->num_cu)));
}

// Launches one stage over every block: per-block launches or one batched
// one.  The passes before the stage kernels (the learned-guess pass, the c2p
// pass of the two-pass stage) cover every block first, since a block's stage
// also reads its neighbours' edge rows.  With with_dt (the first stage of a
// batched run whose dt is not yet computed) the stage runs in the two-pass
// form and its c2p pass is also the CFL reduction of the run (launch_prim
// with the partial minima, then k_dt_finish), in place of a separate dt pass
// over the same state; its failures are logged as the time step's
// (sequence 0), where State::time_step meets them first.
int launch_one_stage(knc_context* c, StageArgs a, bool first_stage, bool segment_start, bool with_dt = false)
{
    const bool stats = c->o.collect_stats != 0;
    const bool prims = two_pass(c) || with_dt;
    const bool lg = guess_read(c);
    // KNC_GUESS_LEARNED_STEP evaluates the network on the first stage's input
    // only; the later stages of the step start Newton from the same guesses.
    // KNC_GUESS_LEARNED_FOLD evaluates it on the first stage of a run of
    // steps with an unchanged block set (a fold, or the part of one between
    // excision events), whose stages all start from those guesses.
    // KNC_GUESS_PREVIOUS has no pass: each stage stores its own pressures.
    const int mode = c->o.c2p_guess;
    const bool guess_pass = learned(c) && (mode == KNC_GUESS_LEARNED || (mode == KNC_GUESS_LEARNED_STEP && first_stage) ||
                                           (mode == KNC_GUESS_LEARNED_FOLD && first_stage && segment_start));
    auto pre_pass = [&](const StageArgs& b, int nitems) -> int {
        const int grid = row_pass_grid(c, nitems);
        if (guess_pass) {
            HIP_TRY(launch_guess(b, grid, c->threads, c->stream));
            c->last_launches += 1;
        }
        if (with_dt) {
            StageArgs d = b;
            d.step = 0;
            HIP_TRY(launch_prim(c->h.system, stats, lg, d, grid, c->threads, c->stream, c->d_dt_partial));
            HIP_TRY(launch_dt_finish(c->d_dt_partial, grid, c->h.cfl_number, c->d_dt, c->stream));
            c->last_launches += 2;
        } else if (prims) {
            HIP_TRY(launch_prim(c->h.system, stats, lg, b, grid, c->threads, c->stream));
            c->last_launches += 1;
        }
        return KNC_OK;
    };
    int rc;
    if (c->o.launch_mode == KNC_LAUNCH_BATCHED) {
        a.items = c->d_items;
        a.nitems_ptr = c->d_nitems;
        if ((rc = pre_pass(a, c->nitems))) return rc;
        // the dt reduction needs a c2p pass: a fused form's first stage of
        // the run takes the two-pass form; the row-unit form is two-pass
        const int form = with_dt && stage_form(c) != STAGE_ROWS ? (int)STAGE_TWO_PASS : stage_form(c);
        const int grid = form == STAGE_ROWS ? row_pass_grid(c, c->nitems)
                                            : with_dt ? c->persistent_grid_two_pass : c->persistent_grid;
        HIP_TRY(launch_stage(c->h.system, c->h.riemann_solver, stats, lg, form, a, grid, c->threads, c->stream));
        c->last_launches += 1;
        return KNC_OK;
    }
    if (with_dt) return fail(KNC_ERR_UNSUPPORTED, "the dt reduction rides on batched stages only");
    for (auto& bi : c->block_items) {
        a.items = c->d_items + bi.first;
        a.nitems_ptr = nullptr;
        a.nitems = bi.second;
        if ((rc = pre_pass(a, bi.second))) return rc;
    }
    for (auto& bi : c->block_items) {
        a.items = c->d_items + bi.first;
        a.nitems_ptr = nullptr;
        a.nitems = bi.second;
        const int grid = stage_form(c) == STAGE_ROWS ? row_pass_grid(c, bi.second) : bi.second;
        HIP_TRY(launch_stage(c->h.system, c->h.riemann_solver, stats, lg, stage_form(c), a, grid, c->threads,
                             c->stream));
        c->last_launches += 1;
    }
    return KNC_OK;
}

// One RK step of RungeKuttaOrder (godunov-core) with the weighted averages of
// src/state.rs:179-212 fused into the last stage.  Rotates c->cur.
// Ghost-row sets of stage `stage` of fold step `fold_step` for the model
// boundaries (the stage kernel sees them as fixed rows).
void stage_boundaries(knc_context* c, StageArgs& a, int fold_step, int stage)
{
    if (c->trial) return;
    const size_t set = (size_t)10 * c->nq;
    const size_t base = ((size_t)fold_step * c->h.runge_kutta_order + stage) * 2 * set;
    if (model_side(c, 0)) {
        a.bnd.inner_mode = BND_FIXED;
        a.bnd.inner_rows = c->d_bseq + base;
    }
    if (model_side(c, 1)) {
        a.bnd.outer_mode = BND_FIXED;
        a.bnd.outer_rows = c->d_bseq + base + set;
    }
}

// Guess planes of a stage: the c2p starts from the plane of the stage's
// input slot (with KNC_GUESS_LEARNED_STEP, the plane the network wrote for
// the step's first stage; with KNC_GUESS_LEARNED_FOLD, the one it wrote for
// the first stage of the run of steps) and KNC_GUESS_PREVIOUS stores the recovered
// pressures into the plane of the output slot, which no workgroup of the
// launch reads.
// Error-key sequence number of stage `stage` of step `step` (1-based) of a
// run of steps: 2k for the k-th stage of the run, 2k - 1 for the model ghost
// rows it reads (evaluated before the stage, as scheme::advance refreshes
// the boundary before each stage); the time-step pass before the run is 0.
int stage_seq(const knc_context* c, int step, int stage, bool ghost = false)
{
    const int64_t k = (int64_t)(step - 1) * c->h.runge_kutta_order + stage + 1;
    return (int)std::min<int64_t>(2 * k - (ghost ? 1 : 0), (int64_t)ERR_SEQ_MAX);
}

void stage_slots(knc_context* c, StageArgs& a, int in, int out, int avg, double bweight)
{
    a.in_slot = in;
    a.out_slot = out;
    a.avg_slot = avg;
    a.bweight = bweight;
    a.guess_in = c->o.c2p_guess == KNC_GUESS_LEARNED_STEP ? c->cur
                 : c->o.c2p_guess == KNC_GUESS_LEARNED_FOLD ? c->segment_slot
                                                            : in;
    a.guess_out = out;
}

// `step`: 1 for the first step of the run of steps; `segment_start`: the
// first step of a run of steps over one block set (c->segment_slot holds its
// input slot).  Each stage logs its failures under its own sequence number.
int enqueue_rk_step(knc_context* c, int step, int fold_step, bool segment_start)
{
    const int f1 = (c->cur + 1) % NSLOT, f2 = (c->cur + 2) % NSLOT;
    StageArgs a = base_args(c);
    int rc;
    // the run's CFL reduction rides on its first stage's c2p pass
    const bool dt_here = step == 1 && c->dt_in_first_stage && !c->trial;
    switch (c->h.runge_kutta_order) {
    case 1:
        stage_slots(c, a, c->cur, f1, -1, 0.0);
        stage_boundaries(c, a, fold_step, 0);
        a.step = stage_seq(c, step, 0);
        if ((rc = launch_one_stage(c, a, true, segment_start, dt_here))) return rc;
        c->cur = f1;
        break;
    case 2:
        stage_slots(c, a, c->cur, f1, -1, 0.0);
        stage_boundaries(c, a, fold_step, 0);
        a.step = stage_seq(c, step, 0);
        if ((rc = launch_one_stage(c, a, true, segment_start, dt_here))) return rc;
        stage_slots(c, a, f1, f2, c->cur, 0.5);
        stage_boundaries(c, a, fold_step, 1);
        a.step = stage_seq(c, step, 1);
        if ((rc = launch_one_stage(c, a, false, segment_start))) return rc;
        c->cur = f2;
        break;
    default:
        stage_slots(c, a, c->cur, f1, -1, 0.0);
        stage_boundaries(c, a, fold_step, 0);
        a.step = stage_seq(c, step, 0);
        if ((rc = launch_one_stage(c, a, true, segment_start, dt_here))) return rc;
        stage_slots(c, a, f1, f2, c->cur, 0.75);
        stage_boundaries(c, a, fold_step, 1);
        a.step = stage_seq(c, step, 1);
        if ((rc = launch_one_stage(c, a, false, segment_start))) return rc;
        stage_slots(c, a, f2, f1, c->cur, 1.0 / 3.0);
        stage_boundaries(c, a, fold_step, 2);
        a.step = stage_seq(c, step, 2);
        if ((rc = launch_one_stage(c, a, false, segment_start))) return rc;
        c->cur = f1;
        break;
    }
    return KNC_OK;
}

// Host time after one RK step, in the reference's arithmetic
// (WeightedAverage on State::time, src/state.rs:204-208).
double advance_time(int order, double t0, double dt)
{
    if (order == 1) return t0 + dt;
    if (order == 2) {
        const double t2 = (t0 + dt) + dt;
        return t2 * (-0.5 + 1.0) + t0 * 0.5;
    }
    const double t1 = t0 + dt;
    const double t2 = (t1 + dt) * (-0.75 + 1.0) + t0 * 0.75;
    const double b3 = 1.0 / 3.0;
    return (t2 + dt) * (-b3 + 1.0) + t0 * b3;
}

// Enqueues the CFL reduction into d_dt (adaptive) or sets the fixed dt.
int enqueue_time_step(knc_context* c)
{
    if (c->h.system == KNC_SRHD && !c->h.adaptive_time_step) {
        const HostBlock& b = c->blocks.begin()->second;
        const double dt = c->h.cfl_number * (b.rv[1] - b.rv[0]) / c->h.light_speed;
        hipLaunchKernelGGL(k_set_double, dim3(1), dim3(64), 0, c->stream, c->d_dt, dt);
        HIP_TRY(hipGetLastError());
        return KNC_OK;
    }
    StageArgs a = base_args(c);
    a.in_slot = c->cur;
    a.step = 0;
    a.nitems_ptr = c->d_nitems;
    const int grid = row_pass_grid(c, c->nitems);
    HIP_TRY(launch_dt(c->h.system, a, c->d_dt_partial, grid, c->stream));
    HIP_TRY(launch_dt_finish(c->d_dt_partial, grid, c->h.cfl_number, c->d_dt, c->stream));
    return KNC_OK;
}

}  // namespace

int knc::collect_error(knc_context* c, knc_error* err)
{
    unsigned long long key = 0;
    HIP_TRY(hipMemcpy(&key, &c->d_err->min_key, sizeof key, hipMemcpyDeviceToHost));
    if (err) std::memset(err, 0, sizeof *err);
    if (key == ~0ull) return KNC_OK;
    ErrLog log;
    HIP_TRY(hipMemcpy(&log, c->d_err, sizeof log, hipMemcpyDeviceToHost));
    const int order = err_key_order(key), row = err_key_row(key), col = err_key_col(key);
    int code = KNC_HYDRO_ROOT_FINDER_FAILED;
    knc_error e{};
    e.i = row;
    e.j = col;
    if ((order == GHOST_INNER_ORDER || order == GHOST_OUTER_ORDER) && !c->blocks.empty()) {
        // a model ghost row: reported as the zone of the neighbour block the
        // reference's mesh puts there (rows block_size - 2, block_size - 1
        // inside the mesh, rows 0, 1 outside it)
        const int side = order == GHOST_OUTER_ORDER;
        const knc_mesh& m = c->bmodel[side].mesh;
        const int32_t inner = c->blocks.begin()->first, outer = c->blocks.rbegin()->first;
        e.block = side ? outer + 1 : inner - 1;
        e.i = side ? row : m.block_size - 2 + row;
        if (c->bmodel[side].on && row < 2 && col < c->nq) {
            e.r = ghost_rows(m, side, inner, outer).rc[row];
            e.theta = 0.5 * (c->qv[col] + c->qv[col + 1]);
        }
    }
    int o = 0;
    for (auto& kv : c->blocks) {
        if (o++ == order) {
            const HostBlock& b = kv.second;
            e.block = b.index;
            if (row < b.nr && col < c->nq) {
                e.r = std::sqrt(b.rv[row] * b.rv[row + 1]);
                e.theta = 0.5 * (c->qv[col] + c->qv[col + 1]);
            }
        }
    }
    const unsigned int n = std::min(log.count, (unsigned int)ERR_LOG_CAP);
    for (unsigned int k = 0; k < n; ++k) {
        if (log.entry[k].key == key) {
            code = log.entry[k].code;
            e.value = log.entry[k].value;
            for (int q = 0; q < 4; ++q) e.conserved[q] = log.entry[k].cons[q];
            break;
        }
    }
    e.code = code;
    if (err) *err = e;
    return fail(code, "hydro error %d at block %d zone (%d,%d) (r, theta) = (%.4e, %.4f)", code, e.block, e.i, e.j,
                e.r, e.theta);
}

int knc::reset_errors(knc_context* c)
{
    hipLaunchKernelGGL(k_reset_err, dim3(1), dim3(64), 0, c->stream, c->d_err);
    HIP_TRY(hipGetLastError());
    return KNC_OK;
}

namespace {

int require_blocks(knc_context* c)
{
    if (!c) return fail(KNC_ERR_INVALID_ARGUMENT, "context is NULL");
    if (c->blocks.empty()) return fail(KNC_ERR_NO_BLOCK, "no blocks");
    if (learned(c) && !c->has_mlp)
        return fail(KNC_ERR_INVALID_ARGUMENT, "c2p_guess is KNC_GUESS_LEARNED but no network is attached");
    return rebuild_tables(c);
}

// Enqueues n RK steps from slot c->cur: replayed from a hipGraph captured
// per (starting slot, n) in batched mode, launched one by one otherwise.
// A run is captured on the second use of its key since the block set last
// changed: the first runs directly, so that runs cut short by excision
// events, which may never repeat, do not pay for a capture.
int enqueue_steps(knc_context* c, int n, int first)
{
    if (n <= 0) return KNC_OK;
    const int start_slot = c->cur;
    c->segment_slot = start_slot;
    bool graph = c->o.launch_mode == KNC_LAUNCH_BATCHED && c->o.use_graph && n > 1;
    // the ghost-row pointers of model boundaries depend on the fold step
    const auto key = std::make_tuple(start_slot, n, any_model_side(c) ? first : 0, (int)c->dt_in_first_stage);
    if (graph && !c->graphs.count(key) && c->graph_keys_seen.insert(key).second) graph = false;
    if (!graph) {
        int rc;
        for (int s = 0; s < n; ++s)
            if ((rc = enqueue_rk_step(c, s + 1, first + s, s == 0))) return rc;
        return KNC_OK;
    }
    auto it = c->graphs.find(key);
    if (it == c->graphs.end()) {
        hipGraph_t g;
        HIP_TRY(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        int crc = KNC_OK;
        const int launches = c->last_launches;
        for (int s = 0; s < n && crc == KNC_OK; ++s) crc = enqueue_rk_step(c, s + 1, first + s, s == 0);
        hipError_t ce = hipStreamEndCapture(c->stream, &g);
        c->cur = start_slot;
        c->last_launches = launches;
        if (crc) return crc;
        HIP_TRY(ce);
        hipGraphExec_t ex;
        HIP_TRY(hipGraphInstantiate(&ex, g, nullptr, nullptr, 0));
        HIP_TRY(hipGraphDestroy(g));
        it = c->graphs.emplace(key, ex).first;
    }
    HIP_TRY(hipGraphLaunch(it->second, c->stream));
    // advance the slot bookkeeping as the captured steps did
    for (int s = 0; s < n; ++s) {
        const int f1 = (c->cur + 1) % NSLOT, f2 = (c->cur + 2) % NSLOT;
        c->cur = c->h.runge_kutta_order == 2 ? f2 : f1;
    }
    c->last_launches += n * c->h.runge_kutta_order;
    return KNC_OK;
}

bool excision_moving(const knc_context* c)
{
    return c->exc_on && (c->exc.inner_speed > 0.0 || c->exc.outer_speed > 0.0);
}

// Mesh::inner_excision_surface / outer_excision_surface (src/mesh.rs:318-336).
double excision_surface(double r0, double speed, double delay, double t)
{
    return r0 + std::max(t - delay, 0.0) * speed;
}

// The outer radius add_remove_blocks compares with the excision surfaces:
// mesh.subgrid_extent(index).outer_radius (src/scheme.rs:190,195,
// src/mesh.rs:358-372) when the context knows the reference mesh (device
// births or model boundaries), else the block's last radial vertex.
double block_outer_radius(const knc_context* c, const HostBlock& b)
{
    const DevModel* dm = c->birth.on ? &c->birth : c->bmodel[0].on ? &c->bmodel[0] : c->bmodel[1].on ? &c->bmodel[1]
                                                                                                        : nullptr;
    if (!dm) return b.rv[b.nr];
    const knc_mesh& m = dm->mesh;
    const double block = (double)m.block_size * mesh_zone_dlogr(m);
    return m.reference_radius * std::pow(1.0 + block, (double)b.index + 1.0);
}

// add_remove_blocks (src/scheme.rs:179-206) at state time t.  With apply
// false only reports whether the block set would change.  The innermost
// block is never removed when it is the only one (the reference would leave
// an empty state).
int excise(knc_context* c, double t, bool apply, bool* change)
{
    *change = false;
    const HostBlock& bi = c->blocks.begin()->second;
    const HostBlock& bo = c->blocks.rbegin()->second;
    const int32_t inner_index = bi.index, outer_index = bo.index;
    const bool remove = c->blocks.size() > 1 && block_outer_radius(c, bi) < excision_surface(c->exc.inner_radius,
                                                                                            c->exc.inner_speed,
                                                                                            c->exc.delay, t);
    const bool add = (c->exc.source != nullptr || c->birth.on) &&
                     block_outer_radius(c, bo) < excision_surface(c->exc.outer_radius, c->exc.outer_speed,
                                                                  c->exc.delay, t);
    *change = remove || add;
    if (!apply) return KNC_OK;
    int rc;
    if (remove) {
        if ((rc = knc_block_remove(c, inner_index))) return rc;
        c->exc_removed += 1;
    }
    if (add && !c->exc.source) {
        if ((rc = insert_block_model(c, outer_index + 1, c->birth, t))) return rc;
        c->exc_added += 1;
    } else if (add) {
        knc_block_data d{};
        const int32_t idx = outer_index + 1;
        const int32_t src_rc = c->exc.source(c->exc.user, idx, t, &d);
        if (src_rc) return fail(KNC_ERR_INVALID_ARGUMENT, "block source returned %d for block %d", src_rc, idx);
        if (!d.radial_vertices || !d.conserved || !d.scalar_mass)
            return fail(KNC_ERR_INVALID_ARGUMENT, "block source left a NULL buffer for block %d", idx);
        const std::vector<double> qv = c->qv;
        if ((rc = knc_block_insert(c, idx, d.nr, c->nq, d.radial_vertices, qv.data(), d.conserved, d.scalar_mass)))
            return rc;
        c->exc_added += 1;
    }
    return KNC_OK;
}

// Room for the ghost-row sets of every stage of a fold, and for their
// descriptors.
int ensure_boundary_sets(knc_context* c, int fold)
{
    const size_t nsets = (size_t)fold * c->h.runge_kutta_order * 2;
    const size_t need = nsets * 10 * c->nq;
    if (nsets > c->cap_rowsets) {
        if (c->d_rowsets) HIP_TRY(hipFree(c->d_rowsets));
        if (c->h_rowsets) HIP_TRY(hipHostFree(c->h_rowsets));
        c->d_rowsets = nullptr;
        c->h_rowsets = nullptr;
        c->cap_rowsets = 0;
        HIP_TRY(hipMalloc(&c->d_rowsets, nsets * sizeof(RowSet)));
        HIP_TRY(hipHostMalloc(&c->h_rowsets, nsets * sizeof(RowSet), hipHostMallocDefault));
        c->cap_rowsets = nsets;
    }
    if (need <= c->cap_bseq) return KNC_OK;
    drop_graphs(c);
    if (c->d_bseq) HIP_TRY(hipFree(c->d_bseq));
    if (c->h_bseq) HIP_TRY(hipHostFree(c->h_bseq));
    c->d_bseq = nullptr;
    c->h_bseq = nullptr;
    c->cap_bseq = 0;
    HIP_TRY(hipMalloc(&c->d_bseq, need * sizeof(double)));
    HIP_TRY(hipHostMalloc(&c->h_bseq, need * sizeof(double), hipHostMallocDefault));
    c->cap_bseq = need;
    return KNC_OK;
}

// Calls the boundary sources for every stage of fold step s, whose input
// state is at time t0: stage times t0, t0 + dt and, for RK3, the averaged
// time of the second stage's output (src/state.rs:189-208).
int fill_boundary_sets(knc_context* c, int s, double t0, double dt)
{
    const int order = c->h.runge_kutta_order;
    const double t1 = t0 + dt;
    const double times[3] = {t0, t1, (t1 + dt) * (-0.75 + 1.0) + t0 * 0.75};
    const int nq = c->nq;
    const int32_t inner = c->blocks.begin()->first, outer = c->blocks.rbegin()->first;
    for (int st = 0; st < order; ++st) c->stage_times[(size_t)s * order + st] = times[st];
    std::vector<double> prim((size_t)8 * nq), scal((size_t)2 * nq);
    for (int st = 0; st < order; ++st)
        for (int side = 0; side < 2; ++side) {
            if (!c->bsrc[side]) continue;
            const int32_t r = c->bsrc[side](c->buser[side], side, times[st], inner, outer, nq, prim.data(),
                                            scal.data());
            if (r) return fail(KNC_ERR_INVALID_ARGUMENT, "boundary source returned %d (side %d, t = %.6e)", r, side,
                               times[st]);
            double* d = c->h_bseq + (((size_t)s * order + st) * 2 + side) * 10 * nq;
            const int n2 = 2 * nq;
            for (int k = 0; k < n2; ++k) {
                for (int q = 0; q < 4; ++q) d[(size_t)q * n2 + k] = prim[4 * k + q];
                d[(size_t)4 * n2 + k] = scal[k];
            }
        }
    return KNC_OK;
}

// Enqueues fold steps [first, first + n), uploading their ghost-row sets
// first when a side has a model boundary.
int flush_steps(knc_context* c, int first, int n)
{
    if (n <= 0) return KNC_OK;
    if (c->bsrc[0] || c->bsrc[1]) {
        const size_t per_step = (size_t)c->h.runge_kutta_order * 2 * 10 * c->nq;
        HIP_TRY(hipMemcpyAsync(c->d_bseq + first * per_step, c->h_bseq + first * per_step,
                               n * per_step * sizeof(double), hipMemcpyHostToDevice, c->stream));
    }
    if (c->bmodel[0].on || c->bmodel[1].on) {
        // device-model sides, written after the host sets they share the
        // buffer with; the block set is fixed within a segment.  Every row
        // set of the segment is known here (the stage times are), so each
        // side's sets are built in one launch.  The descriptors of step s
        // sit at entries [2 order s, 2 order (s + 1)) of the pinned buffer,
        // which earlier segments of the fold do not reuse.
        const int order = c->h.runge_kutta_order;
        const int32_t inner = c->blocks.begin()->first, outer = c->blocks.rbegin()->first;
        for (int s = first; s < first + n; ++s)
            for (int st = 0; st < order; ++st)
                for (int side = 0; side < 2; ++side) {
                    const size_t e = ((size_t)s * order + st) * 2 + side;
                    c->h_rowsets[e] = RowSet{c->stage_times[(size_t)s * order + st],
                                             stage_seq(c, s - first + 1, st, true), 0, (int64_t)(e * 10 * c->nq)};
                }
        const size_t e0 = (size_t)first * order * 2, ne = (size_t)n * order * 2;
        HIP_TRY(hipMemcpyAsync(c->d_rowsets + e0, c->h_rowsets + e0, ne * sizeof(RowSet), hipMemcpyHostToDevice,
                               c->stream));
        for (int side = 0; side < 2; ++side) {
            if (!c->bmodel[side].on) continue;
            const GhostRows g = ghost_rows(c->bmodel[side].mesh, side, inner, outer);
            const int tag0 = (side ? GHOST_OUTER_ORDER : GHOST_INNER_ORDER) << 16;
            HIP_TRY(launch_model_rows(c->h.system, c->bmodel[side].dev, c->zp, c->nq, g, c->d_qc, c->th.dcost,
                                      c->d_rowsets + e0 + side, n * order, 2, tag0, c->d_bseq, c->d_err,
                                      c->stream));
        }
    }
    return enqueue_steps(c, n, first);
}

// Runs `fold` RK steps after an optional dt pass (scheme::advance,
// src/scheme.rs:212-240); returns after the stream drained and the error log
// was read.  With moving excision surfaces the steps run in segments between
// block-set changes; with moving surfaces or model boundaries dt is read back
// first, so the state time of every step is known on the host.
int run_steps(knc_context* c, bool with_dt, double dt_fixed, int fold, double* dt_used, knc_error* err)
{
    int rc;
    if ((rc = require_blocks(c))) return rc;
    c->last_launches = 0;
    const bool moving = with_dt && excision_moving(c);
    const bool model = any_model_side(c);
    if (model && (rc = ensure_boundary_sets(c, fold))) return rc;
    if (model) c->stage_times.assign((size_t)fold * c->h.runge_kutta_order, 0.0);
    hipLaunchKernelGGL(k_reset_err, dim3(1), dim3(64), 0, c->stream, c->d_err);
    HIP_TRY(hipGetLastError());
    const bool host_dt = with_dt && (moving || model);
    // Batched, with dt wanted only on the device and a CFL reduction to do:
    // the first stage's c2p pass computes dt (launch_one_stage), and the
    // state is read once for both instead of twice.
    c->dt_in_first_stage = with_dt && !host_dt && c->o.launch_mode == KNC_LAUNCH_BATCHED &&
                           !(c->h.system == KNC_SRHD && !c->h.adaptive_time_step);
    struct ClearOnExit {
        bool& flag;
        ~ClearOnExit() { flag = false; }
    } clear_dt_flag{c->dt_in_first_stage};
    if (with_dt) {
        if (!c->dt_in_first_stage && (rc = enqueue_time_step(c))) return rc;
    } else {
        hipLaunchKernelGGL(k_set_double, dim3(1), dim3(64), 0, c->stream, c->d_dt, dt_fixed);
        HIP_TRY(hipGetLastError());
    }
    double dt = dt_fixed;
    if (host_dt) {
        HIP_TRY(hipStreamSynchronize(c->stream));
        HIP_TRY(hipMemcpy(&dt, c->d_dt, sizeof dt, hipMemcpyDeviceToHost));
        if ((rc = collect_error(c, err))) return rc;
    }
    if (c->timing) HIP_TRY(hipEventRecord(c->ev0, c->stream));
    int pending = 0, first = 0;
    double t = c->time;
    int64_t iter = c->iteration;
    const int every = std::max(1, c->exc.check_every);
    for (int s = 0; s < fold; ++s) {
        if (moving && iter % every == 0) {
            bool change = false;
            if ((rc = excise(c, t, false, &change))) return rc;
            if (change) {
                if ((rc = flush_steps(c, first, pending))) return rc;
                first = s;
                pending = 0;
                HIP_TRY(hipStreamSynchronize(c->stream));
                if ((rc = collect_error(c, err))) return rc;
                c->time = t;
                c->iteration = iter;
                if ((rc = excise(c, t, true, &change))) return rc;
                if ((rc = require_blocks(c))) return rc;
            }
        }
        if (model && (rc = fill_boundary_sets(c, s, t, dt))) return rc;
        ++pending;
        t = advance_time(c->h.runge_kutta_order, t, dt);
        ++iter;
    }
    if ((rc = flush_steps(c, first, pending))) return rc;
    if (c->timing) HIP_TRY(hipEventRecord(c->ev1, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (!host_dt) HIP_TRY(hipMemcpy(&dt, c->d_dt, sizeof dt, hipMemcpyDeviceToHost));
    if (c->timing) {
        float ms = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, c->ev0, c->ev1));
        const int nst = fold * c->h.runge_kutta_order;
        c->last_stage_ms = nst ? ms / nst : 0.0;
    }
    if (dt_used) *dt_used = dt;
    if ((rc = collect_error(c, err))) return rc;
    if (host_dt || model) {
        c->time = t;
        c->iteration = iter;
    } else {
        for (int s = 0; s < fold; ++s) {
            c->time = advance_time(c->h.runge_kutta_order, c->time, dt);
            c->iteration += 1;
        }
    }
    return KNC_OK;
}

// ---- autotuner ----------------------------------------------------------
// Shapes already measured in this process, keyed by everything the stage
// launches depend on: configuration, device, polar size and block sizes.
std::mutex g_tune_mutex;
std::map<std::vector<int64_t>, std::vector<int>> g_tune_cache;

std::vector<int64_t> tune_key(const knc_context* c)
{
    std::vector<int64_t> k = {c->h.system, c->h.riemann_solver, c->h.runge_kutta_order, c->o.collect_stats,
                              guess_read(c) ? c->o.c2p_guess : 0, c->o.launch_mode, c->o.device, c->num_cu, c->nq};
    for (auto& kv : c->blocks) k.push_back(kv.second.nr);
    return k;
}

// Average device time of one stage over `trials` RK steps with the given
// shape.  The steps read slot c->cur and write the two other slots only, so
// the state is unchanged.
int time_shape(knc_context* c, int threads, int rows, int passes, int trials, double* ms)
{
    c->o.stage_threads = threads;
    c->o.rows_per_item = rows;
    c->o.stage_passes = passes;
    int rc;
    if ((rc = choose_stage_shape(c))) return rc;
    c->dirty = true;
    if ((rc = rebuild_tables(c))) return rc;
    const int saved = c->cur;
    c->trial = true;
    for (int w = 0; w < trials + 1; ++w) {
        if (w == 1) HIP_TRY(hipEventRecord(c->ev0, c->stream));
        c->segment_slot = c->cur;
        rc = enqueue_rk_step(c, 1, 0, true);
        c->cur = saved;
        if (rc) {
            c->trial = false;
            return rc;
        }
    }
    c->trial = false;
    HIP_TRY(hipEventRecord(c->ev1, c->stream));
    HIP_TRY(hipEventSynchronize(c->ev1));
    float t = 0.f;
    HIP_TRY(hipEventElapsedTime(&t, c->ev0, c->ev1));
    *ms = (double)t / (trials * c->h.runge_kutta_order);
    return KNC_OK;
}

int autotune(knc_context* c, int trials, bool force, knc_tune_result* out)
{
    int rc;
    if ((rc = require_blocks(c))) return rc;
    knc_tune_result r{};
    const std::vector<int64_t> key = tune_key(c);
    if (!force) {
        std::lock_guard<std::mutex> lock(g_tune_mutex);
        auto it = g_tune_cache.find(key);
        if (it != g_tune_cache.end()) {
            c->o.stage_threads = it->second[0];
            c->o.rows_per_item = it->second[1];
            c->o.stage_passes = it->second[2];
            if ((rc = choose_stage_shape(c))) return rc;
            c->dirty = true;
            if ((rc = rebuild_tables(c))) return rc;
            r.stage_threads = c->threads;
            r.rows_per_item = c->rows_per_item;
            r.stage_passes = c->o.stage_passes;
            r.from_cache = 1;
            if (out) *out = r;
            return KNC_OK;
        }
    }
    if (trials <= 0) trials = 3;
    HIP_TRY(hipStreamSynchronize(c->stream));
    // keep the statistics and the error log as they were
    C2PStats saved_stats;
    HIP_TRY(hipMemcpy(&saved_stats, c->d_stats, sizeof saved_stats, hipMemcpyDeviceToHost));
    const knc_options saved = c->o;
    // the trial stages use a fixed dt of zero, so every trial reads the same data
    hipLaunchKernelGGL(k_set_double, dim3(1), dim3(64), 0, c->stream, c->d_dt, 0.0);
    HIP_TRY(hipGetLastError());
    double best = 0.0;
    int best_t = 0, best_r = 0, best_p = 1;
    rc = time_shape(c, saved.stage_threads, saved.rows_per_item, saved.stage_passes, trials, &r.default_ms_per_stage);
    int ncand = 0;
    const int widths[] = {64, 128, 256, 512};
    const int rows[] = {4, 8, 16, 32, 64};
    for (int passes = 1; passes <= 4 && rc == KNC_OK; ++passes)
        for (int t : widths) {
            if (rc) break;
            if (!stage_variant_supported(t, passes)) continue;
            if (t > 64 && t / 2 >= c->nq) continue;   // wider than the mesh
            for (int nr : rows) {
                if (nr > stage_max_rows(passes)) continue;
                double ms = 0.0;
                if ((rc = time_shape(c, t, nr, passes, trials, &ms))) break;
                ++ncand;
                if (best_t == 0 || ms < best) {
                    best = ms;
                    best_t = t;
                    best_r = nr;
                    best_p = passes;
                }
            }
        }
    hipLaunchKernelGGL(k_reset_err, dim3(1), dim3(64), 0, c->stream, c->d_err);
    (void)hipMemcpy(c->d_stats, &saved_stats, sizeof saved_stats, hipMemcpyHostToDevice);
    c->o = saved;
    if (rc == KNC_OK && best_t) {
        c->o.stage_threads = best_t;
        c->o.rows_per_item = best_r;
        c->o.stage_passes = best_p;
    }
    int rc2 = choose_stage_shape(c);
    c->dirty = true;
    if (rc2 == KNC_OK) rc2 = rebuild_tables(c);
    if (rc) return rc;
    if (rc2) return rc2;
    {
        std::lock_guard<std::mutex> lock(g_tune_mutex);
        g_tune_cache[key] = {best_t, best_r, best_p};
    }
    r.stage_threads = best_t;
    r.rows_per_item = best_r;
    r.stage_passes = best_p;
    r.candidates = ncand;
    r.ms_per_stage = best;
    if (out) *out = r;
    return KNC_OK;
}

// Temporary device buffers for the per-zone entry points.
struct DevBuf {
    void* p = nullptr;
    ~DevBuf() { if (p) (void)hipFree(p); }
};

int zone_call_prep(const knc_hydro* h)
{
    int rc = check_hydro(h);
    if (rc) return rc;
    return KNC_OK;
}

}  // namespace

// ===========================================================================
extern "C" {

const char* knc_last_error_message(void) { return g_last_error.c_str(); }
int knc_abi_version(void) { return KNC_ABI_VERSION; }

int knc_context_create(const knc_hydro* hydro, const knc_options* options, knc_context** out)
{
    if (!out) return fail(KNC_ERR_INVALID_ARGUMENT, "out is NULL");
    *out = nullptr;
    int rc = check_hydro(hydro);
    if (rc) return rc;
    knc_context* c = new knc_context();
    c->h = *hydro;
    if (options) c->o = *options;
    else {
        c->o.launch_mode = KNC_LAUNCH_PER_BLOCK;
        c->o.use_graph = 0;
    }
    if (c->o.c2p_guess != KNC_GUESS_PLAIN && c->o.c2p_guess != KNC_GUESS_LEARNED &&
        c->o.c2p_guess != KNC_GUESS_LEARNED_STEP && c->o.c2p_guess != KNC_GUESS_PREVIOUS &&
        c->o.c2p_guess != KNC_GUESS_LEARNED_FOLD) {
        delete c;
        return fail(KNC_ERR_INVALID_ARGUMENT, "unknown c2p guess mode");
    }
    if (c->o.stage_passes < 0 || c->o.stage_passes > 4) {
        delete c;
        return fail(KNC_ERR_INVALID_ARGUMENT, "stage_passes must be 0, 1, 2, 3 or 4");
    }
    c->zp = zone_params(c->h);
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev <= 0) {
        delete c;
        return fail(KNC_ERR_HIP, "no HIP device available");
    }
    if (c->o.device < 0 || c->o.device >= ndev) {
        delete c;
        return fail(KNC_ERR_INVALID_ARGUMENT, "device %d out of range", c->o.device);
    }
    auto cleanup = [&](int code) { knc_context_destroy(c); return code; };
    if (hipSetDevice(c->o.device) != hipSuccess) return cleanup(fail(KNC_ERR_HIP, "hipSetDevice failed"));
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, c->o.device) != hipSuccess)
        return cleanup(fail(KNC_ERR_HIP, "hipGetDeviceProperties failed"));
    c->num_cu = prop.multiProcessorCount;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
        return cleanup(fail(KNC_ERR_HIP, "hipStreamCreate failed"));
    if (hipMalloc(&c->d_err, sizeof(ErrLog)) != hipSuccess || hipMalloc(&c->d_stats, sizeof(C2PStats)) != hipSuccess ||
        hipMalloc(&c->d_dt_partial, DT_MAX_GRID * sizeof(double)) != hipSuccess ||
        hipMalloc(&c->d_dt, sizeof(double)) != hipSuccess || hipMalloc(&c->d_nitems, sizeof(int32_t)) != hipSuccess)
        return cleanup(fail(KNC_ERR_HIP, "hipMalloc failed"));
    (void)hipMemset(c->d_stats, 0, sizeof(C2PStats));
    (void)hipMemset(c->d_dt, 0, sizeof(double));
    hipLaunchKernelGGL(k_reset_err, dim3(1), dim3(64), 0, c->stream, c->d_err);
    if (hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess)
        return cleanup(fail(KNC_ERR_HIP, "hipEventCreate failed"));
    if (hipStreamSynchronize(c->stream) != hipSuccess) return cleanup(fail(KNC_ERR_HIP, "init failed"));
    *out = c;
    return KNC_OK;
}

void knc_context_destroy(knc_context* c)
{
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    drop_graphs(c);
    for (auto& kv : c->blocks) {
        (void)hipFree(kv.second.mem);
        (void)hipFree(kv.second.rgeo);
        (void)hipFree(kv.second.zgeo);
        (void)hipFree(kv.second.guess);
        (void)hipFree(kv.second.prim);
    }
    free_model(c->bmodel[0]);
    free_model(c->bmodel[1]);
    free_model(c->birth);
    void* ptrs[] = {c->d_theta, c->d_bnd, c->d_blocks, c->d_items, c->d_nitems, c->d_err, c->d_stats,
                    c->d_dt_partial, c->d_dt, c->d_mlp, c->d_bseq, c->d_qc, c->d_rowsets};
    if (c->h_bseq) (void)hipHostFree(c->h_bseq);
    if (c->h_rowsets) (void)hipHostFree(c->h_rowsets);
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

int knc_block_insert(knc_context* c, int32_t index, int32_t nr, int32_t nq, const double* rv, const double* qv,
                     const double* conserved, const double* scalar_mass)
{
    if (!c || !rv || !qv || !conserved || !scalar_mass) return fail(KNC_ERR_INVALID_ARGUMENT, "NULL argument");
    return insert_block(c, index, nr, nq, rv, qv, conserved, scalar_mass);
}

// knc_block_insert; with conserved NULL the block starts at zero in every
// slot (the caller fills slot c->cur on the device).
extern "C++" int knc::insert_block(knc_context* c, int32_t index, int32_t nr, int32_t nq, const double* rv, const double* qv,
                 const double* conserved, const double* scalar_mass)
{
    if (nr < 2) return fail(KNC_ERR_INVALID_ARGUMENT, "must have at least 2 radial zones per block");
    if (nr >= 65536) return fail(KNC_ERR_INVALID_ARGUMENT, "nr must be below 65536");
    if (nq < 1 || nq > MAX_NQ) return fail(KNC_ERR_INVALID_ARGUMENT, "nq must be in [1, %d]", MAX_NQ);
    if (c->blocks.count(index)) return fail(KNC_ERR_INVALID_ARGUMENT, "block %d already present", index);
    if (int rc = check_cell_volumes(nr, nq, rv, qv)) return rc;
    if (c->blocks.size() >= (size_t)GHOST_INNER_ORDER) return fail(KNC_ERR_INVALID_ARGUMENT, "too many blocks");
    if (c->blocks.empty()) {
        int rc = upload_theta(c, nq, qv);
        if (rc) return rc;
    } else if (nq != c->nq || std::memcmp(qv, c->qv.data(), sizeof(double) * (nq + 1)) != 0) {
        return fail(KNC_ERR_INVALID_ARGUMENT, "every block must share the polar grid");
    } else {
        const int32_t lo = c->blocks.begin()->first, hi = c->blocks.rbegin()->first;
        if (index != lo - 1 && index != hi + 1)
            return fail(KNC_ERR_INVALID_ARGUMENT, "block %d is not adjacent to the chain [%d, %d]", index, lo, hi);
    }
    HIP_TRY(hipStreamSynchronize(c->stream));
    HostBlock b;
    b.index = index;
    b.nr = nr;
    b.nq = nq;
    b.rv.assign(rv, rv + nr + 1);
    const size_t n = (size_t)nr * nq;
    HIP_TRY(hipMalloc(&b.mem, (size_t)NSLOT * 5 * n * sizeof(double)));
    std::vector<double> rt = radial_table(nr, rv);
    if (hipMalloc(&b.rgeo, rt.size() * sizeof(double)) != hipSuccess) {
        (void)hipFree(b.mem);
        return fail(KNC_ERR_HIP, "hipMalloc failed");
    }
    if (hipMalloc(&b.guess, (size_t)NSLOT * n * sizeof(float)) != hipSuccess) {
        (void)hipFree(b.mem);
        (void)hipFree(b.rgeo);
        return fail(KNC_ERR_HIP, "hipMalloc failed");
    }
    if (hipMalloc(&b.prim, 5 * n * sizeof(double)) != hipSuccess) {
        (void)hipFree(b.mem);
        (void)hipFree(b.rgeo);
        (void)hipFree(b.guess);
        return fail(KNC_ERR_HIP, "hipMalloc failed");
    }
    HIP_TRY(hipMemcpy(b.rgeo, rt.data(), rt.size() * sizeof(double), hipMemcpyHostToDevice));
    {
        std::vector<double> zg((size_t)zgeo_size(nr, nq));
        fill_zone_geometry(nr, nq, rv, qv, zg.data());
        if (hipMalloc(&b.zgeo, zg.size() * sizeof(double)) != hipSuccess) {
            (void)hipFree(b.mem);
            (void)hipFree(b.rgeo);
            (void)hipFree(b.guess);
            (void)hipFree(b.prim);
            return fail(KNC_ERR_HIP, "hipMalloc failed");
        }
        HIP_TRY(hipMemcpy(b.zgeo, zg.data(), zg.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    HIP_TRY(hipMemset(b.mem, 0, (size_t)NSLOT * 5 * n * sizeof(double)));
    // log2(p/D) = -inf: no guess yet (2^-inf = 0, the plain guess)
    HIP_TRY(hipMemsetD32((hipDeviceptr_t)b.guess, 0xff800000u, (size_t)NSLOT * n));
    if (conserved) {
        std::vector<double> soa(5 * n);
        for (size_t k = 0; k < n; ++k) {
            soa[k] = conserved[4 * k];
            soa[n + k] = conserved[4 * k + 1];
            soa[2 * n + k] = conserved[4 * k + 2];
            soa[3 * n + k] = conserved[4 * k + 3];
            soa[4 * n + k] = scalar_mass[k];
        }
        HIP_TRY(hipMemcpy(b.mem + (size_t)c->cur * 5 * n, soa.data(), soa.size() * sizeof(double),
                          hipMemcpyHostToDevice));
    }
    c->blocks.emplace(index, b);
    c->dirty = true;
    return KNC_OK;
}

int knc_block_remove(knc_context* c, int32_t index)
{
    if (!c) return fail(KNC_ERR_INVALID_ARGUMENT, "context is NULL");
    auto it = c->blocks.find(index);
    if (it == c->blocks.end()) return fail(KNC_ERR_NO_BLOCK, "no block %d", index);
    if (it != c->blocks.begin() && std::next(it) != c->blocks.end())
        return fail(KNC_ERR_INVALID_ARGUMENT, "only the innermost or outermost block can be removed");
    HIP_TRY(hipStreamSynchronize(c->stream));
    HIP_TRY(hipFree(it->second.mem));
    HIP_TRY(hipFree(it->second.rgeo));
    HIP_TRY(hipFree(it->second.zgeo));
    HIP_TRY(hipFree(it->second.guess));
    HIP_TRY(hipFree(it->second.prim));
    c->blocks.erase(it);
    c->dirty = true;
    return KNC_OK;
}

int knc_block_download(knc_context* c, int32_t index, double* conserved, double* scalar_mass)
{
    if (!c || !conserved || !scalar_mass) return fail(KNC_ERR_INVALID_ARGUMENT, "NULL argument");
    auto it = c->blocks.find(index);
    if (it == c->blocks.end()) return fail(KNC_ERR_NO_BLOCK, "no block %d", index);
    const HostBlock& b = it->second;
    const size_t n = (size_t)b.nr * b.nq;
    std::vector<double> soa(5 * n);
    HIP_TRY(hipStreamSynchronize(c->stream));
    HIP_TRY(hipMemcpy(soa.data(), b.mem + (size_t)c->cur * 5 * n, soa.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (size_t k = 0; k < n; ++k) {
        conserved[4 * k] = soa[k];
        conserved[4 * k + 1] = soa[n + k];
        conserved[4 * k + 2] = soa[2 * n + k];
        conserved[4 * k + 3] = soa[3 * n + k];
        scalar_mass[k] = soa[4 * n + k];
    }
    return KNC_OK;
}

int knc_block_count(knc_context* c, int32_t* count, int32_t* inner_index, int32_t* outer_index)
{
    if (!c) return fail(KNC_ERR_INVALID_ARGUMENT, "context is NULL");
    if (count) *count = (int32_t)c->blocks.size();
    if (!c->blocks.empty()) {
        if (inner_index) *inner_index = c->blocks.begin()->first;
        if (outer_index) *outer_index = c->blocks.rbegin()->first;
    }
    return KNC_OK;
}

int knc_set_boundary(knc_context* c, int32_t side, int32_t mode, const double* primitive, const double* scalar)
{
    if (!c) return fail(KNC_ERR_INVALID_ARGUMENT, "context is NULL");
    if (side != KNC_SIDE_INNER && side != KNC_SIDE_OUTER) return fail(KNC_ERR_INVALID_ARGUMENT, "bad side");
    if (mode != KNC_BOUNDARY_FIXED && mode != KNC_BOUNDARY_OUTFLOW) return fail(KNC_ERR_INVALID_ARGUMENT, "bad mode");
    if (c->nq == 0) return fail(KNC_ERR_NO_BLOCK, "insert a block before setting boundaries");
    if (mode == KNC_BOUNDARY_FIXED) {
        if (!primitive || !scalar) return fail(KNC_ERR_INVALID_ARGUMENT, "fixed rows need primitive and scalar");
        const int n2 = 2 * c->nq;
        std::vector<double> soa((size_t)5 * n2);
        for (int k = 0; k < n2; ++k) {
            for (int q = 0; q < 4; ++q) soa[(size_t)q * n2 + k] = primitive[4 * k + q];
            soa[(size_t)4 * n2 + k] = scalar[k];
        }
        HIP_TRY(hipStreamSynchronize(c->stream));
        HIP_TRY(hipMemcpy(c->d_bnd + (size_t)side * 5 * n2, soa.data(), soa.size() * sizeof(double),
                          hipMemcpyHostToDevice));
    }
    if (side == KNC_SIDE_INNER) c->inner_mode = mode; else c->outer_mode = mode;
    drop_graphs(c);
    return KNC_OK;
}

int knc_set_time(knc_context* c, double time, int64_t iteration)
{
    if (!c) return fail(KNC_ERR_INVALID_ARGUMENT, "context is NULL");
    c->time = time;
    c->iteration = iteration;
    return KNC_OK;
}

int knc_get_time(knc_context* c, double* time, int64_t* iteration)
{
    if (!c) return fail(KNC_ERR_INVALID_ARGUMENT, "context is NULL");
    if (time) *time = c->time;
    if (iteration) *iteration = c->iteration;
    return KNC_OK;
}

int knc_time_step(knc_context* c, double* dt, knc_error* err)
{
    int rc;
    if ((rc = require_blocks(c))) return rc;
    hipLaunchKernelGGL(k_reset_err, dim3(1), dim3(64), 0, c->stream, c->d_err);
    HIP_TRY(hipGetLastError());
    if ((rc = enqueue_time_step(c))) return rc;
    HIP_TRY(hipStreamSynchronize(c->stream));
    double v = 0.0;
    HIP_TRY(hipMemcpy(&v, c->d_dt, sizeof v, hipMemcpyDeviceToHost));
    if ((rc = collect_error(c, err))) return rc;
    if (dt) *dt = v;
    return KNC_OK;
}

int knc_advance_steps(knc_context* c, double dt, int32_t nsteps, knc_error* err)
{
    if (nsteps < 0) return fail(KNC_ERR_INVALID_ARGUMENT, "nsteps < 0");
    if (nsteps == 0) return KNC_OK;
    return run_steps(c, false, dt, nsteps, nullptr, err);
}

int knc_advance(knc_context* c, int32_t fold, double* dt_used, knc_error* err)
{
    if (fold < 1) return fail(KNC_ERR_INVALID_ARGUMENT, "fold must be >= 1");
    return run_steps(c, true, 0.0, fold, dt_used, err);
}

int knc_set_boundary_source(knc_context* c, int32_t side, knc_boundary_fn fn, void* user)
{
    if (!c) return fail(KNC_ERR_INVALID_ARGUMENT, "context is NULL");
    if (side != KNC_SIDE_INNER && side != KNC_SIDE_OUTER) return fail(KNC_ERR_INVALID_ARGUMENT, "bad side");
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->bsrc[side] = fn;
    c->buser[side] = fn ? user : nullptr;
    if (fn) free_model(c->bmodel[side]);
    drop_graphs(c);
    return KNC_OK;
}

int knc_mesh_block_vertices(const knc_mesh* mesh, int32_t index, double* rv, double* qv)
{
    int rc;
    if ((rc = check_mesh(mesh))) return rc;
    if (!rv || !qv) return fail(KNC_ERR_INVALID_ARGUMENT, "NULL argument");
    mesh_vertices(*mesh, index, rv, qv);
    return KNC_OK;
}

int knc_model_eval(knc_context* c, const knc_model* model, int64_t n, const double* r, const double* q, double t,
                   double* prim, double* scal)
{
    if (!c || !model || (n > 0 && (!r || !q || !prim || !scal))) return fail(KNC_ERR_INVALID_ARGUMENT, "NULL argument");
    if (n < 0) return fail(KNC_ERR_INVALID_ARGUMENT, "n must be >= 0");
    if (n == 0) return KNC_OK;
    DevModel dm;
    int rc;
    if ((rc = load_model(dm, nullptr, model))) return rc;
    double* d = nullptr;
    int* dfail = nullptr;
    const size_t bytes = (size_t)n * 7 * sizeof(double);
    if (hipMalloc(&d, bytes) != hipSuccess || hipMalloc(&dfail, 2 * sizeof(int)) != hipSuccess) {
        if (d) (void)hipFree(d);
        free_model(dm);
        return fail(KNC_ERR_HIP, "hipMalloc failed");
    }
    int hfail[2] = {0, 0};
    hipError_t e = hipMemcpy(d, r, n * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d + n, q, n * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(dfail, 0, 2 * sizeof(int));
    if (e == hipSuccess) e = launch_model_eval(dm.dev, n, d, d + n, t, d + 2 * n, d + 6 * n, dfail, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e == hipSuccess) e = hipMemcpy(prim, d + 2 * n, (size_t)n * 4 * sizeof(double), hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(scal, d + 6 * n, n * sizeof(double), hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(hfail, dfail, 2 * sizeof(int), hipMemcpyDeviceToHost);
    (void)hipFree(d);
    (void)hipFree(dfail);
    free_model(dm);
    if (e != hipSuccess) return fail(KNC_ERR_HIP, "HIP error: %s", hipGetErrorString(e));
    if (hfail[0]) return fail(KNC_ERR_MODEL, "the model cannot be evaluated at point %d (the reference panics)", hfail[1]);
    return KNC_OK;
}

// BlockState::from_model of mesh block `index` at time t, inserted as a new
// block (src/state.rs:55-69 on the vertices of src/mesh.rs:202-264).
extern "C++" int knc::insert_block_model(knc_context* c, int32_t index, const DevModel& dm, double t)
{
    const knc_mesh& m = dm.mesh;
    const int nr = m.block_size, nq = m.num_polar_zones;
    std::vector<double> rv((size_t)nr + 1), qv((size_t)nq + 1);
    mesh_vertices(m, index, rv.data(), qv.data());
    int rc;
    if ((rc = insert_block(c, index, nr, nq, rv.data(), qv.data(), nullptr, nullptr))) return rc;
    const HostBlock& b = c->blocks.at(index);
    std::vector<double> tab((size_t)2 * nr);
    for (int i = 0; i < nr; ++i) {
        const double r0 = rv[i], r1 = rv[i + 1];
        tab[i] = std::sqrt(r0 * r1);
        tab[nr + i] = 2.0 * KNC_PI * ((r1 * r1) * r1 - (r0 * r0) * r0) / 3.0;
    }
    double* d = nullptr;
    int* dfail = nullptr;
    if (hipMalloc(&d, tab.size() * sizeof(double)) != hipSuccess || hipMalloc(&dfail, 2 * sizeof(int)) != hipSuccess) {
        if (d) (void)hipFree(d);
        (void)knc_block_remove(c, index);
        return fail(KNC_ERR_HIP, "hipMalloc failed");
    }
    int hfail[2] = {0, 0};
    const size_t n = (size_t)nr * nq;
    hipError_t e = hipMemcpy(d, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(dfail, 0, 2 * sizeof(int));
    if (e == hipSuccess)
        e = launch_block_from_model(c->h.system, dm.dev, c->zp, nr, nq, d, d + nr, c->d_qc, c->th.dcost, t,
                                    b.mem + (size_t)c->cur * 5 * n, dfail, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e == hipSuccess) e = hipMemcpy(hfail, dfail, 2 * sizeof(int), hipMemcpyDeviceToHost);
    (void)hipFree(d);
    (void)hipFree(dfail);
    if (e != hipSuccess) return fail(KNC_ERR_HIP, "HIP error: %s", hipGetErrorString(e));
    if (hfail[0]) {
        (void)knc_block_remove(c, index);
        return fail(KNC_ERR_MODEL, "the model cannot be evaluated in block %d zone %d (the reference panics)", index,
                    hfail[1]);
    }
    return KNC_OK;
}

int knc_block_insert_model(knc_context* c, int32_t index, const knc_mesh* mesh, const knc_model* model, double t)
{
    if (!c) return fail(KNC_ERR_INVALID_ARGUMENT, "context is NULL");
    int rc;
    if ((rc = check_mesh(mesh))) return rc;
    DevModel dm;
    if ((rc = load_model(dm, mesh, model))) return rc;
    rc = insert_block_model(c, index, dm, t);
    free_model(dm);
    return rc;
}

// The mesh's polar grid must be the blocks' (every block shares it).
static int check_mesh_polar(const knc_context* c, const knc_mesh& m)
{
    if (c->blocks.empty()) return KNC_OK;
    std::vector<double> qv((size_t)m.num_polar_zones + 1);
    mesh_vertices(m, c->blocks.begin()->first, nullptr, qv.data());
    if (m.num_polar_zones != c->nq || std::memcmp(qv.data(), c->qv.data(), qv.size() * sizeof(double)) != 0)
        return fail(KNC_ERR_INVALID_ARGUMENT, "the mesh's polar grid differs from the blocks'");
    return KNC_OK;
}

int knc_set_boundary_model(knc_context* c, int32_t side, const knc_mesh* mesh, const knc_model* model)
{
    if (!c) return fail(KNC_ERR_INVALID_ARGUMENT, "context is NULL");
    if (side != KNC_SIDE_INNER && side != KNC_SIDE_OUTER) return fail(KNC_ERR_INVALID_ARGUMENT, "bad side");
    HIP_TRY(hipStreamSynchronize(c->stream));
    drop_graphs(c);
    if (!model) {
        free_model(c->bmodel[side]);
        return KNC_OK;
    }
    int rc;
    if ((rc = check_mesh(mesh)) || (rc = check_mesh_polar(c, *mesh))) return rc;
    if ((rc = load_model(c->bmodel[side], mesh, model))) return rc;
    c->bsrc[side] = nullptr;
    c->buser[side] = nullptr;
    return KNC_OK;
}

int knc_set_birth_model(knc_context* c, const knc_mesh* mesh, const knc_model* model)
{
    if (!c) return fail(KNC_ERR_INVALID_ARGUMENT, "context is NULL");
    if (!model) {
        free_model(c->birth);
        return KNC_OK;
    }
    int rc;
    if ((rc = check_mesh(mesh)) || (rc = check_mesh_polar(c, *mesh))) return rc;
    return load_model(c->birth, mesh, model);
}

int knc_set_excision(knc_context* c, const knc_excision* ex)
{
    if (!c) return fail(KNC_ERR_INVALID_ARGUMENT, "context is NULL");
    if (!ex) {
        c->exc_on = false;
        c->exc = knc_excision{};
        return KNC_OK;
    }
    if (ex->check_every < 1) return fail(KNC_ERR_INVALID_ARGUMENT, "check_every must be >= 1");
    if (!(ex->inner_speed >= 0.0) || !(ex->outer_speed >= 0.0))
        return fail(KNC_ERR_INVALID_ARGUMENT, "excision speeds must be >= 0");
    c->exc = *ex;
    c->exc_on = true;
    return KNC_OK;
}

int knc_excision_counts(knc_context* c, int64_t* removed, int64_t* added)
{
    if (!c) return fail(KNC_ERR_INVALID_ARGUMENT, "context is NULL");
    if (removed) *removed = c->exc_removed;
    if (added) *added = c->exc_added;
    return KNC_OK;
}

int knc_autotune(knc_context* c, int32_t trials, int32_t force, knc_tune_result* out)
{
    if (!c) return fail(KNC_ERR_INVALID_ARGUMENT, "context is NULL");
    return autotune(c, trials, force != 0, out);
}

int knc_context_set_mlp(knc_context* c, const knc_mlp* mm)
{
    const MlpHost* m = reinterpret_cast<const MlpHost*>(mm);
    if (!c || !m) return fail(KNC_ERR_INVALID_ARGUMENT, "NULL argument");
    const size_t bytes = MlpHost::frag_bytes();
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (!c->d_mlp) HIP_TRY(hipMalloc(&c->d_mlp, bytes));
    HIP_TRY(hipMemcpy(c->d_mlp, m->d_frag, bytes, hipMemcpyDeviceToDevice));
    c->mlp = m->dev;
    c->mlp.image = (const uint4*)c->d_mlp;
    c->has_mlp = true;
    drop_graphs(c);
    return KNC_OK;
}

int knc_get_c2p_stats(knc_context* c, knc_c2p_stats* out)
{
    if (!c || !out) return fail(KNC_ERR_INVALID_ARGUMENT, "NULL argument");
    C2PStats s;
    HIP_TRY(hipStreamSynchronize(c->stream));
    HIP_TRY(hipMemcpy(&s, c->d_stats, sizeof s, hipMemcpyDeviceToHost));
    for (int k = 0; k < HIST_BINS; ++k) out->hist[k] = s.hist[k];
    uint64_t z = 0;
    for (int k = 0; k < HIST_BINS; ++k) z += s.hist[k];
    out->zones = z;
    out->fallbacks = s.fallbacks;
    out->floors = s.floors;
    out->learned_iters = s.learned_iters;
    return KNC_OK;
}

int knc_reset_c2p_stats(knc_context* c)
{
    if (!c) return fail(KNC_ERR_INVALID_ARGUMENT, "context is NULL");
    HIP_TRY(hipStreamSynchronize(c->stream));
    HIP_TRY(hipMemset(c->d_stats, 0, sizeof(C2PStats)));
    return KNC_OK;
}

int knc_synchronize(knc_context* c)
{
    if (!c) return fail(KNC_ERR_INVALID_ARGUMENT, "context is NULL");
    HIP_TRY(hipStreamSynchronize(c->stream));
    return KNC_OK;
}

int knc_set_timing(knc_context* c, int32_t enable)
{
    if (!c) return fail(KNC_ERR_INVALID_ARGUMENT, "context is NULL");
    c->timing = enable != 0;
    return KNC_OK;
}

int knc_last_stage_time(knc_context* c, double* ms, int32_t* launches)
{
    if (!c) return fail(KNC_ERR_INVALID_ARGUMENT, "context is NULL");
    if (ms) *ms = c->last_stage_ms;
    if (launches) *launches = c->last_launches;
    return KNC_OK;
}

// ---- per-zone entry points -------------------------------------------------
#define ZONE_ALLOC(buf, bytes) HIP_TRY(hipMalloc(&(buf).p, (bytes) ? (bytes) : 8))

int knc_hydro_try_to_primitive(const knc_hydro* h, int64_t n, const double* cons, const double* guess, double* prim,
                               int32_t* code, int32_t* iters)
{
    int rc = zone_call_prep(h);
    if (rc) return rc;
    if (n < 0 || (n && (!cons || !prim || !code))) return fail(KNC_ERR_INVALID_ARGUMENT, "bad arguments");
    if (n == 0) return KNC_OK;
    DevBuf dc, dg, dp, dcode, dit;
    ZONE_ALLOC(dc, n * 32);
    ZONE_ALLOC(dp, n * 32);
    ZONE_ALLOC(dcode, n * 4);
    if (guess) ZONE_ALLOC(dg, n * 8);
    if (iters) ZONE_ALLOC(dit, n * 4);
    HIP_TRY(hipMemcpy(dc.p, cons, n * 32, hipMemcpyHostToDevice));
    if (guess) HIP_TRY(hipMemcpy(dg.p, guess, n * 8, hipMemcpyHostToDevice));
    HIP_TRY(launch_try_to_primitive(h->system, zone_params(*h), n, (const double*)dc.p, (const double*)dg.p,
                                    (double*)dp.p, (int32_t*)dcode.p, (int32_t*)dit.p, nullptr));
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(hipMemcpy(prim, dp.p, n * 32, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(code, dcode.p, n * 4, hipMemcpyDeviceToHost));
    if (iters) HIP_TRY(hipMemcpy(iters, dit.p, n * 4, hipMemcpyDeviceToHost));
    return KNC_OK;
}

int knc_hydro_to_conserved(const knc_hydro* h, int64_t n, const double* prim, double* cons)
{
    int rc = zone_call_prep(h);
    if (rc) return rc;
    if (n < 0 || (n && (!cons || !prim))) return fail(KNC_ERR_INVALID_ARGUMENT, "bad arguments");
    if (n == 0) return KNC_OK;
    DevBuf dp, dc;
    ZONE_ALLOC(dp, n * 32);
    ZONE_ALLOC(dc, n * 32);
    HIP_TRY(hipMemcpy(dp.p, prim, n * 32, hipMemcpyHostToDevice));
    HIP_TRY(launch_to_conserved(h->system, zone_params(*h), n, (const double*)dp.p, (double*)dc.p, nullptr));
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(hipMemcpy(cons, dc.p, n * 32, hipMemcpyDeviceToHost));
    return KNC_OK;
}

int knc_hydro_intercell_flux(const knc_hydro* h, int64_t n, const double* pl, const double* pr, const double* sl,
                             const double* sr, int32_t direction, double* flux, double* sflux)
{
    int rc = zone_call_prep(h);
    if (rc) return rc;
    if (direction != KNC_DIRECTION_RADIAL && direction != KNC_DIRECTION_POLAR)
        return fail(KNC_ERR_INVALID_ARGUMENT, "bad direction");
    if (n < 0 || (n && (!pl || !pr || !sl || !sr || !flux || !sflux))) return fail(KNC_ERR_INVALID_ARGUMENT, "bad arguments");
    if (n == 0) return KNC_OK;
    DevBuf a, b, c, d, f, g;
    ZONE_ALLOC(a, n * 32);
    ZONE_ALLOC(b, n * 32);
    ZONE_ALLOC(c, n * 8);
    ZONE_ALLOC(d, n * 8);
    ZONE_ALLOC(f, n * 32);
    ZONE_ALLOC(g, n * 8);
    HIP_TRY(hipMemcpy(a.p, pl, n * 32, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(b.p, pr, n * 32, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(c.p, sl, n * 8, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(d.p, sr, n * 8, hipMemcpyHostToDevice));
    HIP_TRY(launch_intercell_flux(h->system, h->riemann_solver, direction, zone_params(*h), n, (const double*)a.p,
                                  (const double*)b.p, (const double*)c.p, (const double*)d.p, (double*)f.p,
                                  (double*)g.p, nullptr));
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(hipMemcpy(flux, f.p, n * 32, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(sflux, g.p, n * 8, hipMemcpyDeviceToHost));
    return KNC_OK;
}

int knc_hydro_geometrical_source_terms(const knc_hydro* h, int64_t n, const double* prim, const double* r,
                                       const double* theta, double* out)
{
    int rc = zone_call_prep(h);
    if (rc) return rc;
    if (n < 0 || (n && (!prim || !r || !theta || !out))) return fail(KNC_ERR_INVALID_ARGUMENT, "bad arguments");
    if (n == 0) return KNC_OK;
    std::vector<double> cot((size_t)n), rinv((size_t)n);
    for (int64_t k = 0; k < n; ++k) {
        cot[k] = std::cos(theta[k]) / std::sin(theta[k]);
        rinv[k] = 1.0 / r[k];
    }
    DevBuf dp, dr, dc, dout;
    ZONE_ALLOC(dp, n * 32);
    ZONE_ALLOC(dr, n * 8);
    ZONE_ALLOC(dc, n * 8);
    ZONE_ALLOC(dout, n * 32);
    HIP_TRY(hipMemcpy(dp.p, prim, n * 32, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(dr.p, rinv.data(), n * 8, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(dc.p, cot.data(), n * 8, hipMemcpyHostToDevice));
    HIP_TRY(launch_sources(h->system, zone_params(*h), n, (const double*)dp.p, (const double*)dr.p,
                           (const double*)dc.p, (double*)dout.p, nullptr));
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(hipMemcpy(out, dout.p, n * 32, hipMemcpyDeviceToHost));
    return KNC_OK;
}

int knc_hydro_max_signal_speed(const knc_hydro* h, int64_t n, const double* prim, double* out)
{
    int rc = zone_call_prep(h);
    if (rc) return rc;
    if (n < 0 || (n && (!prim || !out))) return fail(KNC_ERR_INVALID_ARGUMENT, "bad arguments");
    if (n == 0) return KNC_OK;
    DevBuf dp, dout;
    ZONE_ALLOC(dp, n * 32);
    ZONE_ALLOC(dout, n * 8);
    HIP_TRY(hipMemcpy(dp.p, prim, n * 32, hipMemcpyHostToDevice));
    HIP_TRY(launch_max_signal_speed(h->system, zone_params(*h), n, (const double*)dp.p, (double*)dout.p, nullptr));
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(hipMemcpy(out, dout.p, n * 8, hipMemcpyDeviceToHost));
    return KNC_OK;
}

int knc_hydro_plm_gradient(double theta, int64_t n, const double* a, const double* b, const double* c, double* out)
{
    if (n < 0 || (n && (!a || !b || !c || !out))) return fail(KNC_ERR_INVALID_ARGUMENT, "bad arguments");
    if (n == 0) return KNC_OK;
    DevBuf da, db, dc, dout;
    ZONE_ALLOC(da, n * 8);
    ZONE_ALLOC(db, n * 8);
    ZONE_ALLOC(dc, n * 8);
    ZONE_ALLOC(dout, n * 8);
    HIP_TRY(hipMemcpy(da.p, a, n * 8, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(db.p, b, n * 8, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(dc.p, c, n * 8, hipMemcpyHostToDevice));
    HIP_TRY(launch_plm(theta, n, (const double*)da.p, (const double*)db.p, (const double*)dc.p, (double*)dout.p,
                       nullptr));
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(hipMemcpy(out, dout.p, n * 8, hipMemcpyDeviceToHost));
    return KNC_OK;
}

int knc_fp64_fma_rate(int chains, int waves_per_simd, int iters, double* fma_per_s, float* ms)
{
    if (!(chains == 1 || chains == 2 || chains == 4 || chains == 8) || waves_per_simd < 1 || waves_per_simd > 8 ||
        iters < 1 || iters > (1 << 20) || !fma_per_s)
        return fail(KNC_ERR_INVALID_ARGUMENT, "bad arguments");
    int dev = 0, ncu = 0;
    HIP_TRY(hipGetDevice(&dev));
    HIP_TRY(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    // 256-thread workgroups (a wave on each SIMD), waves_per_simd of them per CU
    const int grid = ncu * waves_per_simd;
    DevBuf out;
    ZONE_ALLOC(out, (size_t)grid * 256 * 8);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIP_TRY(hipEventCreate(&e0));
    hipError_t err = hipEventCreate(&e1);
    if (err == hipSuccess) err = launch_fma_probe(chains, iters, grid, (double*)out.p, nullptr);   // warm-up
    if (err == hipSuccess) err = hipEventRecord(e0, nullptr);
    if (err == hipSuccess) err = launch_fma_probe(chains, iters, grid, (double*)out.p, nullptr);
    if (err == hipSuccess) err = hipEventRecord(e1, nullptr);
    if (err == hipSuccess) err = hipEventSynchronize(e1);
    float t = 0.0f;
    if (err == hipSuccess) err = hipEventElapsedTime(&t, e0, e1);
    (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    HIP_TRY(err);
    *fma_per_s = (double)grid * 256 * chains * iters / (t * 1e-3);
    if (ms) *ms = t;
    return KNC_OK;
}

int knc_fp64_probe(int64_t n, const double* a, const double* b, const double* c, double* div, double* sq,
                   double* mad, double* quot)
{
    if (n < 0 || (n && (!a || !b || !c || !div || !sq || !mad || !quot)))
        return fail(KNC_ERR_INVALID_ARGUMENT, "bad arguments");
    if (n == 0) return KNC_OK;
    DevBuf da, db, dc, d1, d2, d3, d4;
    ZONE_ALLOC(da, n * 8);
    ZONE_ALLOC(db, n * 8);
    ZONE_ALLOC(dc, n * 8);
    ZONE_ALLOC(d1, n * 8);
    ZONE_ALLOC(d2, n * 8);
    ZONE_ALLOC(d3, n * 8);
    ZONE_ALLOC(d4, n * 8);
    HIP_TRY(hipMemcpy(da.p, a, n * 8, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(db.p, b, n * 8, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(dc.p, c, n * 8, hipMemcpyHostToDevice));
    HIP_TRY(launch_fp64_probe(n, (const double*)da.p, (const double*)db.p, (const double*)dc.p, (double*)d1.p,
                              (double*)d2.p, (double*)d3.p, (double*)d4.p, nullptr));
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(hipMemcpy(div, d1.p, n * 8, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(sq, d2.p, n * 8, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(mad, d3.p, n * 8, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(quot, d4.p, n * 8, hipMemcpyDeviceToHost));
    return KNC_OK;
}

int knc_model_rows_check(knc_context* c, int32_t side, double time, double* rows, double* rows_one_thread)
{
    if (!c || !rows || !rows_one_thread) return fail(KNC_ERR_INVALID_ARGUMENT, "bad arguments");
    if (side != KNC_SIDE_INNER && side != KNC_SIDE_OUTER) return fail(KNC_ERR_INVALID_ARGUMENT, "bad side");
    int rc;
    if ((rc = require_blocks(c))) return rc;
    if (!c->bmodel[side].on) return fail(KNC_ERR_INVALID_ARGUMENT, "no device boundary model on this side");
    const size_t n = (size_t)10 * c->nq;
    DevBuf d1, d2, ds;
    ZONE_ALLOC(d1, n * 8);
    ZONE_ALLOC(d2, n * 8);
    ZONE_ALLOC(ds, sizeof(RowSet));
    const RowSet set{time, 0, 0, 0};
    HIP_TRY(hipMemcpy(ds.p, &set, sizeof set, hipMemcpyHostToDevice));
    const GhostRows g = ghost_rows(c->bmodel[side].mesh, side, c->blocks.begin()->first, c->blocks.rbegin()->first);
    const int tag0 = (side ? GHOST_OUTER_ORDER : GHOST_INNER_ORDER) << 16;
    HIP_TRY(hipStreamSynchronize(c->stream));
    HIP_TRY(launch_model_rows(c->h.system, c->bmodel[side].dev, c->zp, c->nq, g, c->d_qc, c->th.dcost,
                              (const RowSet*)ds.p, 1, 1, tag0, (double*)d1.p, c->d_err, c->stream));
    HIP_TRY(launch_model_rows(c->h.system, c->bmodel[side].dev, c->zp, c->nq, g, c->d_qc, c->th.dcost,
                              (const RowSet*)ds.p, 1, 1, tag0, (double*)d2.p, c->d_err, c->stream, true));
    HIP_TRY(hipStreamSynchronize(c->stream));
    HIP_TRY(hipMemcpy(rows, d1.p, n * 8, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(rows_one_thread, d2.p, n * 8, hipMemcpyDeviceToHost));
    // zones the model cannot evaluate are left in the rows as zeros; the
    // check leaves no entries in the context's error log
    return reset_errors(c);
}

}  // extern "C"
