// knc_records.hip -- the reference's output files from a context's resident
// state, and checkpoints read back into a context:
//   checkpoint  App {state, tasks, config, version}   (src/app.rs:179-186)
//   products    Products {time, blocks, config, version} (src/products.rs:32-37)
// in the serde layout ciborium gives them (src/io.rs:24-35).  The serde
// data model decides every byte: a struct is a map of its fields in
// declaration order keyed by name; an enum variant holding data is a
// one-entry map {variant: value}; an ndarray array is the struct
// {"v": 1u8, "dim": [..], "data": [elements in logical order]} (ndarray
// 0.14 serde support); a tuple or fixed array is an array; a HashMap is a
// map; a BlockIndex (i32, usize) key is [i, j]; a Rational64 is
// [numer, denom] (num-rational 0.3 serde support).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cerrno>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/knc_gpu.h"
#include "knc_cbor.hpp"
#include "knc_context.hpp"

using namespace knc;

namespace {

// BlockProducts::try_from_block_state (src/products.rs:44-66) for one block:
// try_to_primitive of U/dV for every zone (the NegativePressure floor
// accepted, src/physics/relativistic_hydro.rs:68-88), the primitive in
// AnyPrimitive order (u_r, u_theta, rho, p) as 4 doubles per zone, then the
// scalar concentration scalar_mass / lab_frame_mass = M / D as a plane of
// nr * nq doubles.  A failure is logged under (order << 16 | row, column).
template <int SYS>
__global__ void k_products(const double* U, int nr, int nq, const double* volA, const double* dcost, ZoneParams zp,
                           int order, ErrLog* err, double* out)
{
    const int64_t n = (int64_t)nr * nq;
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(k / nq), j = (int)(k - (int64_t)i * nq);
        const double ud = U[k];
        const double dv = volA[i] * dcost[j];
        const Cons u{ud / dv, U[n + k] / dv, U[2 * n + k] / dv, U[3 * n + k] / dv};
        Prim P{0.0, 0.0, 0.0, 0.0};
        int code;
        double value = u.d;
        if (SYS == SYS_EULER) {
            code = c2p_euler(u, zp, &P);
        } else if (u.d < 0.0) {
            code = E_NEGATIVE_DENSITY;
        } else if (u.tau < 0.0) {
            code = E_NEGATIVE_ENERGY_DENSITY;
            value = u.tau;
        } else {
            const double S2 = u.sr * u.sr + u.sq * u.sq;
            double p = 0.0, rW = 1.0;
            int it = 0;
            code = c2p_srhd_newton(u.d, S2, u.tau, zp, 0.0, &p, &it);
            if (code == E_OK) {
                code = c2p_srhd_finish(u, S2, p, &P, &rW);
                if (code == E_NEGATIVE_PRESSURE) code = E_OK;
            } else {
                value = u.tau;
            }
        }
        if (code != E_OK) {
            record_error(err, 0, (order << 16) | i, j, code, value, u);
            P = Prim{0.0, 0.0, 0.0, 0.0};
        }
        out[4 * k] = P.ur;
        out[4 * k + 1] = P.uq;
        out[4 * k + 2] = P.rho;
        out[4 * k + 3] = P.p;
        out[4 * n + k] = U[4 * n + k] / ud;
    }
}

int products_grid(int64_t n)
{
    const int64_t g = (n + 255) / 256;
    return (int)std::max<int64_t>(1, std::min<int64_t>(g, 8192));
}

// Two pinned host buffers, their copy events and a device scratch.  The
// destructor waits for the stream first, so that no copy is still writing
// a buffer it frees (an error path can leave one in flight).
struct Staging {
    hipStream_t stream = nullptr;
    double* host[2] = {nullptr, nullptr};
    hipEvent_t ev[2] = {nullptr, nullptr};
    double* dev = nullptr;
    explicit Staging(hipStream_t s) : stream(s) {}
    ~Staging()
    {
        if (stream) (void)hipStreamSynchronize(stream);
        for (int k = 0; k < 2; ++k) {
            if (host[k]) (void)hipHostFree(host[k]);
            if (ev[k]) (void)hipEventDestroy(ev[k]);
        }
        if (dev) (void)hipFree(dev);
    }
};

// Streams one device array per block (5 nr nq doubles from src(block)) to
// the host in block order through two pinned buffers: block k+1's copy is
// in flight while fn(block k, host data) encodes block k.
template <class Src, class Fn>
int stream_blocks(knc_context* c, Staging& st, Src src, Fn fn)
{
    std::vector<const HostBlock*> order;
    size_t maxn = 0;
    for (auto& kv : c->blocks) {
        order.push_back(&kv.second);
        maxn = std::max(maxn, (size_t)5 * kv.second.nr * kv.second.nq);
    }
    for (int k = 0; k < 2; ++k) {
        HIP_TRY(hipHostMalloc((void**)&st.host[k], maxn * sizeof(double), hipHostMallocDefault));
        HIP_TRY(hipEventCreateWithFlags(&st.ev[k], hipEventDisableTiming));
    }
    auto issue = [&](size_t k) -> int {
        const HostBlock& b = *order[k];
        const size_t bytes = (size_t)5 * b.nr * b.nq * sizeof(double);
        HIP_TRY(hipMemcpyAsync(st.host[k & 1], src(b, k), bytes, hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipEventRecord(st.ev[k & 1], c->stream));
        return KNC_OK;
    };
    int rc;
    if (!order.empty() && (rc = issue(0))) return rc;
    for (size_t k = 0; k < order.size(); ++k) {
        if (k + 1 < order.size() && (rc = issue(k + 1))) return rc;
        HIP_TRY(hipEventSynchronize(st.ev[k & 1]));
        if ((rc = fn(*order[k], st.host[k & 1]))) return rc;
    }
    return KNC_OK;
}

int check_meta(const knc_record_meta* m, bool tasks)
{
    if (!m) return fail(KNC_ERR_INVALID_ARGUMENT, "meta is NULL");
    if (!m->config || !m->config_len) return fail(KNC_ERR_INVALID_ARGUMENT, "meta has no config");
    if (tasks && (!m->tasks || !m->tasks_len)) return fail(KNC_ERR_INVALID_ARGUMENT, "meta has no tasks");
    if (!m->version) return fail(KNC_ERR_INVALID_ARGUMENT, "meta has no version");
    // Each supplied item must be one complete CBOR data item.
    const uint8_t* items[2] = {m->config, tasks ? m->tasks : nullptr};
    const uint64_t lens[2] = {m->config_len, tasks ? m->tasks_len : 0};
    for (int k = 0; k < 2; ++k) {
        if (!items[k]) continue;
        try {
            cbor::Reader r(items[k], (size_t)lens[k]);
            r.skip();
            if (!r.at_end()) throw cbor::Error("trailing bytes");
        } catch (const cbor::Error& e) {
            return fail(KNC_ERR_INVALID_ARGUMENT, "%s is not one CBOR item: %s", k ? "tasks" : "config", e.what());
        }
    }
    return KNC_OK;
}

// The ndarray struct header of an (nr, nq) array; the caller writes data.
void ndarray2_head(cbor::Writer& w, int nr, int nq)
{
    w.map(3);
    w.text("v");
    w.uint(1);
    w.text("dim");
    w.array(2);
    w.uint((uint64_t)nr);
    w.uint((uint64_t)nq);
    w.text("data");
    w.array((uint64_t)nr * nq);
}

void ndarray1(cbor::Writer& w, const double* x, size_t n)
{
    w.map(3);
    w.text("v");
    w.uint(1);
    w.text("dim");
    w.array(1);
    w.uint(n);
    w.text("data");
    w.array(n);
    w.f64s(x, n);
}

void block_key(cbor::Writer& w, int32_t index)
{
    w.array(2);
    w.sint(index);
    w.uint(0);
}

// Opens path for writing; on any failure the partial file is removed.
struct OutFile {
    std::string path;
    FILE* f = nullptr;
    bool keep = false;
    ~OutFile()
    {
        if (f) std::fclose(f);
        if (!keep && !path.empty()) std::remove(path.c_str());
    }
    int open(const char* p)
    {
        path = p;
        f = std::fopen(p, "wb");
        if (!f) {
            path.clear();
            return fail(KNC_ERR_IO, "cannot open %s for writing: %s", p, std::strerror(errno));
        }
        return KNC_OK;
    }
    int close(cbor::Writer& w)
    {
        w.flush();
        const bool ok = w.ok() && std::fflush(f) == 0;
        const bool closed = std::fclose(f) == 0;
        f = nullptr;
        if (!ok || !closed) return fail(KNC_ERR_IO, "writing %s failed", path.c_str());
        keep = true;
        return KNC_OK;
    }
};

}  // namespace

// ---------------------------------------------------------------------------
// The parsed checkpoint behind knc_checkpoint.
// ---------------------------------------------------------------------------
struct knc_checkpoint {
    std::vector<uint8_t> file;
    int32_t system = KNC_SRHD;
    double time = 0.0;
    int64_t it_numer = 0, it_denom = 1;
    size_t config_off = 0, config_len = 0, tasks_off = 0, tasks_len = 0;
    std::string version;
    struct Block {
        int32_t index;
        int32_t nr, nq;
        size_t conserved;     // offset of the data array of conserved
        size_t scalar_mass;   // offset of the data array of scalar_mass
    };
    std::vector<Block> blocks;
};

namespace {

// An ndarray struct of rank `rank` whose elements are `arity` numbers (an
// arity-tuple, or one number for arity 1): returns its dims, leaves the
// offset of its data item in *data and checks that the data holds exactly
// prod(dims) such elements.
void parse_ndarray(cbor::Reader& r, int rank, int arity, uint64_t* dims, size_t* data)
{
    bool have_v = false, have_dim = false, have_data = false;
    const uint64_t n = r.open(cbor::MAP);
    for (uint64_t k = 0; r.more(n, k); ++k) {
        const std::string key = r.text();
        if (key == "v") {
            if (r.sint() != 1) throw cbor::Error("unknown array version");
            have_v = true;
        } else if (key == "dim") {
            const uint64_t m = r.open(cbor::ARRAY);
            int q = 0;
            for (uint64_t a = 0; r.more(m, a); ++a) {
                if (q >= rank) throw cbor::Error("array rank mismatch");
                const int64_t d = r.sint();
                if (d < 0 || d > (1 << 24)) throw cbor::Error("array dimension out of range");
                dims[q++] = (uint64_t)d;
            }
            if (q != rank) throw cbor::Error("array rank mismatch");
            have_dim = true;
        } else if (key == "data") {
            *data = r.offset();
            r.skip();
            have_data = true;
        } else {
            throw cbor::Error("unknown field of an array");
        }
    }
    if (!have_v || !have_dim || !have_data) throw cbor::Error("array lacks v, dim or data");
    uint64_t count = 1;
    for (int q = 0; q < rank; ++q) count *= dims[q];
    const size_t end = r.offset();
    r.seek(*data);
    const uint64_t m = r.open(cbor::ARRAY);
    uint64_t k = 0;
    double tmp[4];
    for (; r.more(m, k); ++k) {
        if (k >= count) throw cbor::Error("array data longer than its dimensions");
        if (arity == 1) (void)r.f64();
        else r.f64_array((uint64_t)arity, tmp, 1);
    }
    if (k != count) throw cbor::Error("array data shorter than its dimensions");
    r.seek(end);
}

void parse_block(cbor::Reader& r, knc_checkpoint::Block& b)
{
    bool have_c = false, have_s = false;
    uint64_t dc[2] = {0, 0}, ds[2] = {0, 0};
    const uint64_t n = r.open(cbor::MAP);
    for (uint64_t k = 0; r.more(n, k); ++k) {
        const std::string key = r.text();
        if (key == "conserved") {
            parse_ndarray(r, 2, 4, dc, &b.conserved);
            have_c = true;
        } else if (key == "scalar_mass") {
            parse_ndarray(r, 2, 1, ds, &b.scalar_mass);
            have_s = true;
        } else {
            r.skip();
        }
    }
    if (!have_c || !have_s) throw cbor::Error("block state lacks conserved or scalar_mass");
    if (dc[0] != ds[0] || dc[1] != ds[1]) throw cbor::Error("conserved and scalar_mass shapes differ");
    if (dc[0] < 1 || dc[1] < 1 || dc[1] > (uint64_t)MAX_NQ) throw cbor::Error("block shape out of range");
    b.nr = (int32_t)dc[0];
    b.nq = (int32_t)dc[1];
}

void parse_state(cbor::Reader& r, knc_checkpoint& ck)
{
    // AnyState: {variant: State}, a map of one entry (definite or not)
    const uint64_t n = r.open(cbor::MAP);
    if (n != 1 && n != UINT64_MAX) throw cbor::Error("state is not a one-variant enum");
    if (n == UINT64_MAX && r.at_break()) throw cbor::Error("state is not a one-variant enum");
    const std::string variant = r.text();
    if (variant == "Relativistic") ck.system = KNC_SRHD;
    else if (variant == "Newtonian") ck.system = KNC_EULER;
    else throw cbor::Error("unknown state variant");
    bool have_t = false, have_i = false, have_s = false;
    const uint64_t m = r.open(cbor::MAP);
    for (uint64_t k = 0; r.more(m, k); ++k) {
        const std::string key = r.text();
        if (key == "time") {
            ck.time = r.f64();
            have_t = true;
        } else if (key == "iteration") {
            const uint64_t a = r.open(cbor::ARRAY);
            int64_t v[2];
            int q = 0;
            for (uint64_t e = 0; r.more(a, e); ++e) {
                if (q >= 2) throw cbor::Error("iteration is not [numer, denom]");
                v[q++] = r.sint();
            }
            if (q != 2 || v[1] == 0) throw cbor::Error("iteration is not [numer, denom]");
            ck.it_numer = v[0];
            ck.it_denom = v[1];
            have_i = true;
        } else if (key == "solution") {
            const uint64_t nb = r.open(cbor::MAP);
            for (uint64_t e = 0; r.more(nb, e); ++e) {
                knc_checkpoint::Block b{};
                const uint64_t a = r.open(cbor::ARRAY);
                int64_t idx[2];
                int q = 0;
                for (uint64_t x = 0; r.more(a, x); ++x) {
                    if (q >= 2) throw cbor::Error("block index is not [i, j]");
                    idx[q++] = r.sint();
                }
                if (q != 2 || idx[0] < INT32_MIN || idx[0] > INT32_MAX || idx[1] != 0)
                    throw cbor::Error("block index is not [i, 0]");
                b.index = (int32_t)idx[0];
                parse_block(r, b);
                ck.blocks.push_back(b);
            }
            have_s = true;
        } else {
            r.skip();
        }
    }
    if (!have_t || !have_i || !have_s) throw cbor::Error("state lacks time, iteration or solution");
    if (n == UINT64_MAX) r.take_break();
    std::sort(ck.blocks.begin(), ck.blocks.end(),
              [](const knc_checkpoint::Block& a, const knc_checkpoint::Block& b) { return a.index < b.index; });
    for (size_t k = 1; k < ck.blocks.size(); ++k)
        if (ck.blocks[k].index == ck.blocks[k - 1].index) throw cbor::Error("duplicate block index");
}

void parse_app(knc_checkpoint& ck)
{
    cbor::Reader r(ck.file.data(), ck.file.size());
    bool have[4] = {false, false, false, false};
    const uint64_t n = r.open(cbor::MAP);
    for (uint64_t k = 0; r.more(n, k); ++k) {
        const std::string key = r.text();
        if (key == "state") {
            parse_state(r, ck);
            have[0] = true;
        } else if (key == "tasks") {
            ck.tasks_off = r.offset();
            r.skip();
            ck.tasks_len = r.offset() - ck.tasks_off;
            have[1] = true;
        } else if (key == "config") {
            ck.config_off = r.offset();
            r.skip();
            ck.config_len = r.offset() - ck.config_off;
            have[2] = true;
        } else if (key == "version") {
            ck.version = r.text();
            have[3] = true;
        } else {
            r.skip();
        }
    }
    if (!r.at_end()) throw cbor::Error("trailing bytes after the App");
    if (!have[0] || !have[1] || !have[2] || !have[3]) throw cbor::Error("App lacks state, tasks, config or version");
}

// Decodes block b into AoS conserved (nr, nq, 4) and scalar_mass (nr, nq).
void decode_block(const knc_checkpoint& ck, const knc_checkpoint::Block& b, double* cons, double* sm)
{
    const uint64_t n = (uint64_t)b.nr * b.nq;
    cbor::Reader r(ck.file.data(), ck.file.size());
    r.seek(b.conserved);
    const uint64_t m = r.open(cbor::ARRAY);
    uint64_t k = 0;
    for (; r.more(m, k); ++k) {
        if (k >= n) throw cbor::Error("conserved data longer than its dimensions");
        r.f64_array(4, cons + 4 * k, 1);
    }
    if (k != n) throw cbor::Error("conserved data shorter than its dimensions");
    r.seek(b.scalar_mass);
    r.f64_array(n, sm, 1);
}

}  // namespace

extern "C" {

int knc_write_checkpoint(knc_context* c, const char* path, const knc_record_meta* meta)
{
    if (!c || !path) return fail(KNC_ERR_INVALID_ARGUMENT, "NULL argument");
    int rc;
    if ((rc = check_meta(meta, true))) return rc;
    HIP_TRY(hipStreamSynchronize(c->stream));
    OutFile out;
    if ((rc = out.open(path))) return rc;
    cbor::Writer w(out.f);
    w.map(4);
    w.text("state");
    w.map(1);
    w.text(c->h.system == KNC_SRHD ? "Relativistic" : "Newtonian");
    w.map(3);
    w.text("time");
    w.f64(c->time);
    w.text("iteration");
    w.array(2);
    w.sint(c->iteration);
    w.sint(1);
    w.text("solution");
    w.map(c->blocks.size());
    Staging st(c->stream);
    rc = stream_blocks(
        c, st, [&](const HostBlock& b, size_t) { return b.mem + (size_t)c->cur * 5 * b.nr * b.nq; },
        [&](const HostBlock& b, const double* h) {
            const size_t n = (size_t)b.nr * b.nq;
            block_key(w, b.index);
            w.map(2);
            w.text("conserved");
            ndarray2_head(w, b.nr, b.nq);
            for (size_t k = 0; k < n; ++k) {
                w.array(4);
                w.f64(h[k]);
                w.f64(h[n + k]);
                w.f64(h[2 * n + k]);
                w.f64(h[3 * n + k]);
            }
            w.text("scalar_mass");
            ndarray2_head(w, b.nr, b.nq);
            w.f64s(h + 4 * n, n);
            return KNC_OK;
        });
    if (rc) return rc;
    w.text("tasks");
    w.raw(meta->tasks, (size_t)meta->tasks_len);
    w.text("config");
    w.raw(meta->config, (size_t)meta->config_len);
    w.text("version");
    w.text(meta->version);
    return out.close(w);
}

int knc_write_products(knc_context* c, const char* path, const knc_record_meta* meta, knc_error* err)
{
    if (err) std::memset(err, 0, sizeof *err);
    if (!c || !path) return fail(KNC_ERR_INVALID_ARGUMENT, "NULL argument");
    int rc;
    if ((rc = check_meta(meta, false))) return rc;
    if (c->blocks.empty()) return fail(KNC_ERR_NO_BLOCK, "no blocks");
    // c2p of every block into one device scratch, then check for failures
    // before anything is written.
    Staging st(c->stream);
    size_t total = 0;
    std::vector<size_t> offset;
    for (auto& kv : c->blocks) {
        offset.push_back(total);
        total += (size_t)5 * kv.second.nr * kv.second.nq;
    }
    HIP_TRY(hipMalloc((void**)&st.dev, total * sizeof(double)));
    if ((rc = reset_errors(c))) return rc;
    int order = 0;
    for (auto& kv : c->blocks) {
        const HostBlock& b = kv.second;
        const int64_t n = (int64_t)b.nr * b.nq;
        const double* U = b.mem + (size_t)c->cur * 5 * n;
        BlockDesc d{};
        d.rgeo = b.rgeo;
        d.nr = b.nr;
        const double* volA = rg_volA(d);
        double* o = st.dev + offset[order];
        if (c->h.system == KNC_SRHD)
            hipLaunchKernelGGL(k_products<SYS_SRHD>, dim3(products_grid(n)), dim3(256), 0, c->stream, U, b.nr, b.nq,
                               volA, c->th.dcost, c->zp, order, c->d_err, o);
        else
            hipLaunchKernelGGL(k_products<SYS_EULER>, dim3(products_grid(n)), dim3(256), 0, c->stream, U, b.nr, b.nq,
                               volA, c->th.dcost, c->zp, order, c->d_err, o);
        HIP_TRY(hipGetLastError());
        ++order;
    }
    HIP_TRY(hipStreamSynchronize(c->stream));
    if ((rc = collect_error(c, err))) return rc;

    OutFile out;
    if ((rc = out.open(path))) return rc;
    cbor::Writer w(out.f);
    w.map(4);
    w.text("time");
    w.f64(c->time);
    w.text("blocks");
    w.map(c->blocks.size());
    rc = stream_blocks(
        c, st, [&](const HostBlock&, size_t k) { return st.dev + offset[k]; },
        [&](const HostBlock& b, const double* h) {
            const size_t n = (size_t)b.nr * b.nq;
            block_key(w, b.index);
            w.map(4);
            w.text("radial_vertices");
            ndarray1(w, b.rv.data(), b.rv.size());
            w.text("polar_vertices");
            ndarray1(w, c->qv.data(), c->qv.size());
            w.text("primitive");
            ndarray2_head(w, b.nr, b.nq);
            for (size_t k = 0; k < n; ++k) {
                w.array(4);
                w.f64s(h + 4 * k, 4);
            }
            w.text("scalar");
            ndarray2_head(w, b.nr, b.nq);
            w.f64s(h + 4 * n, n);
            return KNC_OK;
        });
    if (rc) return rc;
    w.text("config");
    w.raw(meta->config, (size_t)meta->config_len);
    w.text("version");
    w.text(meta->version);
    return out.close(w);
}

int knc_checkpoint_open(const char* path, knc_checkpoint** out)
{
    if (!path || !out) return fail(KNC_ERR_INVALID_ARGUMENT, "NULL argument");
    *out = nullptr;
    std::unique_ptr<knc_checkpoint> ck(new (std::nothrow) knc_checkpoint);
    if (!ck) return fail(KNC_ERR_INVALID_ARGUMENT, "out of host memory");
    FILE* f = std::fopen(path, "rb");
    if (!f) return fail(KNC_ERR_IO, "cannot open %s: %s", path, std::strerror(errno));
    std::fseek(f, 0, SEEK_END);
    const long size = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    bool ok = size >= 0;
    if (ok) {
        try {
            ck->file.resize((size_t)size);
        } catch (const std::bad_alloc&) {
            ok = false;
        }
    }
    ok = ok && std::fread(ck->file.data(), 1, (size_t)size, f) == (size_t)size;
    std::fclose(f);
    if (!ok) return fail(KNC_ERR_IO, "cannot read %s", path);
    try {
        parse_app(*ck);
    } catch (const cbor::Error& e) {
        return fail(KNC_ERR_FORMAT, "%s is not a checkpoint: %s", path, e.what());
    }
    *out = ck.release();
    return KNC_OK;
}

int knc_checkpoint_get_info(const knc_checkpoint* ck, knc_checkpoint_info* out)
{
    if (!ck || !out) return fail(KNC_ERR_INVALID_ARGUMENT, "NULL argument");
    out->system = ck->system;
    out->num_blocks = (int32_t)ck->blocks.size();
    out->time = ck->time;
    out->iteration_numer = ck->it_numer;
    out->iteration_denom = ck->it_denom;
    out->config = ck->file.data() + ck->config_off;
    out->config_len = ck->config_len;
    out->tasks = ck->file.data() + ck->tasks_off;
    out->tasks_len = ck->tasks_len;
    out->version = ck->version.c_str();
    return KNC_OK;
}

int knc_checkpoint_block_shape(const knc_checkpoint* ck, int32_t k, int32_t* index, int32_t* nr, int32_t* nq)
{
    if (!ck) return fail(KNC_ERR_INVALID_ARGUMENT, "NULL argument");
    if (k < 0 || k >= (int32_t)ck->blocks.size()) return fail(KNC_ERR_INVALID_ARGUMENT, "block %d out of range", k);
    const knc_checkpoint::Block& b = ck->blocks[(size_t)k];
    if (index) *index = b.index;
    if (nr) *nr = b.nr;
    if (nq) *nq = b.nq;
    return KNC_OK;
}

int knc_checkpoint_block_data(const knc_checkpoint* ck, int32_t k, double* conserved, double* scalar_mass)
{
    if (!ck || !conserved || !scalar_mass) return fail(KNC_ERR_INVALID_ARGUMENT, "NULL argument");
    if (k < 0 || k >= (int32_t)ck->blocks.size()) return fail(KNC_ERR_INVALID_ARGUMENT, "block %d out of range", k);
    try {
        decode_block(*ck, ck->blocks[(size_t)k], conserved, scalar_mass);
    } catch (const cbor::Error& e) {
        return fail(KNC_ERR_FORMAT, "block %d: %s", ck->blocks[(size_t)k].index, e.what());
    }
    return KNC_OK;
}

int knc_checkpoint_load(const knc_checkpoint* ck, knc_context* c, const knc_mesh* mesh)
{
    if (!ck || !c || !mesh) return fail(KNC_ERR_INVALID_ARGUMENT, "NULL argument");
    if (!c->blocks.empty()) return fail(KNC_ERR_INVALID_ARGUMENT, "the context already holds blocks");
    if (ck->system != c->h.system)
        return fail(KNC_ERR_INVALID_ARGUMENT, "the checkpoint's hydrodynamics system is not the context's");
    if (ck->it_denom != 1) return fail(KNC_ERR_UNSUPPORTED, "the checkpoint's iteration is not whole");
    if (ck->blocks.empty()) return fail(KNC_ERR_NO_BLOCK, "the checkpoint holds no blocks");
    for (const knc_checkpoint::Block& b : ck->blocks)
        if (b.nr != mesh->block_size || b.nq != mesh->num_polar_zones)
            return fail(KNC_ERR_INVALID_ARGUMENT, "block %d is %d x %d, the mesh's blocks %d x %d", b.index, b.nr,
                        b.nq, mesh->block_size, mesh->num_polar_zones);
    std::vector<double> rv, qv, cons, sm;
    for (size_t k = 0; k < ck->blocks.size(); ++k) {
        const knc_checkpoint::Block& b = ck->blocks[k];
        rv.resize((size_t)b.nr + 1);
        qv.resize((size_t)b.nq + 1);
        cons.resize((size_t)4 * b.nr * b.nq);
        sm.resize((size_t)b.nr * b.nq);
        int rc = knc_mesh_block_vertices(mesh, b.index, rv.data(), qv.data());
        if (!rc) rc = knc_checkpoint_block_data(ck, (int32_t)k, cons.data(), sm.data());
        if (!rc) rc = insert_block(c, b.index, b.nr, b.nq, rv.data(), qv.data(), cons.data(), sm.data());
        if (rc) {
            // leave the context empty, as it was
            std::vector<int32_t> held;
            for (auto& kv : c->blocks) held.push_back(kv.first);
            const std::string msg = knc_last_error_message();
            for (int32_t i : held) (void)knc_block_remove(c, i);
            return fail(rc, "%s", msg.c_str());
        }
    }
    c->time = ck->time;
    c->iteration = ck->it_numer;
    return KNC_OK;
}

void knc_checkpoint_close(knc_checkpoint* ck) { delete ck; }

int knc_cbor_read_f64(const uint8_t* data, uint64_t len, uint64_t offset, uint64_t count, int32_t arity,
                      double* out, uint64_t* end)
{
    if (!data || !out || arity < 1 || offset > len) return fail(KNC_ERR_INVALID_ARGUMENT, "bad argument");
    try {
        cbor::Reader r(data, (size_t)len);
        r.seek((size_t)offset);
        if (arity == 1) {
            r.f64_array(count, out, 1);
        } else {
            const uint64_t n = r.open(cbor::ARRAY);
            uint64_t k = 0;
            for (; r.more(n, k); ++k) {
                if (k >= count) throw cbor::Error("array longer than expected");
                r.f64_array((uint64_t)arity, out + k * (uint64_t)arity, 1);
            }
            if (k != count) throw cbor::Error("array shorter than expected");
        }
        if (end) *end = r.offset();
    } catch (const cbor::Error& e) {
        return fail(KNC_ERR_FORMAT, "not an array of %llu numbers x %d: %s", (unsigned long long)count, arity,
                    e.what());
    }
    return KNC_OK;
}

}  // extern "C"
