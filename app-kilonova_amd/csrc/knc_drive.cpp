// knc_drive.cpp -- a compiled caller of the C ABI (include/knc_gpu.h), the
// way the reference's Rust host would drive it: build a mesh of radial
// blocks, set the boundaries, optionally autotune, and call knc_advance
// (scheme::advance, src/scheme.rs:212-240) fold by fold, reporting FP64
// zone-updates per second like the reference's Mzps (src/main.rs:42-48).
//
//   knc_drive [nr] [nq] [nblocks] [steps] [fold] [batched 0|1] [checkpoint.cbor]
//
// With a checkpoint path it then writes the state as the reference's App
// file (knc_write_checkpoint; empty configuration and tasks items), reads
// it back (knc_checkpoint_open) and checks every block against the state
// still resident on the device.
// Workload: SRHD, gamma = 4/3, r in [1, 100] log-spaced, theta in [0, pi];
// rho = r^-2 (1 + 0.1 cos 4 theta), radial four-velocity u = 1, p = 1e-2 rho;
// a fixed inner inflow (u = 3) and an outflow outer boundary; HLLE, RK2,
// CFL 0.4, adaptive time step.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/knc_gpu.h"

namespace {

const double PI = 3.14159265358979323846;

// to_conserved of hydro_srhd (u = Gamma beta), per unit volume.
void prim_to_cons(double gamma, const double* P, double* U)
{
    const double W = std::sqrt(1.0 + (P[1] * P[1] + P[2] * P[2]));
    const double rhoh = P[0] + P[3] * (gamma / (gamma - 1.0));
    const double hw = rhoh * W;
    U[0] = P[0] * W;
    U[1] = hw * P[1];
    U[2] = hw * P[2];
    U[3] = (hw * W - P[3]) - U[0];
}

int check(int rc, const char* what)
{
    if (rc != KNC_OK) std::fprintf(stderr, "knc_drive: %s failed (%d): %s\n", what, rc, knc_last_error_message());
    return rc;
}

// The state as a checkpoint file, read back and compared block by block
// with the device state.
int write_and_verify(knc_context* ctx, const char* path, int nblocks, int nr, int nq)
{
    const uint8_t empty_map = 0xa0;   // CBOR {}
    knc_record_meta meta{&empty_map, 1, &empty_map, 1, "knc_drive"};
    if (check(knc_write_checkpoint(ctx, path, &meta), "knc_write_checkpoint")) return 1;
    knc_checkpoint* ck = nullptr;
    if (check(knc_checkpoint_open(path, &ck), "knc_checkpoint_open")) return 1;
    knc_checkpoint_info info{};
    int rc = check(knc_checkpoint_get_info(ck, &info), "knc_checkpoint_get_info");
    double t = 0.0;
    int64_t iteration = 0;
    knc_get_time(ctx, &t, &iteration);
    if (!rc && (info.num_blocks != nblocks || info.time != t || info.iteration_numer != iteration ||
                info.iteration_denom != 1 || info.system != KNC_SRHD)) {
        std::fprintf(stderr, "knc_drive: checkpoint header differs from the state\n");
        rc = 1;
    }
    const size_t n = (size_t)nr * nq;
    std::vector<double> c0(4 * n), s0(n), c1(4 * n), s1(n);
    for (int k = 0; k < nblocks && !rc; ++k) {
        int32_t index = -1, bnr = 0, bnq = 0;
        rc = check(knc_checkpoint_block_shape(ck, k, &index, &bnr, &bnq), "knc_checkpoint_block_shape");
        if (!rc && (index != k || bnr != nr || bnq != nq)) rc = 1;
        if (!rc) rc = check(knc_checkpoint_block_data(ck, k, c1.data(), s1.data()), "knc_checkpoint_block_data");
        if (!rc) rc = check(knc_block_download(ctx, k, c0.data(), s0.data()), "knc_block_download");
        if (!rc && (std::memcmp(c0.data(), c1.data(), c0.size() * sizeof(double)) ||
                    std::memcmp(s0.data(), s1.data(), s0.size() * sizeof(double)))) {
            std::fprintf(stderr, "knc_drive: block %d of the checkpoint differs from the state\n", k);
            rc = 1;
        }
    }
    knc_checkpoint_close(ck);
    if (!rc) std::printf("knc_drive: checkpoint %s written and read back, %d blocks identical\n", path, nblocks);
    return rc;
}

}  // namespace

int main(int argc, char** argv)
{
    const int nr = argc > 1 ? std::atoi(argv[1]) : 1024;
    const int nq = argc > 2 ? std::atoi(argv[2]) : 256;
    const int nblocks = argc > 3 ? std::atoi(argv[3]) : 8;
    const int steps = argc > 4 ? std::atoi(argv[4]) : 100;
    const int fold = argc > 5 ? std::atoi(argv[5]) : 10;
    const int batched = argc > 6 ? std::atoi(argv[6]) : 1;
    const char* checkpoint = argc > 7 ? argv[7] : nullptr;
    if (nr <= 0 || nq <= 0 || nblocks <= 0 || nr % nblocks || steps <= 0 || fold <= 0) {
        std::fprintf(stderr,
                     "usage: knc_drive [nr] [nq] [nblocks] [steps] [fold] [batched 0|1] [checkpoint.cbor]\n");
        return 2;
    }
    const double gamma = 4.0 / 3.0;
    knc_hydro h{KNC_SRHD, KNC_HLLE, 2, 1, gamma, 1.5, 0.4, 3e10};
    knc_options o{};
    o.launch_mode = batched ? KNC_LAUNCH_BATCHED : KNC_LAUNCH_PER_BLOCK;
    o.c2p_guess = KNC_GUESS_PLAIN;
    o.use_graph = 1;
    knc_context* ctx = nullptr;
    if (check(knc_context_create(&h, &o, &ctx), "knc_context_create")) return 1;

    std::vector<double> qv(nq + 1);
    for (int j = 0; j <= nq; ++j) qv[j] = PI / nq * j;
    const int nb = nr / nblocks;
    const double dy = 2.0 / nr;
    int rc = 0;
    for (int b = 0; b < nblocks && !rc; ++b) {
        std::vector<double> rv(nb + 1), cons((size_t)nb * nq * 4), sm((size_t)nb * nq, 0.0);
        for (int i = 0; i <= nb; ++i) rv[i] = std::pow(10.0, dy * (b * nb + i));
        for (int i = 0; i < nb; ++i)
            for (int j = 0; j < nq; ++j) {
                const double r = std::sqrt(rv[i] * rv[i + 1]);
                const double q = 0.5 * (qv[j] + qv[j + 1]);
                const double rho = 1.0 / (r * r) * (1.0 + 0.1 * std::cos(4.0 * q));
                const double P[4] = {rho, 1.0, 0.0, 1e-2 * rho};
                const double dv = 2.0 * PI * (rv[i + 1] * rv[i + 1] * rv[i + 1] - rv[i] * rv[i] * rv[i]) / 3.0 *
                                  (std::cos(qv[j]) - std::cos(qv[j + 1]));
                double* U = &cons[((size_t)i * nq + j) * 4];
                prim_to_cons(gamma, P, U);
                for (int k = 0; k < 4; ++k) U[k] *= dv;
            }
        rc = check(knc_block_insert(ctx, b, nb, nq, rv.data(), qv.data(), cons.data(), sm.data()),
                   "knc_block_insert");
    }
    std::vector<double> gp((size_t)2 * nq * 4), gs((size_t)2 * nq, 1.0);
    for (int k = 0; k < 2 * nq; ++k) {
        gp[4 * k] = 1.0;
        gp[4 * k + 1] = 3.0;
        gp[4 * k + 2] = 0.0;
        gp[4 * k + 3] = 1e-2;
    }
    if (!rc) rc = check(knc_set_boundary(ctx, KNC_SIDE_INNER, KNC_BOUNDARY_FIXED, gp.data(), gs.data()),
                        "knc_set_boundary");
    if (!rc) rc = check(knc_set_boundary(ctx, KNC_SIDE_OUTER, KNC_BOUNDARY_OUTFLOW, nullptr, nullptr),
                        "knc_set_boundary");
    knc_tune_result tr{};
    if (!rc) rc = check(knc_autotune(ctx, 2, 0, &tr), "knc_autotune");
    if (!rc) rc = check(knc_advance(ctx, fold, nullptr, nullptr), "warm-up knc_advance");
    if (!rc) rc = check(knc_synchronize(ctx), "knc_synchronize");
    const auto t0 = std::chrono::steady_clock::now();
    int done = 0;
    knc_error err{};
    double dt = 0.0;
    while (!rc && done < steps) {
        const int f = steps - done < fold ? steps - done : fold;
        rc = knc_advance(ctx, f, &dt, &err);
        if (rc) std::fprintf(stderr, "knc_drive: hydro error %d at block %d zone (%d, %d): %s\n", rc, err.block,
                             err.i, err.j, knc_last_error_message());
        done += f;
    }
    const double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    double t = 0.0;
    int64_t iteration = 0;
    knc_get_time(ctx, &t, &iteration);
    if (!rc)
        std::printf("knc_drive: %d x %d zones, %d blocks, %s, shape %d x %d (%d-pass), %d steps in %.3f s: "
                    "%.3f G zone-updates/s, t = %.6e s, iteration %lld\n",
                    nr, nq, nblocks, batched ? "batched" : "per-block", tr.stage_threads, tr.rows_per_item,
                    tr.stage_passes, steps, wall, (double)nr * nq * steps / wall / 1e9, t, (long long)iteration);
    if (!rc && checkpoint) rc = write_and_verify(ctx, checkpoint, nblocks, nb, nq);
    knc_context_destroy(ctx);
    return rc ? 1 : 0;
}
