/* knc_gpu.h -- C ABI of the MI355X (gfx950) implementation of the KNC
 * per-block RK stage.
 *
 * Drop-in boundary for the hot path of clemson-cal/app-kilonova:
 *   - the per-scheme seam  scheme::advance      (src/scheme.rs:212-240),
 *                          State::time_step     (src/state.rs:143-162),
 *                          the block set of State (src/state.rs:24-43) and
 *                          add_remove_blocks    (src/scheme.rs:179-206);
 *   - the per-zone seam    trait Hydrodynamics  (src/traits.rs:41-136),
 *     implemented by RelativisticHydro (src/physics/relativistic_hydro.rs)
 *     and NewtonianHydro (src/physics/newtonian_hydro.rs).
 *
 * Conventions: every function returns 0 on success or a KNC_ERR_* /
 * KNC_HYDRO_* code, never unwinds and never aborts; knc_last_error_message()
 * describes the last failure of the calling thread.  Host buffers belong to
 * the caller and are copied in or out before the call returns; the context
 * owns all device memory.  One context drives one HIP stream on one device
 * and must not be used from two threads at once.
 *
 * Array layouts follow the reference's ndarray layouts: a block state is
 * (nr, nq) row-major with theta fastest; conserved is an array of 4-double
 * tuples (D, S_r, S_theta, tau) for SRHD or (rho, m_r, m_theta, E) for
 * Euler; primitives are (rho, u_r, u_theta, p) with u = Gamma*beta for SRHD
 * and u = v for Euler (hydro_srhd::srhd_2d::Primitive,
 * src/physics/relativistic_hydro.rs:110-112).
 */
#ifndef KNC_GPU_H
#define KNC_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define KNC_ABI_VERSION 4

/* hydrodynamics system and options */
enum { KNC_EULER = 0, KNC_SRHD = 1 };
enum { KNC_HLLE = 0, KNC_HLLC = 1 };
enum { KNC_DIRECTION_RADIAL = 0, KNC_DIRECTION_POLAR = 1 };
enum { KNC_BOUNDARY_FIXED = 0, KNC_BOUNDARY_OUTFLOW = 1 };
enum { KNC_SIDE_INNER = 0, KNC_SIDE_OUTER = 1 };
enum { KNC_LAUNCH_PER_BLOCK = 0, KNC_LAUNCH_BATCHED = 1 };
/* c2p initial pressure: PLAIN p = 0; LEARNED the network's guess for every
 * stage's input; LEARNED_STEP the network evaluated on the first stage's
 * input of each RK step, its guesses reused by the step's later stages;
 * LEARNED_FOLD the network evaluated on the first stage's input of each run
 * of steps over one block set (the steps of a knc_advance fold, split at
 * excision events), its guesses reused by every later stage of the run;
 * PREVIOUS the pressure each zone's previous stage recovered (p = 0 for a
 * zone with none yet: a new block's first stage).  Newton ends in a
 * guess-independent polish, so every mode gives the plain mode's state. */
enum {
    KNC_GUESS_PLAIN = 0,
    KNC_GUESS_LEARNED = 1,
    KNC_GUESS_LEARNED_STEP = 2,
    KNC_GUESS_PREVIOUS = 3,
    KNC_GUESS_LEARNED_FOLD = 4
};

/* return codes: hydro errors mirror HydroErrorType (src/physics/mod.rs:82-95) */
enum {
    KNC_OK = 0,
    KNC_HYDRO_NEGATIVE_DENSITY = 1,
    KNC_HYDRO_NEGATIVE_PRESSURE = 2,
    KNC_HYDRO_NEGATIVE_ENERGY_DENSITY = 3,
    KNC_HYDRO_ROOT_FINDER_FAILED = 4,
    KNC_ERR_INVALID_ARGUMENT = 10,
    KNC_ERR_HIP = 11,
    KNC_ERR_NO_BLOCK = 12,
    KNC_ERR_UNSUPPORTED = 13,
    /* an initial model was evaluated where the reference panics: wind_shock
     * outside its flares without a table or outside the table's range,
     * halo_kilonova below the galactic midplane */
    KNC_ERR_MODEL = 14,
    /* a file could not be opened, read or written (io::Error, src/io.rs:16-17) */
    KNC_ERR_IO = 15,
    /* a checkpoint is not an App of the expected shape (ciborium::de::Error,
     * src/io.rs:13-14) */
    KNC_ERR_FORMAT = 16
};

/* Hydrodynamics configuration: the fields of RelativisticHydro
 * (src/physics/relativistic_hydro.rs:13-36) and NewtonianHydro
 * (src/physics/newtonian_hydro.rs:13-28).  light_speed scales SRHD fluxes,
 * sources and signal speeds (LIGHT_SPEED = 3e10, src/physics/mod.rs:7); it
 * must be 1 for Euler. */
typedef struct knc_hydro {
    int32_t system;              /* KNC_EULER | KNC_SRHD */
    int32_t riemann_solver;      /* KNC_HLLE | KNC_HLLC */
    int32_t runge_kutta_order;   /* 1, 2 or 3 */
    int32_t adaptive_time_step;  /* SRHD: 0 = cfl*dr0/c, 1 = CFL reduction; Euler: always reduction */
    double gamma_law_index;
    double plm_theta;            /* [1, 2] */
    double cfl_number;           /* [0, 0.7] */
    double light_speed;
} knc_hydro;

/* Execution options (no counterpart in the reference). */
typedef struct knc_options {
    int32_t launch_mode;     /* KNC_LAUNCH_PER_BLOCK: one stage launch per block;
                                KNC_LAUNCH_BATCHED: one persistent launch per stage
                                over every block, fold steps replayed as a hipGraph */
    int32_t c2p_guess;       /* KNC_GUESS_PLAIN | _LEARNED | _LEARNED_STEP | _PREVIOUS | _LEARNED_FOLD */
    int32_t collect_stats;   /* 1: Newton iteration histogram and fallback counts */
    int32_t rows_per_item;   /* radial rows per work item (<= 64); 0 = chosen from the mesh */
    int32_t device;          /* HIP device ordinal */
    int32_t use_graph;       /* batched mode: 1 = capture fold steps in a hipGraph */
    int32_t stage_threads;   /* stage workgroup size = polar columns per work item
                                (64, 128, 256, 512); 0 = chosen from nq */
    int32_t stage_passes;    /* 1: fused stage (c2p inside the stage kernel);
                                2: a c2p pass writing primitive planes, then the
                                fused reconstruction -> flux -> update pass;
                                3: fused stage with the c2p and the flux work on
                                separate wavefronts of each workgroup (128 or 256
                                polar columns per work item, at most 32 rows);
                                4: the c2p pass, then a flux pass over (work
                                item, row) units, all rows at once (for meshes
                                of few-row blocks);
                                0 = 1 (knc_autotune may choose any) */
} knc_options;

/* HydroError (src/physics/mod.rs:109-117): code, the failing zone and its
 * (r, theta) centre, the offending value, and U/dV at that zone. */
typedef struct knc_error {
    int32_t code;
    int32_t block;        /* caller's block index (BlockIndex.0) */
    int32_t i;
    int32_t j;
    double r;
    double theta;
    double value;
    double conserved[4];
} knc_error;

/* Newton statistics of the SRHD c2p. */
typedef struct knc_c2p_stats {
    uint64_t hist[64];       /* zones by Newton iteration count */
    uint64_t zones;
    uint64_t fallbacks;      /* learned guesses that fell back to the plain guess */
    uint64_t floors;         /* NegativePressure floors (p = 1e-3 rho) */
    uint64_t learned_iters;  /* Newton iterations spent from learned guesses */
} knc_c2p_stats;

typedef struct knc_context knc_context;

const char* knc_last_error_message(void);
int knc_abi_version(void);

/* ---- context and block set (State, src/state.rs:26-43) ----------------- */
int knc_context_create(const knc_hydro* hydro, const knc_options* options, knc_context** out);
void knc_context_destroy(knc_context* ctx);

/* Inserts block `index` (BlockIndex.0; the chain must stay contiguous).
 * radial_vertices: nr+1, polar_vertices: nq+1 (GridGeometry::radial_vertices
 * / polar_vertices, src/mesh.rs:21-29; every block shares the polar grid).
 * conserved: (nr, nq, 4) volume-integrated; scalar_mass: (nr, nq)
 * (BlockState, src/state.rs:26-30).  Replaces add_remove_blocks' insert
 * (src/scheme.rs:197-205). */
int knc_block_insert(knc_context* ctx, int32_t index, int32_t nr, int32_t nq, const double* radial_vertices,
                     const double* polar_vertices, const double* conserved, const double* scalar_mass);
/* Removes block `index` (src/scheme.rs:191-194). */
int knc_block_remove(knc_context* ctx, int32_t index);
int knc_block_download(knc_context* ctx, int32_t index, double* conserved, double* scalar_mass);
int knc_block_count(knc_context* ctx, int32_t* count, int32_t* inner_index, int32_t* outer_index);

/* Ghost rows beyond the inner (side 0) or outer (side 1) block: FIXED takes
 * primitive (2*nq*4) and scalar concentration rho*X (2*nq) rows in radial
 * order; OUTFLOW repeats the edge row (zero gradient).  Stands in for the
 * model-built ghost blocks of src/scheme.rs:46-52. */
int knc_set_boundary(knc_context* ctx, int32_t side, int32_t mode, const double* primitive, const double* scalar);

/* Ghost rows from the initial model at every stage's input time, as
 * try_advance_rk builds them (BlockState::from_model on the boundary block
 * mesh.subgrid(inner_index - 1) for side 0 or mesh.subgrid(outer_index + 1)
 * for side 1 at state.time, then c2p; src/scheme.rs:46-52).  The callback
 * fills the two rows next to the mesh edge, in radial order: primitive
 * (2*nq*4) and scalar concentration rho*X (2*nq).  It is called for every
 * stage of an advance call before the steps are launched (the stage times
 * follow from the fold's dt); it returns 0, or nonzero to abort the call. */
typedef int32_t (*knc_boundary_fn)(void* user, int32_t side, double time, int32_t inner_index,
                                   int32_t outer_index, int32_t nq, double* primitive, double* scalar);
/* Makes `side` a model boundary fed by fn; fn = NULL returns it to the rows
 * of knc_set_boundary. */
int knc_set_boundary_source(knc_context* ctx, int32_t side, knc_boundary_fn fn, void* user);

int knc_set_time(knc_context* ctx, double time, int64_t iteration);
int knc_get_time(knc_context* ctx, double* time, int64_t* iteration);

/* State::time_step (src/state.rs:143-162). */
int knc_time_step(knc_context* ctx, double* dt, knc_error* err);

/* ---- allocation-free time step (no counterpart in the reference) ------- */
/* For callers that cannot allocate device memory: code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc.  The CFL
 * time step that knc_time_step (State::time_step, src/state.rs:143-162)
 * returns for a context holding this one block, of a block whose state lives
 * in the caller's device memory: `conserved` is a device pointer, on the
 * current device, to the block's volume-integrated conserved planes
 * D | S_r | S_theta | tau, each nr*nq doubles with theta fastest (the layout
 * a context keeps resident); the vertices are host arrays (nr+1, nq+1).  The
 * work runs on `stream` (a hipStream_t of the current device, or NULL) and
 * the call returns after it.  The library allocates nothing: the kernels'
 * tables, partial minima, dt and error log live in a file-scope device array
 * of KNC_NOALLOC_SCRATCH_BYTES that each device holds until the process
 * ends; nr above KNC_NOALLOC_MAX_NR or nq above KNC_NOALLOC_MAX_NQ is
 * KNC_ERR_INVALID_ARGUMENT.  Every call on a device shares that device's
 * array, which keeps the last call's contents: calls from one process take
 * turns under a lock, and calls that overlapped on one device would race on
 * it.  Errors come back as from knc_time_step, with err->block = index. */
#define KNC_NOALLOC_MAX_NR 8192
#define KNC_NOALLOC_MAX_NQ 1024
#define KNC_NOALLOC_SCRATCH_BYTES (576 * 1024)
int knc_block_time_step_noalloc(const knc_hydro* hydro, int32_t index, int32_t nr, int32_t nq,
                                const double* radial_vertices, const double* polar_vertices,
                                const double* conserved, void* stream, double* dt, knc_error* err);

/* nsteps RK steps (RungeKuttaOrder::try_advance_async with the stage of
 * try_advance_rk, src/scheme.rs:16-173) with a fixed dt, without excision. */
int knc_advance_steps(knc_context* ctx, double dt, int32_t nsteps, knc_error* err);

/* scheme::advance (src/scheme.rs:212-240): dt from time_step once, then
 * `fold` RK steps, each preceded by the excision check of
 * add_remove_blocks when excision is set (knc_set_excision).  dt_used may
 * be NULL.  The time and iteration advance with every completed step. */
int knc_advance(knc_context* ctx, int32_t fold, double* dt_used, knc_error* err);

/* ---- moving excision surfaces ------------------------------------------ */
/* A new outer block, written by a knc_block_source_fn: the buffers must stay
 * valid until the callback is called again or the advance call returns. */
typedef struct knc_block_data {
    int32_t nr;
    int32_t pad;
    const double* radial_vertices;   /* nr+1 */
    const double* conserved;         /* (nr, nq, 4), volume-integrated */
    const double* scalar_mass;       /* (nr, nq) */
} knc_block_data;

/* Block birth (BlockState::from_model on mesh.subgrid(index) at `time`,
 * src/scheme.rs:197-205).  Returns 0, or nonzero to abort the advance call
 * with KNC_ERR_INVALID_ARGUMENT. */
typedef int32_t (*knc_block_source_fn)(void* user, int32_t index, double time, knc_block_data* out);

/* Excision surfaces of Mesh (src/mesh.rs:312-336): IES(t) = inner_radius +
 * max(t - delay, 0) * inner_speed, OES(t) likewise from outer_radius.  With
 * either speed positive (moving_excision_surfaces), before each RK step whose
 * iteration is a multiple of check_every (1: every step, as
 * src/scheme.rs:228-232) the innermost block is removed if its outer radius
 * is below IES(t), and block outer+1 is added from `source` if the outermost
 * block's outer radius is below OES(t) (src/scheme.rs:179-206).  The
 * innermost block is never removed when it is the only one.  source may be
 * NULL (no births). */
typedef struct knc_excision {
    double inner_radius;
    double outer_radius;
    double inner_speed;
    double outer_speed;
    double delay;
    int32_t check_every;
    int32_t pad;
    knc_block_source_fn source;
    void* user;
} knc_excision;

/* Sets (or with NULL clears) the excision surfaces of a context. */
int knc_set_excision(knc_context* ctx, const knc_excision* excision);
/* Blocks removed and added by excision since the context was created. */
int knc_excision_counts(knc_context* ctx, int64_t* removed, int64_t* added);

/* ---- initial models on the device (the files of src/models) ----------- */
/* The InitialModel implementations of the reference, evaluated on the GPU:
 * kinetic_bomb, jet_in_cloud, jet_in_star, halo_kilonova and wind_shock.
 * params holds the fields of the model's struct in declaration order:
 *   KINETIC_BOMB   external_medium_density, launch_radius, shell_thickness,
 *                  kinetic_energy, shell_mass (src/models/kinetic_bomb.rs:16-22)
 *   JET_IN_CLOUD   cloud_mass, engine_delay, engine_duration, engine_strength,
 *                  engine_theta, engine_u, envelop_m1, envelop_slowest_beta,
 *                  envelop_fastest_beta, envelop_psi (src/models/jet_in_cloud.rs:18-49)
 *   JET_IN_STAR    star_mass, engine_duration, engine_energy, engine_theta,
 *                  engine_u, envelope_radius, envelope_mass, volume_factor
 *                  (src/models/jet_in_star.rs:37-62)
 *   HALO_KILONOVA  altitude, launch_radius, shell_thickness, kinetic_energy,
 *                  shell_mass, radial_distance (src/models/halo_kilonova.rs:14-25)
 *   WIND_SHOCK     wind_mass_outflow_rate, wind_gamma_beta, wind_pressure,
 *                  post_shock_pressure, shock_location, post_shock_gamma_beta,
 *                  flare_outflow_rate, flare_gamma_beta, flare_time,
 *                  flare_duration (src/models/wind_shock.rs:14-59); table
 *                  holds table_rows rows (radius, gamma-beta, density,
 *                  specific enthalpy) of the initial data table, the first
 *                  column strictly increasing (LookupTable::from_rows,
 *                  src/lookup_table_v2.rs:35-48), or table_rows = 0. */
enum {
    KNC_MODEL_KINETIC_BOMB = 1,
    KNC_MODEL_JET_IN_CLOUD = 2,
    KNC_MODEL_JET_IN_STAR = 3,
    KNC_MODEL_HALO_KILONOVA = 4,
    KNC_MODEL_WIND_SHOCK = 5
};
#define KNC_MODEL_NPARAM 12
typedef struct knc_model {
    int32_t kind;
    int32_t table_rows;
    double params[KNC_MODEL_NPARAM];
    const double* table;        /* copied in; may be NULL when table_rows = 0 */
} knc_model;

/* The reference's radial block mesh (Mesh, src/mesh.rs:84-113): block i
 * spans [R (1 + b)^i, R (1 + b)^(i+1)] with b = block_size * zone_dlogr,
 * zone_dlogr = 1/num_radial_zones, or pi/num_polar_zones when
 * num_radial_zones = 0 (None); block_size log-spaced radial zones; polar
 * extent [0, pi], or pi/2 -+ zone_dlogr when num_polar_zones = 1
 * (src/mesh.rs:341-375, 202-229). */
typedef struct knc_mesh {
    double reference_radius;
    int32_t num_radial_zones;
    int32_t num_polar_zones;
    int32_t block_size;
    int32_t pad;
} knc_mesh;

/* Vertices of block `index` of the mesh: block_size+1 radial and
 * num_polar_zones+1 polar values (SphericalPolarGrid::geometry,
 * src/mesh.rs:246-264).  Host only; no context needed. */
int knc_mesh_block_vertices(const knc_mesh* mesh, int32_t index, double* radial_vertices,
                            double* polar_vertices);

/* primitive_at / scalar_at at n points (r, theta) and `time` on the device:
 * primitive (n, 4) as (rho, u_r, u_theta, p), scalar (n). */
int knc_model_eval(knc_context* ctx, const knc_model* model, int64_t n, const double* r, const double* theta,
                   double time, double* primitive, double* scalar);

/* Inserts block `index` of `mesh` built on the device by
 * BlockState::from_model at `time` (src/state.rs:55-69). */
int knc_block_insert_model(knc_context* ctx, int32_t index, const knc_mesh* mesh, const knc_model* model,
                           double time);

/* Makes `side` a model boundary evaluated on the device: at every stage's
 * input time the ghost rows are BlockState::from_model on the two rows of
 * mesh block inner_index - 1 (side 0) or outer_index + 1 (side 1) next to
 * the live mesh, then try_to_primitive (src/scheme.rs:30-52).  No host call
 * per stage.  model = NULL returns the side to the rows of knc_set_boundary.
 * The mesh's polar grid must be the blocks'. */
int knc_set_boundary_model(knc_context* ctx, int32_t side, const knc_mesh* mesh, const knc_model* model);

/* Block births of excision (src/scheme.rs:197-205) from the model on the
 * device, for a context whose excision has no block source; NULL clears. */
int knc_set_birth_model(knc_context* ctx, const knc_mesh* mesh, const knc_model* model);

/* ---- checkpoints and products (src/io.rs, src/app.rs:174-186,
 *      src/products.rs) ---------------------------------------------------- */
/* The reference's output files, serde-serialised through ciborium 0.1.0
 * (src/io.rs:24-35): every number in its shortest lossless CBOR form,
 * structs as maps keyed by field name, enums externally tagged, ndarray
 * arrays as {"v": 1, "dim": [...], "data": [...]}, BlockIndex map keys as
 * [i, j], the Rational64 iteration as [numer, denom].  The library encodes
 * the state from HBM; the Configuration (src/app.rs:163-172) and the Tasks
 * (src/tasks.rs:10-47) are small and arrive already CBOR-encoded from the
 * caller, copied into the file verbatim. */
typedef struct knc_record_meta {
    const uint8_t* config;      /* CBOR of app::Configuration */
    uint64_t config_len;
    const uint8_t* tasks;       /* CBOR of tasks::Tasks (checkpoints only) */
    uint64_t tasks_len;
    const char* version;        /* the version string of App / Products */
} knc_record_meta;

/* chkpt.NNNN.cbor (side_effects, src/main.rs:61-68): App {state, tasks,
 * config, version} (src/app.rs:179-186) with state = AnyState {"Relativistic" | "Newtonian":
 * State {time, iteration, solution {[i, 0]: BlockState {conserved,
 * scalar_mass}}}} (src/state.rs:24-43); conserved elements are 4-tuples
 * (D, S_r, S_theta, tau).  The blocks stream from HBM through pinned
 * buffers while the previous block is encoded. */
int knc_write_checkpoint(knc_context* ctx, const char* path, const knc_record_meta* meta);

/* prods.NNNN.cbor (src/main.rs:50-59): Products {time, blocks {[i, 0]:
 * BlockProducts {radial_vertices, polar_vertices, primitive, scalar}},
 * config, version} (src/products.rs:14-37, 44-91): try_to_primitive of every zone
 * on the device, primitive elements in AnyPrimitive order (u_r, u_theta,
 * rho, p) (src/physics/mod.rs:36-75), scalar = scalar_mass / D.  A failed
 * c2p writes nothing and returns its KNC_HYDRO_* code with err filled. */
int knc_write_products(knc_context* ctx, const char* path, const knc_record_meta* meta, knc_error* err);

/* Reading a checkpoint (App::from_file on a .cbor, src/app.rs:344-350).
 * The handle parses and checks the whole file when opened; block data are
 * decoded on request.  Byte ranges and strings returned by
 * knc_checkpoint_get_info stay valid until knc_checkpoint_close. */
typedef struct knc_checkpoint knc_checkpoint;
typedef struct knc_checkpoint_info {
    int32_t system;             /* KNC_SRHD (Relativistic) or KNC_EULER (Newtonian) */
    int32_t num_blocks;
    double time;
    int64_t iteration_numer;
    int64_t iteration_denom;
    const uint8_t* config;      /* the CBOR item of the config field */
    uint64_t config_len;
    const uint8_t* tasks;       /* the CBOR item of the tasks field */
    uint64_t tasks_len;
    const char* version;
} knc_checkpoint_info;
int knc_checkpoint_open(const char* path, knc_checkpoint** out);
int knc_checkpoint_get_info(const knc_checkpoint* ck, knc_checkpoint_info* out);
/* Block k (0 <= k < num_blocks, increasing block index). */
int knc_checkpoint_block_shape(const knc_checkpoint* ck, int32_t k, int32_t* index, int32_t* nr, int32_t* nq);
/* conserved (nr, nq, 4) and scalar_mass (nr, nq), as knc_block_download. */
int knc_checkpoint_block_data(const knc_checkpoint* ck, int32_t k, double* conserved, double* scalar_mass);
/* Inserts every block of the checkpoint into an empty context, with the
 * vertices of `mesh` (knc_mesh_block_vertices; the blocks' shapes must be
 * block_size x num_polar_zones), and sets the context's time and iteration
 * (the iteration must be whole). */
int knc_checkpoint_load(const knc_checkpoint* ck, knc_context* ctx, const knc_mesh* mesh);
void knc_checkpoint_close(knc_checkpoint* ck);

/* Decoding helper for readers of these files (no counterpart in the
 * reference; host only, no context): the CBOR array at data[offset] of
 * exactly `count` items, each a number (arity 1) or an array of `arity`
 * numbers, any float width or integer, into out[count * arity]; *end
 * receives the offset just past the array.  KNC_ERR_FORMAT if the item is
 * not such an array. */
int knc_cbor_read_f64(const uint8_t* data, uint64_t len, uint64_t offset, uint64_t count, int32_t arity,
                      double* out, uint64_t* end);

/* ---- autotuner (no counterpart in the reference) ------------------------ */
/* Times the stage kernels of the context's current mesh for each supported
 * workgroup shape (polar columns per work item), radial rows per work item
 * and stage form (fused or two-pass), keeps the fastest for later advance
 * calls, and reports it.  The
 * state, time and statistics are unchanged.  `trials` timed RK steps per
 * candidate (0: 3).  Results are cached per mesh shape and configuration
 * within the process; force = 1 re-measures. */
typedef struct knc_tune_result {
    int32_t stage_threads;
    int32_t rows_per_item;
    int32_t stage_passes;
    int32_t candidates;     /* shapes timed (0 when taken from the cache) */
    int32_t from_cache;
    int32_t pad;
    double ms_per_stage;    /* of the chosen shape */
    double default_ms_per_stage;   /* of the library's default choice */
} knc_tune_result;
int knc_autotune(knc_context* ctx, int32_t trials, int32_t force, knc_tune_result* out);

int knc_get_c2p_stats(knc_context* ctx, knc_c2p_stats* out);
int knc_reset_c2p_stats(knc_context* ctx);
int knc_synchronize(knc_context* ctx);
/* Average device time of the last advance call's stages (ms), from HIP
 * events on the context's stream around all of its stage launches (the
 * guess and c2p passes before a stage included, and in batched runs the
 * CFL reduction that rides on the first stage's c2p pass); 0 if not
 * measured. */
int knc_last_stage_time(knc_context* ctx, double* ms_per_stage, int32_t* launches);
int knc_set_timing(knc_context* ctx, int32_t enable);

/* ---- per-zone Hydrodynamics methods over host arrays ------------------- */
/* try_to_primitive (src/traits.rs:77-82): code[k] is 0 or a KNC_HYDRO_*
 * code; KNC_HYDRO_NEGATIVE_PRESSURE marks the 1e-3 rho floor and returns
 * the floored primitive.  guess may be NULL (plain guess p = 0). */
int knc_hydro_try_to_primitive(const knc_hydro* h, int64_t n, const double* conserved, const double* guess,
                               double* primitive, int32_t* code, int32_t* iterations);
/* to_conserved (src/traits.rs:90-93) */
int knc_hydro_to_conserved(const knc_hydro* h, int64_t n, const double* primitive, double* conserved);
/* intercell_flux (src/traits.rs:118-124), scaled by light_speed */
int knc_hydro_intercell_flux(const knc_hydro* h, int64_t n, const double* pl, const double* pr, const double* sl,
                             const double* sr, int32_t direction, double* flux, double* scalar_flux);
/* geometrical_source_terms (src/traits.rs:126-130) at (r, theta) */
int knc_hydro_geometrical_source_terms(const knc_hydro* h, int64_t n, const double* primitive, const double* r,
                                       const double* theta, double* out);
/* max_signal_speed (src/traits.rs:95-98) */
int knc_hydro_max_signal_speed(const knc_hydro* h, int64_t n, const double* primitive, double* out);
/* plm_gradient_scalar / the per-component rule of plm_gradient_primitive
 * (src/traits.rs:66-75) */
int knc_hydro_plm_gradient(double plm_theta, int64_t n, const double* a, const double* b, const double* c,
                           double* out);

/* ---- learned c2p guess (no counterpart in the reference) --------------- */
/* A 4 -> 64 -> 64 -> 1 tanh MLP mapping log2-scaled features of a conserved
 * state (D, |S|/D, thermal part of tau/D, gamma) to log2(p/D), used as the
 * initial pressure of the SRHD Newton recovery when a context's c2p_guess is
 * KNC_GUESS_LEARNED; Newton always refines it and restarts from the plain
 * guess if it does not converge.  Parameters: 4545 FP32 values
 * W1[64][4] | b1[64] | W2[64][64] | b2[64] | W3[64] | b3; norm: 10 FP32
 * values shift[4] | scale[4] | tshift | tscale (x = (f - shift) * scale,
 * target = (log2(p/D) - tshift) * tscale). */
typedef struct knc_mlp knc_mlp;
const char* knc_mlp_last_error_message(void);
int knc_mlp_num_params(void);
int knc_mlp_create(int32_t device, knc_mlp** out);
void knc_mlp_destroy(knc_mlp* mlp);
int knc_mlp_init(knc_mlp* mlp, uint64_t seed);
int knc_mlp_set_params(knc_mlp* mlp, const float* params, const float* norm);
int knc_mlp_get_params(knc_mlp* mlp, float* params, float* norm);
/* Mean squared error and its gradient over n normalised samples
 * (x: n*4, t: n), computed by the training kernels. */
int knc_mlp_loss_grad(knc_mlp* mlp, int64_t n, const float* x, const float* t, float* loss, float* grad);
/* Adam training on the GPU; epoch_loss (may be NULL) gets one value per epoch. */
int knc_mlp_train(knc_mlp* mlp, int64_t n, const float* x, const float* t, int32_t epochs, int32_t batch, float lr,
                  uint64_t seed, float* epoch_loss);
/* Raw (unnormalised) features of per-volume conserved states, n*4. */
int knc_mlp_features(int64_t n, const double* conserved, const double* gamma, float* features);
/* BF16-MFMA inference: normalised output y (may be NULL) and the pressure
 * guess (may be NULL). */
int knc_mlp_predict(knc_mlp* mlp, int64_t n, const double* conserved, const double* gamma, float* y, double* guess);
/* Attaches a trained network to a context (copied; the network may be
 * destroyed afterwards).  Required before advancing a context created with
 * c2p_guess = KNC_GUESS_LEARNED. */
int knc_context_set_mlp(knc_context* ctx, const knc_mlp* mlp);

/* FP64 vector FMA rate of the current device: 256-thread workgroups,
 * waves_per_simd (1..8) of them per compute unit, every lane running
 * `chains` (1, 2, 4 or 8) independent chains of `iters` dependent FMAs.
 * Writes FMAs per second and the timed launch's milliseconds (ms may be
 * NULL).  A measurement utility with no counterpart in the reference: it
 * confirms the FP64 issue peak that the stage kernel is measured against and
 * the rate a dependent chain leaves at a given occupancy. */
int knc_fp64_fma_rate(int chains, int waves_per_simd, int iters, double* fma_per_s, float* ms);

/* FP64 arithmetic probe of the device (a/b, sqrt|a|, a*b+c with two
 * roundings, and a/|b| as the stage divides U by dV: from the correctly
 * rounded 1/|b| with one correction), used to check that device arithmetic
 * rounds like the host's. */
int knc_fp64_probe(int64_t n, const double* a, const double* b, const double* c, double* div, double* sqrt_abs_a,
                   double* mul_add, double* quot_abs_b);

/* The ghost rows of one side's device boundary model (knc_set_boundary_model)
 * at `time` for the context's current block set, each written as 5 planes of
 * 2 nq doubles (rho, u_r, u_theta, p, s; the row next to the mesh edge
 * second on the inner side, first on the outer), twice: as the stage builds
 * them (rows) and with one thread per zone (rows_one_thread).  A check with
 * no counterpart in the reference: the library builds halo_kilonova's rows
 * with a workgroup per zone, and the two must agree bit for bit.  Leaves the
 * context's error log empty. */
int knc_model_rows_check(knc_context* c, int32_t side, double time, double* rows, double* rows_one_thread);

#ifdef __cplusplus
}
#endif

#endif /* KNC_GPU_H */
