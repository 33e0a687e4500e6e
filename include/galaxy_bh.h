/*
 * galaxy_bh.h — C ABI of the MI355X (gfx950) Barnes-Hut evolution library.
 *
 * Drop-in boundary for GALAXY's BH #2D resident Barnes-Hut step.  Every entry
 * point below names the reference interface it replaces (file:line in the
 * GALAXY repository).  Conventions shared by all entry points:
 *   - return 0 on success, nonzero on failure; the failure text is available
 *     from gbh_last_error() on the calling thread (thread-local), matching the
 *     human-readable Box<dyn Error> strings of runtime/src/config.rs:7;
 *   - calls are synchronous: they return after the GPU work they issued has
 *     completed, like the reference's submit + device.poll(Wait);
 *   - one host thread drives a device handle at a time;
 *   - only plain pointers and sizes cross the boundary.
 */
#ifndef GALAXY_BH_H
#define GALAXY_BH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GBH_NO_CHILD 0xffffffffu
/* nbody_parallel.rs:20 (MAX_PARALLEL_TREE_BODIES) */
#define GBH_MAX_PARALLEL_TREE_BODIES 65536u
/* Upper bound of the beyond-cap mode (gbh_tree_create_large): 2^23 bodies, so
 * that every subtree size of the traversal's node stream (at most the cell
 * capacity bound plus one node per body, < 2^26) fits its 26-bit field; about
 * 10 GB of device memory per simulation at the bound. */
#define GBH_MAX_LARGE_TREE_BODIES (1u << 23)

/* ---------------------------------------------------------------------------
 * Frozen f32/u32 transfer records, byte-identical to runtime/src/nbody_gpu.rs:22-83
 * (sizes pinned at nbody_gpu.rs:1441-1451, 144-byte fixture at :1453-1489).
 * ------------------------------------------------------------------------- */
typedef struct { uint32_t info[4]; float physics[4]; } GbhBhSettings;      /* 32 B: {N, cells, root, 0} {theta, eps, G, 0} */
typedef struct { float position_mass[4]; uint32_t meta[4]; } GbhGpuBody;  /* 32 B: {x, y, m, 0} {sorted position, 0, 0, 0} */
typedef struct { uint32_t data[4]; } GbhGpuEntry;                         /* 16 B: {Morton code, body index, 0, 0} */
typedef struct {
    float center_half_mass[4]; /* cx, cy, half, mass */
    float com[4];              /* com_x, com_y, 0, 0 */
    uint32_t range_depth[4];   /* start, end, depth, valid */
    uint32_t children[4];      /* child ids or GBH_NO_CHILD */
} GbhGpuCell;                                                             /* 64 B */
typedef struct { float value[4]; } GbhGpuAccel;                           /* 16 B: {ax, ay, 0, 0} */
typedef struct { uint32_t info[4]; float motion[4]; } GbhEvolveSettings;  /* 32 B: {N, 0, 0, 0} {dt, 0, 0, 0} */
typedef struct { float position_mass[4]; float velocity[4]; } GbhEvolveState; /* 32 B */
typedef struct { uint32_t info[4]; float physics[4]; } GbhTreeSettings;   /* 32 B */
typedef struct { uint32_t data[4]; float bounds[4]; } GbhTreeMeta;        /* 32 B: data 0; {cx, cy, half, 0} */

/* galaxy_nbody::Config (nbody/src/lib.rs:27-46); validated as Config::validate (:48-67). */
typedef struct {
    double theta;
    double softening;
    double g;
    uint32_t bucket;
    uint32_t max_depth;
} gbh_config;

/* StageTiming (nbody_gpu.rs:125-130) */
typedef struct {
    double transfer_seconds;
    double dispatch_seconds;
    double readback_seconds;
} gbh_stage_timing;

/* ParallelTreeBuildTiming (nbody_parallel.rs:45-53): host wall time per stage
 * family; stages that the GPU library fuses report their time in the stage that
 * launches them. */
typedef struct {
    double bounds_seconds;
    double morton_seconds;
    double radix_seconds;
    double positions_seconds;
    double topology_seconds;
    double aggregate_seconds;
} gbh_build_timing;

typedef struct gbh_device gbh_device;
typedef struct gbh_state gbh_state;
typedef struct gbh_runtime gbh_runtime;
typedef struct gbh_tree gbh_tree;
typedef struct gbh_batch gbh_batch;
typedef struct gbh_server gbh_server;

/* Thread-local text of the last failure (runtime/src/config.rs:7 error strings). */
const char* gbh_last_error(void);

/* ---------------------------------------------------------------------------
 * Device (NbodyGpu::new, nbody_gpu.rs:331-621; select_adapter :278-296)
 * selector: NULL, a decimal device index, or a case-insensitive name substring.
 * allow_software is accepted for interface parity; ROCm exposes no software
 * adapter, so "software" is always false.
 * ------------------------------------------------------------------------- */
int gbh_device_open(const char* selector, int allow_software, gbh_device** out);
void gbh_device_close(gbh_device* dev);
/* describe()/info() (nbody_gpu.rs:298-310, 782-784): JSON object with keys
 * index, adapter_count, name, backend, device_type, driver, driver_info, software.
 * Writes at most cap bytes (NUL-terminated) and the full length to *len. */
int gbh_device_info_json(gbh_device* dev, char* buf, size_t cap, size_t* len);

/* ---------------------------------------------------------------------------
 * Persistent evolving state (EvolvingState, nbody_gpu.rs:132-138)
 * ------------------------------------------------------------------------- */
/* create_evolving_state (nbody_gpu.rs:786-852): n in 2..=u32::MAX, finite values,
 * mass > 0 ("body {i} contains invalid evolving state"). */
int gbh_state_create(gbh_device* dev, const GbhEvolveState* states, uint32_t n,
                     gbh_state** out, gbh_stage_timing* timing);
void gbh_state_destroy(gbh_state* st);
uint32_t gbh_state_count(const gbh_state* st);
/* evolving_buffer_bytes (nbody_gpu.rs:1116-1118) */
uint64_t gbh_state_bytes(const gbh_state* st);
/* read_evolving_state (nbody_gpu.rs:1019-1069): rejects non-finite or mass <= 0. */
int gbh_state_read(gbh_state* st, GbhEvolveState* out, gbh_stage_timing* timing);
/* read_evolving_accelerations (nbody_gpu.rs:1071-1114): rejects non-finite. */
int gbh_accel_read(gbh_state* st, GbhGpuAccel* out, gbh_stage_timing* timing);
/* kick_drift / final_kick (nbody_gpu.rs:1011-1017, integrate_stage :940-1009):
 * dt finite and > 0, applied as f32. */
int gbh_kick_drift(gbh_state* st, double dt, gbh_stage_timing* timing);
int gbh_final_kick(gbh_state* st, double dt, gbh_stage_timing* timing);

/* ---------------------------------------------------------------------------
 * BH #2D parallel tree (ParallelTreeRuntime, nbody_parallel.rs:93-993)
 * ------------------------------------------------------------------------- */
/* ParallelTreeRuntime::new (nbody_parallel.rs:156-307) */
int gbh_runtime_create(gbh_device* dev, gbh_runtime** out);
void gbh_runtime_destroy(gbh_runtime* rt);
/* create_tree (nbody_parallel.rs:361-493): 2..=65,536 resident bodies. */
int gbh_tree_create(gbh_runtime* rt, gbh_state* st, gbh_tree** out);
/* Beyond-cap mode (no reference counterpart): 2..=GBH_MAX_LARGE_TREE_BODIES,
 * compact cell layout; same kernels and results contract. */
int gbh_tree_create_large(gbh_runtime* rt, gbh_state* st, gbh_tree** out);
void gbh_tree_destroy(gbh_tree* tree);
/* rebuild (nbody_parallel.rs:555-826).  The stages run as one HIP graph (captured
 * at the tree's first rebuild), so the call's wall time is reported in
 * timing->bounds_seconds and the other stage slots are 0. */
int gbh_tree_rebuild(gbh_runtime* rt, gbh_state* st, gbh_tree* tree, const gbh_config* cfg,
                     gbh_build_timing* timing);
/* force (nbody_parallel.rs:828-897): writes the state's acceleration buffer. */
int gbh_force(gbh_runtime* rt, gbh_state* st, gbh_tree* tree, const gbh_config* cfg,
              gbh_stage_timing* timing);
/* Force with per-target interaction evidence (no reference counterpart; the
 * parity check of the traversal): stats[3*i] = {visited cells, accepted cells,
 * direct terms} and hash[i] = FNV-mix64 over body i's ordered interaction list
 * (direct term: (1<<60)|source body; accepted cell: (2<<60)|(depth<<32)|start).
 * Either pointer may be NULL. */
int gbh_force_evidence(gbh_runtime* rt, gbh_state* st, gbh_tree* tree, const gbh_config* cfg,
                       uint32_t* stats, uint64_t* hash);
/* No reference counterpart.  For callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc): gbh_force_evidence with the same arguments and results, but the
 * per-target evidence is staged in a fixed device array of the library (one per
 * GPU, 20 bytes per target, held until the process ends) instead of device
 * memory allocated for the call.  The call allocates nothing.  Trees of more
 * than GBH_NOALLOC_MAX_BODIES bodies are rejected with an error.  Every call on
 * one GPU shares that array: calls that overlap on one GPU (several host
 * threads) race, and must be serialized by the caller. */
#define GBH_NOALLOC_MAX_BODIES 65536u
int gbh_force_evidence_noalloc(gbh_runtime* rt, gbh_state* st, gbh_tree* tree,
                               const gbh_config* cfg, uint32_t* stats, uint64_t* hash);
/* Cell statistics of the last rebuild: active cells, leaves, max depth. */
int gbh_tree_counts(gbh_tree* tree, uint32_t* active, uint32_t* leaves, uint32_t* max_depth);
/* read_evidence (nbody_parallel.rs:929-988): sparse level-order layout, cells in
 * slot order (slot = depth*N + start), children as slot ids, valid = 1,
 * meta.data = 0.  cells must hold cell_capacity records (>= active). */
int gbh_tree_read_evidence(gbh_tree* tree, GbhTreeMeta* meta, GbhGpuEntry* entries,
                           GbhGpuCell* cells, uint32_t cell_capacity, uint32_t* active,
                           uint32_t* leaves, uint32_t* max_depth);
/* BH #2A-layout readback (no reference counterpart): the tree in flat.rs preorder,
 * packed like pack_flat_tree (nbody_gpu.rs:175-250, range_depth[3] = 0), and the
 * sorted position of every body (entry_position_by_body, flat.rs:236-239). */
int gbh_tree_read_flat(gbh_tree* tree, GbhGpuEntry* entries, GbhGpuCell* cells,
                       uint32_t cell_capacity, uint32_t* ncells, uint32_t* position_by_body);
/* GPU time (ms, HIP events on the library's stream) of the traversal kernel of
 * the last gbh_force on this tree (no reference counterpart; profiling). */
double gbh_tree_last_force_ms(gbh_tree* tree);
/* buffer_bytes (nbody_parallel.rs:990-992): bytes this library actually holds. */
uint64_t gbh_tree_buffer_bytes(gbh_tree* tree);
/* FNV-mix64 of the receipt evidence (checksum_word, galaxy-bh-gpu-tree-parallel.rs:
 * 145-149) folded over `count` 32-bit words from `seed`; parallel_tree_checksum
 * (:150-187) is this over the meta, entry and active-cell words in order. */
int gbh_fnv_mix64(uint64_t seed, const uint32_t* words, uint64_t count, uint64_t* out);
/* The reference's sparse-layout size for n bodies (nbody_parallel.rs:485-491),
 * kept for receipt-field compatibility. */
uint64_t gbh_sparse_layout_bytes(uint32_t n);

/* ---------------------------------------------------------------------------
 * BH #2B1 / #2B2 host-built tree ingest (nbody_gpu.rs:640-780, 854-938)
 * ------------------------------------------------------------------------- */
/* traverse (nbody_gpu.rs:640-780): packed records in, accelerations out. */
int gbh_traverse_packed(gbh_device* dev, const GbhBhSettings* settings, const GbhGpuBody* bodies,
                        const GbhGpuEntry* entries, const GbhGpuCell* cells, GbhGpuAccel* out,
                        gbh_stage_timing* timing);
/* force_into (nbody_gpu.rs:854-938): packed records in, state acceleration out. */
int gbh_force_packed(gbh_state* st, const GbhBhSettings* settings, const GbhGpuBody* bodies,
                     const GbhGpuEntry* entries, const GbhGpuCell* cells,
                     gbh_stage_timing* timing);

/* Host BH #2A tree packed for the GPU: build_flat_tree (nbody/src/flat.rs:127-268)
 * of the f64 bodies (x, y, mass), then pack_flat_tree (nbody_gpu.rs:175-250) into
 * settings, bodies[n], entries[n] and cells (preorder, root 0).  *cell_count
 * receives the tree's cell count; the call fails if cap is smaller (at most
 * 16 n + 1 cells).  The packed records feed gbh_force_packed (BH #2B2). */
int gbh_host_flat_tree(uint32_t n, const double* x, const double* y, const double* mass,
                       const gbh_config* cfg, GbhBhSettings* settings, GbhGpuBody* bodies,
                       GbhGpuEntry* entries, GbhGpuCell* cells, uint32_t cap,
                       uint32_t* cell_count);

/* ---------------------------------------------------------------------------
 * BH #2C correctness-first GPU tree (NbodyGpu::create_gpu_tree, rebuild_gpu_tree,
 * force_from_gpu_tree, read_gpu_tree_evidence, gpu_tree_buffer_bytes;
 * nbody_gpu.rs:1145-1431; kernels nbody_tree.wgsl:77-370): f32 root bounds,
 * keys and (code, index) order, cells in level order, f32 aggregates; the force
 * follows bh_traverse (nbody.wgsl:43-123) over these cells.  Used as a GPU
 * oracle of the BH #2D path (galaxy-bh-gpu-tree-parallel.rs:225-246, 422-474).
 * ------------------------------------------------------------------------- */
typedef struct gbh_gpu_tree gbh_gpu_tree;
/* create_gpu_tree (nbody_gpu.rs:1146): MAX_GPU_TREE_BODIES */
#define GBH_MAX_GPU_TREE_BODIES 4096u
/* create_gpu_tree (nbody_gpu.rs:1145-1212): 2..=4096 resident bodies, cell
 * capacity 16 n + 1. */
int gbh_gpu_tree_create(gbh_device* dev, gbh_state* st, gbh_gpu_tree** out);
void gbh_gpu_tree_destroy(gbh_gpu_tree* tree);
/* rebuild_gpu_tree (nbody_gpu.rs:1214-1301).  The six stages run as one
 * launch; their wall time is reported in timing->bounds_seconds (the slot
 * order of TreeBuildTiming, nbody_gpu.rs:85-93, with sort in radix_seconds). */
int gbh_gpu_tree_rebuild(gbh_state* st, gbh_gpu_tree* tree, const gbh_config* cfg,
                         gbh_build_timing* timing);
/* force_from_gpu_tree (nbody_gpu.rs:1303-1366): writes the state's acceleration. */
int gbh_gpu_tree_force(gbh_state* st, gbh_gpu_tree* tree, const gbh_config* cfg,
                       gbh_stage_timing* timing);
/* read_gpu_tree_evidence (nbody_gpu.rs:1392-1428): meta, the n sorted entries and
 * the meta.data[0] cells in level order (cells must hold cap >= that many). */
int gbh_gpu_tree_read_evidence(gbh_gpu_tree* tree, GbhTreeMeta* meta, GbhGpuEntry* entries,
                               GbhGpuCell* cells, uint32_t cap, uint32_t* cell_count);
/* gpu_tree_buffer_bytes (nbody_gpu.rs:1430-1432): bytes this library holds. */
uint64_t gbh_gpu_tree_buffer_bytes(gbh_gpu_tree* tree);

/* ---------------------------------------------------------------------------
 * Batched ensembles (no reference counterpart): many independent simulations of
 * uneven N advanced together, one launch per phase, whole step captured in a
 * HIP graph.  Each simulation's results are bit-identical to the same
 * simulation run alone through the drop-in entry points.
 * ------------------------------------------------------------------------- */
typedef struct {
    gbh_config config;
    double dt;
} gbh_sim_params;

/* Launch-shape choices; none of them changes any result.  A zero field keeps
 * its default (2048, 256, 32, one group, 512). */
typedef struct {
    uint32_t sort_tile;      /* keys per sort workgroup: 1024 or 2048 */
    uint32_t traverse_block; /* traversal workgroup size: 64, 128, 256 */
    uint32_t traverse_waves_per_cu; /* persistent traversal wavefronts per CU, 1..32 */
    uint32_t use_graph;      /* capture each step in a HIP graph */
    uint32_t streams;        /* 1 or 2 groups of simulations (a server: of in-flight
                                requests), each stepped on a stream of its own so their
                                kernels overlap (0 = 1) */
    uint32_t topo_tile;      /* sorted positions per topology workgroup in a set of
                                several simulations: 512 or 1024 (a set of one
                                simulation always uses 1024) */
} gbh_tuning;

/* counts[nsims]; states = all simulations' records concatenated in order. */
int gbh_batch_create(gbh_device* dev, uint32_t nsims, const uint32_t* counts,
                     const GbhEvolveState* states, const gbh_sim_params* params,
                     const gbh_tuning* tuning, gbh_batch** out);
void gbh_batch_destroy(gbh_batch* b);
/* Advance every simulation by `steps` KDK steps (the first call also performs
 * the initial build + force, like galaxy-bh-gpu-tree-parallel.rs:338-339). */
int gbh_batch_step(gbh_batch* b, uint32_t steps);
int gbh_batch_read(gbh_batch* b, uint32_t sim, GbhEvolveState* out, GbhGpuAccel* accel);
int gbh_batch_force_evidence(gbh_batch* b, uint32_t sim, uint32_t* stats, uint64_t* hash);
uint64_t gbh_batch_device_bytes(gbh_batch* b);
/* Cells and traversal nodes (internal cells + bodies) of a simulation's current
 * tree (the last force state). */
int gbh_batch_tree_counts(gbh_batch* b, uint32_t sim, uint32_t* cells, uint32_t* nodes);
/* Profiling: runs ONE MORE KDK step of every simulation, un-captured and with
 * each phase's launches in sequence, and writes its kernel time per phase (ms,
 * HIP events): kick/drift + bounds, keys, sort, topology, top aggregates,
 * traversal.  The ensemble advances by that step, like a gbh_batch_step(b, 1). */
int gbh_batch_profile_step(gbh_batch* b, double* ms, uint32_t cap, uint32_t* count);

/* ---------------------------------------------------------------------------
 * Serving (no reference counterpart): continuous batching of "advance this
 * galaxy k steps" requests with HBM-resident state and HBM-budget admission.
 * ------------------------------------------------------------------------- */
typedef struct {
    uint64_t hbm_budget_bytes;  /* admission budget for resident request state */
    uint32_t steps_per_round;   /* steps every in-flight request advances per round */
    uint32_t max_inflight;      /* cap on concurrently resident requests (0 = none) */
    gbh_tuning tuning;
} gbh_server_options;

typedef struct {
    uint64_t id;
    double submit_time;   /* seconds, server clock */
    double admit_time;
    double finish_time;
    uint32_t n;
    uint32_t steps;
    int32_t status;       /* 0 = ok, nonzero = failed (gbh_server_take_result gives the text) */
} gbh_completion;

/* Device bytes one resident simulation of n bodies holds (state + tree buffers);
 * the unit of HBM-budget admission. */
uint64_t gbh_sim_device_bytes(uint32_t n);
/* The server keeps the device arenas of retired requests for reuse by later ones;
 * pooled and resident arenas together stay within hbm_budget_bytes.  The pool is
 * released when an allocation fails and at gbh_server_destroy (which first stops
 * the worker thread, if one runs). */
int gbh_server_create(gbh_device* dev, const gbh_server_options* opt, gbh_server** out);
void gbh_server_destroy(gbh_server* s);
/* Server clock in seconds (monotonic, starts at create). */
double gbh_server_now(gbh_server* s);
/* Queue a request; the initial state is copied.  Refused immediately (nonzero)
 * if the request alone exceeds the HBM budget.  Safe to call from any thread,
 * also while the worker thread runs. */
int gbh_server_submit(gbh_server* s, const GbhEvolveState* states, uint32_t n,
                      const gbh_sim_params* params, uint32_t steps, uint64_t* id);
/* One serving round on the calling thread: admit queued requests while budget
 * and slots allow, advance every in-flight request, retire finished ones.
 * Returns the round's completions (*ndone of them; at most cap are written).
 * Refused while the worker thread runs. */
int gbh_server_round(gbh_server* s, gbh_completion* done, uint32_t cap, uint32_t* ndone);
/* Continuous serving: a worker thread of the library runs rounds whenever
 * requests are queued or in flight and sleeps otherwise, so submissions never
 * wait for the caller to drive a round.  gbh_server_stop lets the current round
 * finish, joins the thread and reports a failure of the worker, if any (a device
 * error fails the requests concerned; a failed HIP call ends the worker). */
int gbh_server_start(gbh_server* s);
int gbh_server_stop(gbh_server* s);
/* Completions the worker thread produced since the last poll (at most cap; the
 * rest stay for the next poll).  With none ready, waits up to wait_seconds for
 * one while the worker runs. */
int gbh_server_poll(gbh_server* s, gbh_completion* out, uint32_t cap, uint32_t* ndone,
                    double wait_seconds);
/* Final state of a completed request (then released). */
int gbh_server_take_result(gbh_server* s, uint64_t id, GbhEvolveState* out, GbhGpuAccel* accel);
/* queued requests; in-flight requests and their resident bytes as of the last round */
int gbh_server_stats(gbh_server* s, uint32_t* queued, uint32_t* inflight, uint64_t* resident);

/* FP64 diagnostics over a resident state (no reference counterpart; the
 * measurement side of the driver's oracle checks).  gbh_diag_direct: f64 direct
 * accelerations of the listed bodies, as direct_acceleration_for
 * (nbody/src/lib.rs:315-329) on the state widened to f64; probe_error
 * (lib.rs:370-395) is computed from them by the caller.  gbh_diag_energy:
 * out = {kinetic + Plummer-softened pair potential, Lz, kinetic}, f64. */
int gbh_diag_direct(gbh_state* st, const gbh_config* cfg, uint32_t count, const uint32_t* index,
                    double* ax, double* ay);
int gbh_diag_energy(gbh_state* st, double softening, double g, double out[3]);
/* Exhaustive device check (no reference counterpart) that the traversal's f32
 * square root and reciprocals are correctly rounded: mismatches[0] = sqrt and
 * [1] = reciprocal over every non-negative f32, [2] = the reciprocal of the
 * guard-free interaction over every f32 in [2^-126, 2^126); first_bad = smallest
 * failing pattern of each. */
int gbh_check_rounding(gbh_device* dev, uint64_t mismatches[3], uint32_t first_bad[3]);
/* Device check (no reference counterpart) of the per-depth opening thresholds:
 * out[i] = the smallest positive f32 d2 with fl(2 half[i] / fl(sqrt(d2))) <
 * theta[i] as the builds compute it, NaN when none is (theta <= 0), so that
 * d2 >= out[i] <=> the criterion of nbody.wgsl:105-108 holds. */
int gbh_check_thresholds(gbh_device* dev, const float* half, const float* theta, uint32_t count, float* out);

/* Device-to-device copy of `bytes` bytes by a plain streaming kernel
 * (k_copy_calibration); GPU time in *ms.  Used to calibrate hardware traffic
 * counters (no reference counterpart). */
int gbh_copy_calibration(gbh_device* dev, uint64_t bytes, double* ms);

/* ---------------------------------------------------------------------------
 * Seeded synthetic initial conditions (host side; no GPU needed).
 * The f32 records are the canonical state; generation is f64 then one rounding.
 * ------------------------------------------------------------------------- */
/* Exponential disc Sigma ~ exp(-R/scale) truncated at `truncation`*scale, uniform
 * angles, circular speed from the enclosed mass with softened radius plus
 * `dispersion` x v_c Gaussian components; equal masses total_mass/n. */
int gbh_ic_rotating_disc(uint32_t n, uint64_t seed, double total_mass, double scale,
                         double truncation, double softening, double g, double dispersion,
                         GbhEvolveState* out);
/* Uniform unit-radius disc (R = radius*sqrt(u)), zero velocity, equal masses. */
int gbh_ic_uniform_collapse(uint32_t n, uint64_t seed, double total_mass, double radius,
                            GbhEvolveState* out);
/* Two exponential discs on a parabolic approach (COM frame). */
int gbh_ic_binary_disc(uint32_t n1, uint32_t n2, uint64_t seed1, uint64_t seed2, double m1,
                       double m2, double separation, double pericentre, double softening,
                       double g, double dispersion, GbhEvolveState* out);
/* make_disc / make_collision presets (nbody/src/lib.rs:454-528), kpc/Msun/Myr;
 * f64 bodies {x, y, vx, vy, mass} are written to `bodies` (5 doubles each). */
int gbh_ic_reference_disc(uint32_t n, uint64_t seed, double total_mass, double scale_kpc,
                          double* bodies);
int gbh_ic_reference_collision(uint32_t n, uint64_t seed, double* bodies);

#ifdef __cplusplus
}
#endif
#endif
