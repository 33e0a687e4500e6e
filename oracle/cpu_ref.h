/*
 * cpu_ref — CPU restatement of the GALAXY Barnes-Hut reference path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the oracle that the HIP path is
 * checked against.  Only tests/, __graft_entry__.smoke() and the cpu_baseline
 * leg of bench.py may load it.  Nothing in the galaxy_amd package links,
 * loads or calls it.
 *
 * What it restates (reference file:line):
 *   - Config / validation / root bounds / point-mass law / direct sums /
 *     probe error / KDK / BH #1 recursive tree .......... nbody/src/lib.rs
 *   - BH #2A Morton flat tree (f64) ...................... nbody/src/flat.rs
 *   - BH #2B1 packing into the f32/u32 transfer ABI ...... runtime/src/nbody_gpu.rs:175-250
 *   - BH #2B1/2D f32 flat traversal ...................... runtime/src/nbody.wgsl:35-123
 *   - BH #2B2 f32 kick/drift and final kick .............. runtime/src/nbody_evolve.wgsl:19-46
 *   - BH #2C correctness-first GPU tree (f32) ........... runtime/src/nbody_tree.wgsl:77-370
 *   - BH #2D sparse level-order evidence + checksum ...... runtime/src/nbody_parallel.rs:929-988,
 *                                                         runtime/src/bin/galaxy-bh-gpu-tree-parallel.rs:150-203
 *   - error metrics / cpu_flat_evolve .................... galaxy-bh-gpu-tree-parallel.rs:62-148
 *
 * Arithmetic discipline: compiled without FMA contraction; every f32 and f64
 * operation is an IEEE-754 round-to-nearest operation in source order.  Where
 * WGSL leaves rounding open (inverseSqrt), this oracle defines it as the two
 * correctly rounded operations 1.0f / sqrtf(r2).
 */
#ifndef GALAXY_CPU_REF_H
#define GALAXY_CPU_REF_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CREF_NO_CHILD 0xffffffffu
#define CREF_MORTON_TREE_DEPTH 16u

/* galaxy_nbody::Config (nbody/src/lib.rs:27-46) */
typedef struct {
    double theta;
    double softening;
    double g;
    uint32_t bucket;
    uint32_t max_depth;
} cref_config;

/* Frozen transfer records (runtime/src/nbody_gpu.rs:22-83). */
typedef struct { uint32_t info[4]; float physics[4]; } cref_bh_settings;          /* 32 B */
typedef struct { float position_mass[4]; uint32_t meta[4]; } cref_gpu_body;      /* 32 B */
typedef struct { uint32_t data[4]; } cref_gpu_entry;                             /* 16 B */
typedef struct {
    float center_half_mass[4];
    float com[4];
    uint32_t range_depth[4];
    uint32_t children[4];
} cref_gpu_cell;                                                                 /* 64 B */
typedef struct { float value[4]; } cref_gpu_accel;                               /* 16 B */
typedef struct { float position_mass[4]; float velocity[4]; } cref_evolve_state; /* 32 B */
typedef struct { uint32_t data[4]; float bounds[4]; } cref_tree_meta;            /* 32 B */

typedef struct cref_tree cref_tree;

const char* cref_last_error(void);
int cref_config_validate(const cref_config* cfg);

/* ---- BH #2A flat tree (f64) ------------------------------------------- */
uint32_t cref_morton2(uint32_t x, uint32_t y);                       /* flat.rs:77-80 */
uint32_t cref_quantize_axis(double value, double center, double half); /* flat.rs:82-87 */
cref_tree* cref_build_flat_tree(uint32_t n, const double* x, const double* y,
                                const double* m, const cref_config* cfg);
void cref_tree_free(cref_tree* t);
/* counts[0]=cells, [1]=node_count, [2]=leaf_count, [3]=max_depth */
void cref_tree_counts(const cref_tree* t, uint32_t counts[4]);
/* root[0..2] = f64 (cx, cy, half) from root_bounds */
void cref_tree_root(const cref_tree* t, double root[3]);
void cref_tree_entries(const cref_tree* t, uint32_t* code, uint32_t* body_index,
                       uint32_t* position_by_body);
/* geom: 6 doubles per cell (cx, cy, half, mass, com_x, com_y);
 * topo: 7 u32 per cell (start, end, depth, child0..child3). */
void cref_tree_cells(const cref_tree* t, double* geom, uint32_t* topo);
uint64_t cref_flat_topology_checksum(const cref_tree* t);

/* pack_flat_tree (nbody_gpu.rs:175-250): range_depth[3] = 0. Returns 0 on success. */
int cref_pack_flat_tree(const cref_tree* t, uint32_t n, const double* x, const double* y,
                        const double* m, const cref_config* cfg, cref_bh_settings* settings,
                        cref_gpu_body* bodies, cref_gpu_entry* entries, cref_gpu_cell* cells);

/* BH #2D sparse level-order evidence (slot = depth*N + start), cells ordered by slot,
 * children as slot ids, valid = 1, meta bounds = f32 of the tree's root bounds. */
void cref_sparse_evidence(const cref_tree* t, const cref_config* cfg, cref_tree_meta* meta,
                          cref_gpu_entry* entries, cref_gpu_cell* cells);
/* parallel_tree_checksum (galaxy-bh-gpu-tree-parallel.rs:156-187) */
uint64_t cref_parallel_tree_checksum(const cref_tree_meta* meta, uint32_t n,
                                     const cref_gpu_entry* entries, uint32_t ncells,
                                     const cref_gpu_cell* cells);
/* validate_parallel_order (galaxy-bh-gpu-tree-parallel.rs:189-203): 0 if strict */
int cref_validate_parallel_order(uint32_t n, const cref_gpu_entry* entries);

/* ---- traversal ---------------------------------------------------------- */
/* f32 flat traversal of packed records (nbody.wgsl:35-123).  Optional per-target
 * evidence: stats[3*i] = {visited, accepted cells, direct terms}; hash[i] =
 * FNV-mix64 over the ordered interaction list (direct: (1<<60)|body,
 * accepted cell: (2<<60)|(depth<<32)|start). */
int cref_traverse_f32(const cref_bh_settings* settings, const cref_gpu_body* bodies,
                      const cref_gpu_entry* entries, const cref_gpu_cell* cells,
                      cref_gpu_accel* out, uint32_t* stats, uint64_t* hash);
/* Threads the f32 traversal above uses (default 1; results do not depend on it). */
void cref_set_threads(int threads);
/* f64 flat traversal (flat.rs:270-371); stats_total = {visited, approximated, direct} */
int cref_flat_accelerations(const cref_tree* t, uint32_t n, const double* x, const double* y,
                            const double* m, const cref_config* cfg, double* ax, double* ay,
                            uint64_t stats_total[3]);
/* BH #1 recursive tree + walk (lib.rs:156-313) */
int cref_recursive_accelerations(uint32_t n, const double* x, const double* y,
                                 const double* m, const cref_config* cfg, double* ax,
                                 double* ay, uint64_t stats_total[3], uint32_t tree_stats[3]);
/* direct_accelerations (lib.rs:331-351), pairwise symmetric order */
int cref_direct_accelerations(uint32_t n, const double* x, const double* y, const double* m,
                              const cref_config* cfg, double* ax, double* ay);
/* direct_acceleration_for (lib.rs:315-329), for a list of target indices
 * (threads > 1 splits targets across OpenMP threads; each result is identical). */
int cref_direct_for(uint32_t n, const double* x, const double* y, const double* m,
                    const cref_config* cfg, uint32_t count, const uint32_t* index,
                    double* ax, double* ay, int threads);
/* probe_error (lib.rs:370-395): out[0]=rms, out[1]=max */
int cref_probe_error(uint32_t n, const double* x, const double* y, const double* m,
                     const double* ax, const double* ay, const cref_config* cfg,
                     uint32_t probes, double out[2], int threads);
/* Plummer-softened total energy (kinetic + pair potential) and Lz, f64. */
void cref_energy(uint32_t n, const double* x, const double* y, const double* vx,
                 const double* vy, const double* m, double softening, double g,
                 double out[3], int threads);

/* ---- integration -------------------------------------------------------- */
/* bh_kick_drift / bh_final_kick (nbody_evolve.wgsl:19-46) on the frozen records */
/* BH #2C correctness-first GPU tree (runtime/src/nbody_tree.wgsl:77-370) on f32
 * states: meta.data = {cells, leaves, max depth seen, overflow}, meta.bounds =
 * {cx, cy, half, 0}; bodies[n] with meta[0] = sorted position; entries[padded]
 * (padded = next power of two >= n); cells in level order (capacity 16n + 1). */
int cref_gpu_tree_build(uint32_t n, const cref_evolve_state* states, uint32_t bucket,
                        uint32_t max_depth, cref_tree_meta* meta, cref_gpu_body* bodies,
                        cref_gpu_entry* entries, uint32_t padded, cref_gpu_cell* cells,
                        uint32_t capacity);
void cref_kick_drift_f32(uint32_t n, cref_evolve_state* s, const cref_gpu_accel* a, float dt);
void cref_final_kick_f32(uint32_t n, cref_evolve_state* s, const cref_gpu_accel* a, float dt);
/* cpu_flat_evolve (galaxy-bh-gpu-tree-parallel.rs:124-148): f64 BH #2A trajectory.
 * body arrays are updated in place. */
int cref_flat_evolve(uint32_t n, double* x, double* y, double* vx, double* vy,
                     const double* m, uint32_t steps, double dt, const cref_config* cfg);
/* The drop-in f32 trajectory: the tree is rebuilt from the widened f32 state every
 * force state (2A-exact build), forces come from the f32 traversal, integration is the
 * f32 KDK.  stats (optional, 3*n per target) and hash (optional) describe the final
 * force state. */
int cref_f32_evolve(uint32_t n, cref_evolve_state* s, cref_gpu_accel* a, uint32_t steps,
                    double dt, const cref_config* cfg, uint32_t* stats, uint64_t* hash);

/* ---- checksums / metrics ------------------------------------------------ */
uint64_t cref_state_checksum(uint32_t n, const double* x, const double* y, const double* vx,
                             const double* vy, const double* m);
void cref_force_error(uint32_t n, const double* ax, const double* ay, const double* bx,
                      const double* by, double out[2]);
void cref_state_error(uint32_t n, const double* ax, const double* ay, const double* avx,
                      const double* avy, const double* bx, const double* by,
                      const double* bvx, const double* bvy, double out[4]);

#ifdef __cplusplus
}
#endif
#endif
