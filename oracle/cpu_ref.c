/*
 * cpu_ref.c — CPU restatement of the GALAXY Barnes-Hut reference path.
 * TEST INFRASTRUCTURE ONLY (the oracle); see cpu_ref.h for the scope and the
 * reference files each function follows.  Build: oracle/Makefile
 * (gcc -O2 -ffp-contract=off, no -march, so no FMA contraction).
 */
#include "cpu_ref.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

static __thread char g_error[256];

static void set_error(const char* msg) {
    snprintf(g_error, sizeof g_error, "%s", msg);
}

const char* cref_last_error(void) { return g_error; }

/* Config::validate (nbody/src/lib.rs:48-67) */
int cref_config_validate(const cref_config* c) {
    if (!isfinite(c->theta) || c->theta < 0.0) { set_error("theta must be finite and >= 0"); return -1; }
    if (!isfinite(c->softening) || c->softening < 0.0) {
        set_error("softening_kpc must be finite and >= 0");
        return -1;
    }
    if (!isfinite(c->g) || c->g <= 0.0) { set_error("g must be positive and finite"); return -1; }
    if (c->bucket == 0 || c->bucket > 1024) { set_error("bucket must be in 1..=1024"); return -1; }
    if (c->max_depth == 0 || c->max_depth > 128) { set_error("max_depth must be in 1..=128"); return -1; }
    return 0;
}

/* validate_bodies (lib.rs:122-135), restricted to the fields the tree uses */
static int validate_bodies(uint32_t n, const double* x, const double* y, const double* m) {
    if (n < 2) { set_error("Barnes-Hut requires at least two resident bodies"); return -1; }
    for (uint32_t i = 0; i < n; ++i) {
        if (!isfinite(x[i]) || !isfinite(y[i])) {
            snprintf(g_error, sizeof g_error, "body %u contains non-finite state", i);
            return -1;
        }
        if (!isfinite(m[i]) || m[i] <= 0.0) {
            snprintf(g_error, sizeof g_error, "body %u mass must be positive and finite", i);
            return -1;
        }
    }
    return 0;
}

/* Rust f64::min / f64::max ignore a NaN operand; inputs here are finite. */
static double rmin(double a, double b) { return b < a ? b : a; }
static double rmax(double a, double b) { return b > a ? b : a; }

/* root_bounds (lib.rs:137-154) */
static void root_bounds(uint32_t n, const double* x, const double* y, double* cx, double* cy,
                        double* half) {
    double min_x = INFINITY, min_y = INFINITY, max_x = -INFINITY, max_y = -INFINITY;
    for (uint32_t i = 0; i < n; ++i) {
        min_x = rmin(min_x, x[i]);
        min_y = rmin(min_y, y[i]);
        max_x = rmax(max_x, x[i]);
        max_y = rmax(max_y, y[i]);
    }
    double span = rmax(rmax(max_x - min_x, max_y - min_y), 1e-12);
    *cx = (min_x + max_x) * 0.5;
    *cy = (min_y + max_y) * 0.5;
    *half = span * 0.500000000001 + 1e-12;
}

/* add_point_mass (lib.rs:247-255) */
static void add_point_mass(double* ax, double* ay, double tx, double ty, double x, double y,
                           double mass, const cref_config* c) {
    double dx = x - tx;
    double dy = y - ty;
    double r2 = dx * dx + dy * dy + c->softening * c->softening;
    double inv_r = 1.0 / sqrt(r2);
    double scale = c->g * mass * inv_r * inv_r * inv_r;
    *ax += dx * scale;
    *ay += dy * scale;
}

static void checksum_word(uint64_t* h, uint64_t w) {
    *h ^= w;
    *h *= 0x00000100000001b3ull;
    *h ^= (w << 23) | (w >> 41);
}

#define FNV_OFFSET 0xcbf29ce484222325ull

/* ======================= BH #2A (nbody/src/flat.rs) ======================= */

/* spread16 / morton2 (flat.rs:67-80) */
static uint32_t spread16(uint32_t v) {
    v &= 0x0000ffffu;
    v = (v | (v << 8)) & 0x00ff00ffu;
    v = (v | (v << 4)) & 0x0f0f0f0fu;
    v = (v | (v << 2)) & 0x33333333u;
    v = (v | (v << 1)) & 0x55555555u;
    return v;
}

static uint32_t morton2(uint32_t x, uint32_t y) { return spread16(x) | (spread16(y) << 1); }

/* quantize_axis (flat.rs:82-87).  Rust clamp keeps NaN and `NaN as u16` is 0. */
static uint32_t quantize_axis(double value, double center, double half) {
    double low = center - half;
    double width = half * 2.0;
    double scaled = floor((value - low) / width * 65536.0);
    if (!(scaled >= 0.0)) return 0;
    if (scaled > 65535.0) return 65535;
    return (uint32_t)scaled;
}

/* exported for the known-answer tests */
uint32_t cref_morton2(uint32_t x, uint32_t y) { return morton2(x & 0xffffu, y & 0xffffu); }
uint32_t cref_quantize_axis(double value, double center, double half) {
    return quantize_axis(value, center, half);
}

struct cref_tree {
    uint32_t n;
    uint32_t *code, *index, *pos;
    uint32_t ncells, cap;
    double *cx, *cy, *half, *mass, *comx, *comy;
    uint32_t *start, *end, *depth, *child; /* child: 4 per cell */
    uint32_t node_count, leaf_count, max_depth;
    double root_cx, root_cy, root_half;
};

typedef struct { uint32_t code, index; } entry_t;

static int entry_cmp(const void* a, const void* b) {
    const entry_t* p = (const entry_t*)a;
    const entry_t* q = (const entry_t*)b;
    if (p->code != q->code) return p->code < q->code ? -1 : 1;
    if (p->index != q->index) return p->index < q->index ? -1 : 1;
    return 0;
}

static int grow_cells(cref_tree* t) {
    uint32_t cap = t->cap ? t->cap * 2 : 64;
    double** dp[] = {&t->cx, &t->cy, &t->half, &t->mass, &t->comx, &t->comy};
    for (int i = 0; i < 6; ++i) {
        double* p = (double*)realloc(*dp[i], (size_t)cap * sizeof(double));
        if (!p) return -1;
        *dp[i] = p;
    }
    uint32_t** up[] = {&t->start, &t->end, &t->depth};
    for (int i = 0; i < 3; ++i) {
        uint32_t* p = (uint32_t*)realloc(*up[i], (size_t)cap * sizeof(uint32_t));
        if (!p) return -1;
        *up[i] = p;
    }
    uint32_t* c = (uint32_t*)realloc(t->child, (size_t)cap * 4 * sizeof(uint32_t));
    if (!c) return -1;
    t->child = c;
    t->cap = cap;
    return 0;
}

/* build_topology (flat.rs:127-177): recursive preorder with 0->3 child order */
static uint32_t build_topology(cref_tree* t, uint32_t start, uint32_t end, double cx, double cy,
                               double half, uint32_t depth, uint32_t bucket, uint32_t max_depth) {
    if (t->ncells == t->cap && grow_cells(t) != 0) return CREF_NO_CHILD;
    uint32_t ci = t->ncells++;
    t->cx[ci] = cx;
    t->cy[ci] = cy;
    t->half[ci] = half;
    t->mass[ci] = 0.0;
    t->comx[ci] = 0.0;
    t->comy[ci] = 0.0;
    t->start[ci] = start;
    t->end[ci] = end;
    t->depth[ci] = depth;
    for (int q = 0; q < 4; ++q) t->child[4 * ci + q] = CREF_NO_CHILD;
    t->node_count += 1;
    if (depth > t->max_depth) t->max_depth = depth;

    if (end - start <= bucket || depth >= max_depth) {
        t->leaf_count += 1;
        return ci;
    }
    uint32_t shift = 30u - 2u * depth;
    uint32_t cursor = start;
    while (cursor < end) {
        uint32_t quadrant = (t->code[cursor] >> shift) & 3u;
        uint32_t group_start = cursor;
        cursor += 1;
        while (cursor < end && ((t->code[cursor] >> shift) & 3u) == quadrant) cursor += 1;
        /* child_bounds (flat.rs:96-103) */
        double child_half = half * 0.5;
        double ccx = cx + ((quadrant & 1u) ? child_half : -child_half);
        double ccy = cy + ((quadrant & 2u) ? child_half : -child_half);
        uint32_t child = build_topology(t, group_start, cursor, ccx, ccy, child_half, depth + 1,
                                        bucket, max_depth);
        if (child == CREF_NO_CHILD) return CREF_NO_CHILD;
        t->child[4 * ci + quadrant] = child;
    }
    return ci;
}

static int is_leaf(const cref_tree* t, uint32_t c) {
    const uint32_t* ch = t->child + 4 * (size_t)c;
    return ch[0] == CREF_NO_CHILD && ch[1] == CREF_NO_CHILD && ch[2] == CREF_NO_CHILD &&
           ch[3] == CREF_NO_CHILD;
}

/* fill_aggregates (flat.rs:179-215): reverse preorder */
static void fill_aggregates(cref_tree* t, const double* x, const double* y, const double* m) {
    for (uint32_t k = t->ncells; k-- > 0;) {
        double mass = 0.0, wx = 0.0, wy = 0.0;
        if (is_leaf(t, k)) {
            for (uint32_t p = t->start[k]; p < t->end[k]; ++p) {
                uint32_t b = t->index[p];
                mass += m[b];
                wx += m[b] * x[b];
                wy += m[b] * y[b];
            }
        } else {
            for (int q = 0; q < 4; ++q) {
                uint32_t c = t->child[4 * (size_t)k + q];
                if (c == CREF_NO_CHILD) continue;
                mass += t->mass[c];
                wx += t->mass[c] * t->comx[c];
                wy += t->mass[c] * t->comy[c];
            }
        }
        t->mass[k] = mass;
        t->comx[k] = wx / mass;
        t->comy[k] = wy / mass;
    }
}

void cref_tree_free(cref_tree* t) {
    if (!t) return;
    free(t->code); free(t->index); free(t->pos);
    free(t->cx); free(t->cy); free(t->half); free(t->mass); free(t->comx); free(t->comy);
    free(t->start); free(t->end); free(t->depth); free(t->child);
    free(t);
}

/* build_flat_tree (flat.rs:217-268) */
cref_tree* cref_build_flat_tree(uint32_t n, const double* x, const double* y, const double* m,
                                const cref_config* cfg) {
    if (validate_bodies(n, x, y, m) != 0 || cref_config_validate(cfg) != 0) return NULL;
    cref_tree* t = (cref_tree*)calloc(1, sizeof(cref_tree));
    if (!t) { set_error("out of memory"); return NULL; }
    t->n = n;
    root_bounds(n, x, y, &t->root_cx, &t->root_cy, &t->root_half);
    entry_t* e = (entry_t*)malloc((size_t)n * sizeof(entry_t));
    t->code = (uint32_t*)malloc((size_t)n * 4);
    t->index = (uint32_t*)malloc((size_t)n * 4);
    t->pos = (uint32_t*)malloc((size_t)n * 4);
    if (!e || !t->code || !t->index || !t->pos) {
        free(e);
        cref_tree_free(t);
        set_error("out of memory");
        return NULL;
    }
    for (uint32_t i = 0; i < n; ++i) {
        e[i].code = morton2(quantize_axis(x[i], t->root_cx, t->root_half),
                            quantize_axis(y[i], t->root_cy, t->root_half));
        e[i].index = i;
    }
    /* sort_by_key(code) is stable; (code, index) is the same total order. */
    qsort(e, n, sizeof(entry_t), entry_cmp);
    for (uint32_t p = 0; p < n; ++p) {
        t->code[p] = e[p].code;
        t->index[p] = e[p].index;
        t->pos[e[p].index] = p;
    }
    free(e);
    uint32_t max_depth = cfg->max_depth < CREF_MORTON_TREE_DEPTH ? cfg->max_depth : CREF_MORTON_TREE_DEPTH;
    if (build_topology(t, 0, n, t->root_cx, t->root_cy, t->root_half, 0, cfg->bucket, max_depth) ==
        CREF_NO_CHILD) {
        cref_tree_free(t);
        set_error("out of memory");
        return NULL;
    }
    fill_aggregates(t, x, y, m);
    return t;
}

void cref_tree_counts(const cref_tree* t, uint32_t counts[4]) {
    counts[0] = t->ncells;
    counts[1] = t->node_count;
    counts[2] = t->leaf_count;
    counts[3] = t->max_depth;
}

void cref_tree_root(const cref_tree* t, double root[3]) {
    root[0] = t->root_cx;
    root[1] = t->root_cy;
    root[2] = t->root_half;
}

void cref_tree_entries(const cref_tree* t, uint32_t* code, uint32_t* body_index,
                       uint32_t* position_by_body) {
    if (code) memcpy(code, t->code, (size_t)t->n * 4);
    if (body_index) memcpy(body_index, t->index, (size_t)t->n * 4);
    if (position_by_body) memcpy(position_by_body, t->pos, (size_t)t->n * 4);
}

void cref_tree_cells(const cref_tree* t, double* geom, uint32_t* topo) {
    for (uint32_t k = 0; k < t->ncells; ++k) {
        if (geom) {
            double* g = geom + 6 * (size_t)k;
            g[0] = t->cx[k]; g[1] = t->cy[k]; g[2] = t->half[k];
            g[3] = t->mass[k]; g[4] = t->comx[k]; g[5] = t->comy[k];
        }
        if (topo) {
            uint32_t* o = topo + 7 * (size_t)k;
            o[0] = t->start[k]; o[1] = t->end[k]; o[2] = t->depth[k];
            for (int q = 0; q < 4; ++q) o[3 + q] = t->child[4 * (size_t)k + q];
        }
    }
}

/* flat_topology_checksum (flat.rs:416-433) */
uint64_t cref_flat_topology_checksum(const cref_tree* t) {
    uint64_t h = FNV_OFFSET;
    checksum_word(&h, 16);
    checksum_word(&h, t->n);
    for (uint32_t p = 0; p < t->n; ++p) {
        checksum_word(&h, t->code[p]);
        checksum_word(&h, t->index[p]);
    }
    for (uint32_t k = 0; k < t->ncells; ++k) {
        checksum_word(&h, t->start[k]);
        checksum_word(&h, t->end[k]);
        checksum_word(&h, t->depth[k]);
        for (int q = 0; q < 4; ++q) checksum_word(&h, t->child[4 * (size_t)k + q]);
    }
    return h;
}

/* checked_f32 (nbody_gpu.rs:167-173) */
static int checked_f32(double v, float* out, const char* name) {
    float f = (float)v;
    if (!isfinite(f)) {
        snprintf(g_error, sizeof g_error, "%s cannot be represented as finite f32", name);
        return -1;
    }
    *out = f;
    return 0;
}

/* pack_flat_tree (nbody_gpu.rs:175-250) */
int cref_pack_flat_tree(const cref_tree* t, uint32_t n, const double* x, const double* y,
                        const double* m, const cref_config* cfg, cref_bh_settings* settings,
                        cref_gpu_body* bodies, cref_gpu_entry* entries, cref_gpu_cell* cells) {
    if (cref_config_validate(cfg) != 0) return -1;
    if (n != t->n) { set_error("flat tree/body cardinality mismatch"); return -1; }
    for (uint32_t i = 0; i < n; ++i) {
        cref_gpu_body* b = bodies + i;
        if (checked_f32(x[i], &b->position_mass[0], "body.x") ||
            checked_f32(y[i], &b->position_mass[1], "body.y") ||
            checked_f32(m[i], &b->position_mass[2], "body.mass"))
            return -1;
        b->position_mass[3] = 0.0f;
        b->meta[0] = t->pos[i];
        b->meta[1] = b->meta[2] = b->meta[3] = 0;
    }
    for (uint32_t p = 0; p < n; ++p) {
        entries[p].data[0] = t->code[p];
        entries[p].data[1] = t->index[p];
        entries[p].data[2] = entries[p].data[3] = 0;
    }
    for (uint32_t k = 0; k < t->ncells; ++k) {
        cref_gpu_cell* c = cells + k;
        if (checked_f32(t->cx[k], &c->center_half_mass[0], "cell.cx") ||
            checked_f32(t->cy[k], &c->center_half_mass[1], "cell.cy") ||
            checked_f32(t->half[k], &c->center_half_mass[2], "cell.half") ||
            checked_f32(t->mass[k], &c->center_half_mass[3], "cell.mass") ||
            checked_f32(t->comx[k], &c->com[0], "cell.com_x") ||
            checked_f32(t->comy[k], &c->com[1], "cell.com_y"))
            return -1;
        c->com[2] = c->com[3] = 0.0f;
        c->range_depth[0] = t->start[k];
        c->range_depth[1] = t->end[k];
        c->range_depth[2] = t->depth[k];
        c->range_depth[3] = 0;
        for (int q = 0; q < 4; ++q) c->children[q] = t->child[4 * (size_t)k + q];
    }
    settings->info[0] = n;
    settings->info[1] = t->ncells;
    settings->info[2] = 0;
    settings->info[3] = 0;
    if (checked_f32(cfg->theta, &settings->physics[0], "theta") ||
        checked_f32(cfg->softening, &settings->physics[1], "softening_kpc") ||
        checked_f32(cfg->g, &settings->physics[2], "G"))
        return -1;
    settings->physics[3] = 0.0f;
    return 0;
}

/* BH #2D evidence layout (nbody_tree_parallel_topology.wgsl:148-215,
 * nbody_parallel.rs:929-988): slot = depth*N + start, filtered to valid cells
 * and read back in slot order. */
typedef struct { uint64_t slot; uint32_t cell; } slot_t;

static int slot_cmp(const void* a, const void* b) {
    const slot_t* p = (const slot_t*)a;
    const slot_t* q = (const slot_t*)b;
    return p->slot < q->slot ? -1 : (p->slot > q->slot ? 1 : 0);
}

void cref_sparse_evidence(const cref_tree* t, const cref_config* cfg, cref_tree_meta* meta,
                          cref_gpu_entry* entries, cref_gpu_cell* cells) {
    (void)cfg;
    uint32_t n = t->n;
    memset(meta, 0, sizeof *meta);
    meta->bounds[0] = (float)t->root_cx;
    meta->bounds[1] = (float)t->root_cy;
    meta->bounds[2] = (float)t->root_half;
    meta->bounds[3] = 0.0f;
    for (uint32_t p = 0; p < n; ++p) {
        entries[p].data[0] = t->code[p];
        entries[p].data[1] = t->index[p];
        entries[p].data[2] = entries[p].data[3] = 0;
    }
    slot_t* s = (slot_t*)malloc((size_t)t->ncells * sizeof(slot_t));
    for (uint32_t k = 0; k < t->ncells; ++k) {
        s[k].slot = (uint64_t)t->depth[k] * n + t->start[k];
        s[k].cell = k;
    }
    qsort(s, t->ncells, sizeof(slot_t), slot_cmp);
    for (uint32_t i = 0; i < t->ncells; ++i) {
        uint32_t k = s[i].cell;
        cref_gpu_cell* c = cells + i;
        c->center_half_mass[0] = (float)t->cx[k];
        c->center_half_mass[1] = (float)t->cy[k];
        c->center_half_mass[2] = (float)t->half[k];
        c->center_half_mass[3] = (float)t->mass[k];
        c->com[0] = (float)t->comx[k];
        c->com[1] = (float)t->comy[k];
        c->com[2] = c->com[3] = 0.0f;
        c->range_depth[0] = t->start[k];
        c->range_depth[1] = t->end[k];
        c->range_depth[2] = t->depth[k];
        c->range_depth[3] = 1;
        for (int q = 0; q < 4; ++q) {
            uint32_t ch = t->child[4 * (size_t)k + q];
            c->children[q] = ch == CREF_NO_CHILD ? CREF_NO_CHILD : t->depth[ch] * n + t->start[ch];
        }
    }
    free(s);
}

static uint32_t fbits(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    return u;
}

/* parallel_tree_checksum (galaxy-bh-gpu-tree-parallel.rs:156-187) */
uint64_t cref_parallel_tree_checksum(const cref_tree_meta* meta, uint32_t n,
                                     const cref_gpu_entry* entries, uint32_t ncells,
                                     const cref_gpu_cell* cells) {
    uint64_t h = FNV_OFFSET;
    for (int i = 0; i < 4; ++i) checksum_word(&h, meta->data[i]);
    for (int i = 0; i < 4; ++i) checksum_word(&h, fbits(meta->bounds[i]));
    for (uint32_t p = 0; p < n; ++p)
        for (int i = 0; i < 4; ++i) checksum_word(&h, entries[p].data[i]);
    for (uint32_t k = 0; k < ncells; ++k) {
        const cref_gpu_cell* c = cells + k;
        for (int i = 0; i < 4; ++i) checksum_word(&h, fbits(c->center_half_mass[i]));
        for (int i = 0; i < 4; ++i) checksum_word(&h, fbits(c->com[i]));
        for (int i = 0; i < 4; ++i) checksum_word(&h, c->range_depth[i]);
        for (int i = 0; i < 4; ++i) checksum_word(&h, c->children[i]);
    }
    return h;
}

/* validate_parallel_order (galaxy-bh-gpu-tree-parallel.rs:189-203) */
int cref_validate_parallel_order(uint32_t n, const cref_gpu_entry* e) {
    for (uint32_t p = 1; p < n; ++p) {
        const uint32_t* a = e[p - 1].data;
        const uint32_t* b = e[p].data;
        if (a[0] > b[0] || (a[0] == b[0] && a[1] >= b[1])) {
            snprintf(g_error, sizeof g_error,
                     "radix order is not strict by (Morton code, body index) at %u", p);
            return -1;
        }
    }
    return 0;
}

/* ====================== f32 traversal (nbody.wgsl) ======================= */

/* point_mass (nbody.wgsl:35-41).  inverseSqrt is defined here as 1/sqrt, each
 * correctly rounded. */
static void point_mass_f32(float tx, float ty, float sx, float sy, float mass, float eps,
                           float g, float* ox, float* oy) {
    float dx = sx - tx;
    float dy = sy - ty;
    float dd = dx * dx;
    float ee = dy * dy;
    float r2 = (dd + ee) + eps * eps;
    float s = sqrtf(r2);
    float inv_r = 1.0f / s;
    float scale = g * mass;
    scale = scale * inv_r;
    scale = scale * inv_r;
    scale = scale * inv_r;
    *ox = dx * scale;
    *oy = dy * scale;
}

/* bh_traverse (nbody.wgsl:43-123) */
/* One target of bh_traverse (nbody.wgsl:43-123): returns 0, or -1 with the
 * error set. */
static int traverse_one_f32(const cref_bh_settings* st, const cref_gpu_body* bodies,
                            const cref_gpu_entry* entries, const cref_gpu_cell* cells,
                            uint32_t t, cref_gpu_accel* out, uint32_t* stats, uint64_t* hash) {
    uint32_t ncells = st->info[1];
    float theta = st->physics[0], eps = st->physics[1], g = st->physics[2];
    float tx = bodies[t].position_mass[0];
    float ty = bodies[t].position_mass[1];
    uint32_t tpos = bodies[t].meta[0];
    float ax = 0.0f, ay = 0.0f;
    uint32_t stack[64];
    uint32_t len = 1;
    stack[0] = st->info[2];
    uint32_t visited = 0, accepted = 0, direct = 0;
    uint64_t h = FNV_OFFSET;
    while (len > 0) {
        uint32_t ci = stack[--len];
        if (ci >= ncells) { set_error("cell index out of range"); return -1; }
        const cref_gpu_cell* c = cells + ci;
        visited += 1;
        int leaf = c->children[0] == CREF_NO_CHILD && c->children[1] == CREF_NO_CHILD &&
                   c->children[2] == CREF_NO_CHILD && c->children[3] == CREF_NO_CHILD;
        if (leaf) {
            for (uint32_t p = c->range_depth[0]; p < c->range_depth[1]; ++p) {
                uint32_t other = entries[p].data[1];
                if (other != t) {
                    float px, py;
                    const float* o = bodies[other].position_mass;
                    point_mass_f32(tx, ty, o[0], o[1], o[2], eps, g, &px, &py);
                    ax = ax + px;
                    ay = ay + py;
                    direct += 1;
                    checksum_word(&h, (1ull << 60) | other);
                }
            }
            continue;
        }
        float dx = c->com[0] - tx;
        float dy = c->com[1] - ty;
        float d2 = dx * dx + dy * dy;
        int contains = tpos >= c->range_depth[0] && tpos < c->range_depth[1];
        if (!contains && d2 > 0.0f) {
            float distance = sqrtf(d2);
            float w = 2.0f * c->center_half_mass[2];
            if (w / distance < theta) {
                float px, py;
                point_mass_f32(tx, ty, c->com[0], c->com[1], c->center_half_mass[3], eps, g,
                               &px, &py);
                ax = ax + px;
                ay = ay + py;
                accepted += 1;
                checksum_word(&h, (2ull << 60) | ((uint64_t)c->range_depth[2] << 32) |
                                      c->range_depth[0]);
                continue;
            }
        }
        for (int q = 3; q >= 0; --q) {
            if (c->children[q] != CREF_NO_CHILD) {
                if (len >= 64) { set_error("traversal stack overflow"); return -1; }
                stack[len++] = c->children[q];
            }
        }
    }
    out[t].value[0] = ax;
    out[t].value[1] = ay;
    out[t].value[2] = 0.0f;
    out[t].value[3] = 0.0f;
    if (stats) {
        stats[3 * (size_t)t + 0] = visited;
        stats[3 * (size_t)t + 1] = accepted;
        stats[3 * (size_t)t + 2] = direct;
    }
    if (hash) hash[t] = h;
    return 0;
}

/* Threads of the f32 traversal (cref_set_threads); 1 by default.  Targets are
 * independent, so every thread count gives identical results. */
static int g_threads = 1;

void cref_set_threads(int threads) { g_threads = threads > 0 ? threads : 1; }

int cref_traverse_f32(const cref_bh_settings* st, const cref_gpu_body* bodies,
                      const cref_gpu_entry* entries, const cref_gpu_cell* cells,
                      cref_gpu_accel* out, uint32_t* stats, uint64_t* hash) {
    const int64_t n = st->info[0];
    int failed = 0;
    char first_error[sizeof g_error] = {0};
#pragma omp parallel for schedule(dynamic, 256) num_threads(g_threads)
    for (int64_t t = 0; t < n; ++t) {
        int stop;
#pragma omp atomic read
        stop = failed;
        if (stop) continue;
        if (traverse_one_f32(st, bodies, entries, cells, (uint32_t)t, out, stats, hash) != 0) {
#pragma omp critical
            {
                if (!failed) memcpy(first_error, g_error, sizeof first_error);
                failed = 1;
            }
        }
    }
    if (failed) {
        memcpy(g_error, first_error, sizeof g_error);
        return -1;
    }
    return 0;
}

/* ===================== f64 traversal (flat.rs:270-371) ==================== */
int cref_flat_accelerations(const cref_tree* t, uint32_t n, const double* x, const double* y,
                            const double* m, const cref_config* cfg, double* ax_out,
                            double* ay_out, uint64_t stats_total[3]) {
    if (validate_bodies(n, x, y, m) != 0 || cref_config_validate(cfg) != 0) return -1;
    if (t->n != n) { set_error("flat tree body count does not match resident bodies"); return -1; }
    uint64_t visited = 0, approx = 0, direct = 0;
    uint32_t* stack = (uint32_t*)malloc(((size_t)t->max_depth * 3 + 8) * sizeof(uint32_t));
    for (uint32_t ti = 0; ti < n; ++ti) {
        double tx = x[ti], ty = y[ti];
        uint32_t tpos = t->pos[ti];
        double ax = 0.0, ay = 0.0;
        uint32_t len = 1;
        stack[0] = 0;
        while (len > 0) {
            uint32_t k = stack[--len];
            visited += 1;
            if (is_leaf(t, k)) {
                for (uint32_t p = t->start[k]; p < t->end[k]; ++p) {
                    uint32_t o = t->index[p];
                    if (o == ti) continue;
                    add_point_mass(&ax, &ay, tx, ty, x[o], y[o], m[o], cfg);
                    direct += 1;
                }
                continue;
            }
            double dx = t->comx[k] - tx;
            double dy = t->comy[k] - ty;
            double distance = hypot(dx, dy);
            double width = t->half[k] * 2.0;
            int contains = tpos >= t->start[k] && tpos < t->end[k];
            if (!contains && distance > 0.0 && width / distance < cfg->theta) {
                add_point_mass(&ax, &ay, tx, ty, t->comx[k], t->comy[k], t->mass[k], cfg);
                approx += 1;
                continue;
            }
            for (int q = 3; q >= 0; --q) {
                uint32_t c = t->child[4 * (size_t)k + q];
                if (c != CREF_NO_CHILD) stack[len++] = c;
            }
        }
        ax_out[ti] = ax;
        ay_out[ti] = ay;
    }
    free(stack);
    if (stats_total) {
        stats_total[0] = visited;
        stats_total[1] = approx;
        stats_total[2] = direct;
    }
    return 0;
}

/* ===================== BH #1 recursive (lib.rs:156-313) =================== */
typedef struct rnode {
    double cx, cy, half, mass, comx, comy;
    uint32_t count, depth;
    struct rnode* child[4];
    uint32_t* idx; /* non-NULL only for leaves */
} rnode;

typedef struct { uint32_t node_count, leaf_count, max_depth; } rstats;

static rnode* rbuild(const double* x, const double* y, const double* m, uint32_t* idx,
                     uint32_t count, double cx, double cy, double half, uint32_t depth,
                     const cref_config* c, rstats* st) {
    rnode* nd = (rnode*)calloc(1, sizeof(rnode));
    st->node_count += 1;
    if (depth > st->max_depth) st->max_depth = depth;
    double mass = 0.0, wx = 0.0, wy = 0.0;
    for (uint32_t k = 0; k < count; ++k) {
        uint32_t i = idx[k];
        mass += m[i];
        wx += m[i] * x[i];
        wy += m[i] * y[i];
    }
    nd->cx = cx; nd->cy = cy; nd->half = half; nd->mass = mass;
    nd->comx = wx / mass; nd->comy = wy / mass;
    nd->count = count; nd->depth = depth;
    if (count <= c->bucket || depth >= c->max_depth || half <= 2.220446049250313e-16 * 32.0) {
        st->leaf_count += 1;
        nd->idx = (uint32_t*)malloc((size_t)count * 4);
        memcpy(nd->idx, idx, (size_t)count * 4);
        return nd;
    }
    uint32_t* groups[4];
    uint32_t gn[4] = {0, 0, 0, 0};
    for (int q = 0; q < 4; ++q) groups[q] = (uint32_t*)malloc((size_t)count * 4);
    for (uint32_t k = 0; k < count; ++k) {
        uint32_t i = idx[k];
        int q = (x[i] >= cx) | ((y[i] >= cy) << 1);
        groups[q][gn[q]++] = i;
    }
    for (int q = 0; q < 4; ++q) {
        if (gn[q] > 0) {
            double h = half * 0.5;
            double ccx = cx + ((q & 1) ? h : -h);
            double ccy = cy + ((q & 2) ? h : -h);
            nd->child[q] = rbuild(x, y, m, groups[q], gn[q], ccx, ccy, h, depth + 1, c, st);
        }
        free(groups[q]);
    }
    return nd;
}

static void rfree(rnode* nd) {
    if (!nd) return;
    for (int q = 0; q < 4; ++q) rfree(nd->child[q]);
    free(nd->idx);
    free(nd);
}

static void rwalk(uint32_t ti, double tx, double ty, const double* x, const double* y,
                  const double* m, const rnode* nd, const cref_config* c, double* ax, double* ay,
                  uint64_t* st) {
    st[0] += 1;
    if (nd->idx) {
        for (uint32_t k = 0; k < nd->count; ++k) {
            uint32_t o = nd->idx[k];
            if (o == ti) continue;
            add_point_mass(ax, ay, tx, ty, x[o], y[o], m[o], c);
            st[2] += 1;
        }
        return;
    }
    double dx = nd->comx - tx;
    double dy = nd->comy - ty;
    double distance = hypot(dx, dy);
    double width = nd->half * 2.0;
    int contains = tx >= nd->cx - nd->half && tx <= nd->cx + nd->half &&
                   ty >= nd->cy - nd->half && ty <= nd->cy + nd->half;
    if (!contains && distance > 0.0 && width / distance < c->theta) {
        add_point_mass(ax, ay, tx, ty, nd->comx, nd->comy, nd->mass, c);
        st[1] += 1;
        return;
    }
    for (int q = 0; q < 4; ++q)
        if (nd->child[q]) rwalk(ti, tx, ty, x, y, m, nd->child[q], c, ax, ay, st);
}

int cref_recursive_accelerations(uint32_t n, const double* x, const double* y,
                                 const double* m, const cref_config* cfg, double* ax,
                                 double* ay, uint64_t stats_total[3], uint32_t tree_stats[3]) {
    if (validate_bodies(n, x, y, m) != 0 || cref_config_validate(cfg) != 0) return -1;
    double cx, cy, half;
    root_bounds(n, x, y, &cx, &cy, &half);
    uint32_t* idx = (uint32_t*)malloc((size_t)n * 4);
    for (uint32_t i = 0; i < n; ++i) idx[i] = i;
    rstats rs = {0, 0, 0};
    rnode* root = rbuild(x, y, m, idx, n, cx, cy, half, 0, cfg, &rs);
    free(idx);
    uint64_t st[3] = {0, 0, 0};
    for (uint32_t i = 0; i < n; ++i) {
        ax[i] = 0.0;
        ay[i] = 0.0;
        rwalk(i, x[i], y[i], x, y, m, root, cfg, ax + i, ay + i, st);
    }
    rfree(root);
    if (stats_total) memcpy(stats_total, st, sizeof st);
    if (tree_stats) {
        tree_stats[0] = rs.node_count;
        tree_stats[1] = rs.leaf_count;
        tree_stats[2] = rs.max_depth;
    }
    return 0;
}

/* ======================== direct sums (lib.rs) =========================== */
int cref_direct_accelerations(uint32_t n, const double* x, const double* y, const double* m,
                              const cref_config* cfg, double* ax, double* ay) {
    if (validate_bodies(n, x, y, m) != 0 || cref_config_validate(cfg) != 0) return -1;
    for (uint32_t i = 0; i < n; ++i) ax[i] = ay[i] = 0.0;
    double e2 = cfg->softening * cfg->softening;
    for (uint32_t i = 0; i < n; ++i) {
        for (uint32_t j = i + 1; j < n; ++j) {
            double dx = x[j] - x[i];
            double dy = y[j] - y[i];
            double r2 = dx * dx + dy * dy + e2;
            double inv_r = 1.0 / sqrt(r2);
            double common = cfg->g * inv_r * inv_r * inv_r;
            ax[i] += dx * common * m[j];
            ay[i] += dy * common * m[j];
            ax[j] -= dx * common * m[i];
            ay[j] -= dy * common * m[i];
        }
    }
    return 0;
}

static void direct_one(uint32_t n, const double* x, const double* y, const double* m,
                       const cref_config* cfg, uint32_t i, double* ax, double* ay) {
    double sx = 0.0, sy = 0.0;
    for (uint32_t j = 0; j < n; ++j)
        if (j != i) add_point_mass(&sx, &sy, x[i], y[i], x[j], y[j], m[j], cfg);
    *ax = sx;
    *ay = sy;
}

int cref_direct_for(uint32_t n, const double* x, const double* y, const double* m,
                    const cref_config* cfg, uint32_t count, const uint32_t* index, double* ax,
                    double* ay, int threads) {
    if (validate_bodies(n, x, y, m) != 0 || cref_config_validate(cfg) != 0) return -1;
    for (uint32_t k = 0; k < count; ++k)
        if (index[k] >= n) { set_error("body index is out of range"); return -1; }
    (void)threads;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 4) num_threads(threads > 0 ? threads : 1)
#endif
    for (int64_t k = 0; k < (int64_t)count; ++k) direct_one(n, x, y, m, cfg, index[k], ax + k, ay + k);
    return 0;
}

int cref_probe_error(uint32_t n, const double* x, const double* y, const double* m,
                     const double* ax, const double* ay, const cref_config* cfg,
                     uint32_t probes, double out[2], int threads) {
    if (probes < 1) probes = 1;
    if (probes > n) probes = n;
    uint32_t* idx = (uint32_t*)calloc((size_t)probes, 4);
    double* ex = (double*)malloc((size_t)probes * 8);
    double* ey = (double*)malloc((size_t)probes * 8);
    uint32_t den = probes > 1 ? probes - 1 : 1;
    for (uint32_t k = 0; k < probes; ++k)
        idx[k] = (uint32_t)(((uint64_t)k * (uint64_t)(n - 1)) / den);
    int rc = cref_direct_for(n, x, y, m, cfg, probes, idx, ex, ey, threads);
    if (rc == 0) {
        double sum_sq = 0.0, mx = 0.0;
        for (uint32_t k = 0; k < probes; ++k) {
            uint32_t i = idx[k];
            double denom = rmax(hypot(ex[k], ey[k]), 1e-30);
            double rel = hypot(ax[i] - ex[k], ay[i] - ey[k]) / denom;
            sum_sq += rel * rel;
            mx = rmax(mx, rel);
        }
        out[0] = sqrt(sum_sq / (double)probes);
        out[1] = mx;
    }
    free(idx); free(ex); free(ey);
    return rc;
}

void cref_energy(uint32_t n, const double* x, const double* y, const double* vx,
                 const double* vy, const double* m, double softening, double g, double out[3],
                 int threads) {
    double* pot = (double*)malloc((size_t)n * 8);
    double e2 = softening * softening;
    (void)threads;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 64) num_threads(threads > 0 ? threads : 1)
#endif
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        double s = 0.0;
        for (uint32_t j = (uint32_t)i + 1; j < n; ++j) {
            double dx = x[j] - x[i];
            double dy = y[j] - y[i];
            s += m[j] / sqrt(dx * dx + dy * dy + e2);
        }
        pot[i] = -g * m[i] * s;
    }
    double kin = 0.0, potential = 0.0, lz = 0.0;
    for (uint32_t i = 0; i < n; ++i) {
        kin += 0.5 * m[i] * (vx[i] * vx[i] + vy[i] * vy[i]);
        potential += pot[i];
        lz += m[i] * (x[i] * vy[i] - y[i] * vx[i]);
    }
    free(pot);
    out[0] = kin + potential;
    out[1] = lz;
    out[2] = kin;
}

/* ===================== integration (nbody_evolve.wgsl) =================== */
void cref_kick_drift_f32(uint32_t n, cref_evolve_state* s, const cref_gpu_accel* a, float dt) {
    float h = 0.5f * dt;
    for (uint32_t i = 0; i < n; ++i) {
        float hvx = s[i].velocity[0] + h * a[i].value[0];
        float hvy = s[i].velocity[1] + h * a[i].value[1];
        s[i].position_mass[0] = s[i].position_mass[0] + dt * hvx;
        s[i].position_mass[1] = s[i].position_mass[1] + dt * hvy;
        s[i].velocity[0] = hvx;
        s[i].velocity[1] = hvy;
    }
}

void cref_final_kick_f32(uint32_t n, cref_evolve_state* s, const cref_gpu_accel* a, float dt) {
    float h = 0.5f * dt;
    for (uint32_t i = 0; i < n; ++i) {
        s[i].velocity[0] = s[i].velocity[0] + h * a[i].value[0];
        s[i].velocity[1] = s[i].velocity[1] + h * a[i].value[1];
    }
}

/* cpu_flat_evolve (galaxy-bh-gpu-tree-parallel.rs:124-148) */
int cref_flat_evolve(uint32_t n, double* x, double* y, double* vx, double* vy, const double* m,
                     uint32_t steps, double dt, const cref_config* cfg) {
    double* ax = (double*)malloc((size_t)n * 8);
    double* ay = (double*)malloc((size_t)n * 8);
    int rc = -1;
    cref_tree* t = cref_build_flat_tree(n, x, y, m, cfg);
    if (!t) goto out;
    if (cref_flat_accelerations(t, n, x, y, m, cfg, ax, ay, NULL) != 0) goto out;
    for (uint32_t s = 0; s < steps; ++s) {
        for (uint32_t i = 0; i < n; ++i) {
            vx[i] += 0.5 * dt * ax[i];
            vy[i] += 0.5 * dt * ay[i];
            x[i] += dt * vx[i];
            y[i] += dt * vy[i];
        }
        cref_tree_free(t);
        t = cref_build_flat_tree(n, x, y, m, cfg);
        if (!t) goto out;
        if (cref_flat_accelerations(t, n, x, y, m, cfg, ax, ay, NULL) != 0) goto out;
        for (uint32_t i = 0; i < n; ++i) {
            vx[i] += 0.5 * dt * ax[i];
            vy[i] += 0.5 * dt * ay[i];
        }
    }
    rc = 0;
out:
    cref_tree_free(t);
    free(ax);
    free(ay);
    return rc;
}

/* One f32 force state: widen the f32 state, build the 2A tree, pack it, traverse f32. */
static int f32_force(uint32_t n, const cref_evolve_state* s, cref_gpu_accel* a,
                     const cref_config* cfg, uint32_t* stats, uint64_t* hash, double* x,
                     double* y, double* m, cref_gpu_body* bodies, cref_gpu_entry* entries) {
    for (uint32_t i = 0; i < n; ++i) {
        x[i] = s[i].position_mass[0];
        y[i] = s[i].position_mass[1];
        m[i] = s[i].position_mass[2];
    }
    cref_tree* t = cref_build_flat_tree(n, x, y, m, cfg);
    if (!t) return -1;
    cref_gpu_cell* cells = (cref_gpu_cell*)malloc((size_t)t->ncells * sizeof(cref_gpu_cell));
    cref_bh_settings st;
    int rc = cref_pack_flat_tree(t, n, x, y, m, cfg, &st, bodies, entries, cells);
    if (rc == 0) rc = cref_traverse_f32(&st, bodies, entries, cells, a, stats, hash);
    free(cells);
    cref_tree_free(t);
    return rc;
}

int cref_f32_evolve(uint32_t n, cref_evolve_state* s, cref_gpu_accel* a, uint32_t steps,
                    double dt, const cref_config* cfg, uint32_t* stats, uint64_t* hash) {
    float dtf;
    if (!isfinite(dt) || dt <= 0.0) { set_error("dt_myr must be positive and finite"); return -1; }
    if (checked_f32(dt, &dtf, "dt_myr") != 0) return -1;
    double* x = (double*)malloc((size_t)n * 8);
    double* y = (double*)malloc((size_t)n * 8);
    double* m = (double*)malloc((size_t)n * 8);
    cref_gpu_body* bodies = (cref_gpu_body*)malloc((size_t)n * sizeof(cref_gpu_body));
    cref_gpu_entry* entries = (cref_gpu_entry*)malloc((size_t)n * sizeof(cref_gpu_entry));
    int rc = f32_force(n, s, a, cfg, steps == 0 ? stats : NULL, steps == 0 ? hash : NULL, x, y,
                       m, bodies, entries);
    for (uint32_t k = 0; rc == 0 && k < steps; ++k) {
        cref_kick_drift_f32(n, s, a, dtf);
        rc = f32_force(n, s, a, cfg, k + 1 == steps ? stats : NULL,
                       k + 1 == steps ? hash : NULL, x, y, m, bodies, entries);
        if (rc == 0) cref_final_kick_f32(n, s, a, dtf);
    }
    free(x); free(y); free(m); free(bodies); free(entries);
    return rc;
}

/* ====================== checksums / metrics ============================== */
static uint64_t dbits(double d) {
    uint64_t u;
    memcpy(&u, &d, 8);
    return u;
}

/* state_checksum (lib.rs:530-543) */
uint64_t cref_state_checksum(uint32_t n, const double* x, const double* y, const double* vx,
                             const double* vy, const double* m) {
    uint64_t h = FNV_OFFSET;
    for (uint32_t i = 0; i < n; ++i) {
        checksum_word(&h, dbits(x[i]));
        checksum_word(&h, dbits(y[i]));
        checksum_word(&h, dbits(vx[i]));
        checksum_word(&h, dbits(vy[i]));
        checksum_word(&h, dbits(m[i]));
    }
    return h;
}

/* force_error (galaxy-bh-gpu-tree-parallel.rs:101-114) */
void cref_force_error(uint32_t n, const double* ax, const double* ay, const double* bx,
                      const double* by, double out[2]) {
    double sum_sq = 0.0, mx = 0.0;
    for (uint32_t i = 0; i < n; ++i) {
        double denom = rmax(hypot(bx[i], by[i]), 1e-30);
        double rel = hypot(ax[i] - bx[i], ay[i] - by[i]) / denom;
        sum_sq += rel * rel;
        mx = rmax(mx, rel);
    }
    out[0] = sqrt(sum_sq / (double)n);
    out[1] = mx;
}

/* state_error (galaxy-bh-gpu-tree-parallel.rs:62-90) */
void cref_state_error(uint32_t n, const double* ax, const double* ay, const double* avx,
                      const double* avy, const double* bx, const double* by,
                      const double* bvx, const double* bvy, double out[4]) {
    double p_err = 0.0, p_ref = 0.0, v_err = 0.0, v_ref = 0.0, p_max = 0.0, v_max = 0.0;
    for (uint32_t i = 0; i < n; ++i) {
        double pd = hypot(ax[i] - bx[i], ay[i] - by[i]);
        double vd = hypot(avx[i] - bvx[i], avy[i] - bvy[i]);
        double pr = hypot(bx[i], by[i]);
        double vr = hypot(bvx[i], bvy[i]);
        p_err += pd * pd;
        p_ref += pr * pr;
        v_err += vd * vd;
        v_ref += vr * vr;
        p_max = rmax(p_max, pd / rmax(pr, 1e-6));
        v_max = rmax(v_max, vd / rmax(vr, 1e-9));
    }
    out[0] = sqrt(p_err / rmax(p_ref, 1e-30));
    out[1] = p_max;
    out[2] = sqrt(v_err / rmax(v_ref, 1e-30));
    out[3] = v_max;
}

/* ============ BH #2C correctness-first GPU tree (nbody_tree.wgsl) =========
 * The reference's serialized algorithm, step for step, in f32:
 * bh_tree_bounds (:77-108), bh_tree_morton (:110-131), bh_tree_sort (:133-167,
 * bitonic network over the padded entries with the (code, index) order of
 * :72-78), bh_tree_positions (:169-176), bh_tree_topology (:204-302, level by
 * level, children in quadrant order), bh_tree_aggregate (:304-370, reverse
 * cell order).
 */
static float quantize_f32(float value, float center, float half) {
    float low = center - half;
    float width = 2.0f * half;
    float scaled = floorf((value - low) / width * 65536.0f);
    if (!(scaled > 0.0f)) scaled = 0.0f; /* clamp(scaled, 0, 65535) */
    if (scaled > 65535.0f) scaled = 65535.0f;
    return scaled;
}

static int entry_after(const cref_gpu_entry* a, const cref_gpu_entry* b) {
    return a->data[0] > b->data[0] || (a->data[0] == b->data[0] && a->data[1] > b->data[1]);
}

static int entry_before(const cref_gpu_entry* a, const cref_gpu_entry* b) {
    return a->data[0] < b->data[0] || (a->data[0] == b->data[0] && a->data[1] < b->data[1]);
}

static cref_gpu_cell empty_cell_f32(float cx, float cy, float half, uint32_t start, uint32_t end,
                                    uint32_t depth) {
    cref_gpu_cell c;
    memset(&c, 0, sizeof c);
    c.center_half_mass[0] = cx;
    c.center_half_mass[1] = cy;
    c.center_half_mass[2] = half;
    c.range_depth[0] = start;
    c.range_depth[1] = end;
    c.range_depth[2] = depth;
    for (int q = 0; q < 4; ++q) c.children[q] = CREF_NO_CHILD;
    return c;
}

int cref_gpu_tree_build(uint32_t n, const cref_evolve_state* states, uint32_t bucket,
                        uint32_t max_depth, cref_tree_meta* meta, cref_gpu_body* bodies,
                        cref_gpu_entry* entries, uint32_t padded, cref_gpu_cell* cells,
                        uint32_t capacity) {
    if (n < 2 || padded < n || (padded & (padded - 1)) != 0 || capacity < 1) {
        set_error("invalid BH #2C tree sizes");
        return -1;
    }
    /* bounds */
    float min_x = states[0].position_mass[0], max_x = min_x;
    float min_y = states[0].position_mass[1], max_y = min_y;
    for (uint32_t i = 1; i < n; ++i) {
        float x = states[i].position_mass[0], y = states[i].position_mass[1];
        min_x = fminf(min_x, x);
        max_x = fmaxf(max_x, x);
        min_y = fminf(min_y, y);
        max_y = fmaxf(max_y, y);
    }
    float span = fmaxf(fmaxf(max_x - min_x, max_y - min_y), 1.0e-12f);
    memset(meta, 0, sizeof *meta);
    meta->bounds[0] = 0.5f * (min_x + max_x);
    meta->bounds[1] = 0.5f * (min_y + max_y);
    meta->bounds[2] = 0.5f * span + 1.0e-12f;
    meta->bounds[3] = 0.0f;
    /* Morton entries and body copies */
    for (uint32_t i = 0; i < padded; ++i) {
        memset(&entries[i], 0, sizeof entries[i]);
        if (i >= n) {
            entries[i].data[0] = 0xffffffffu;
            entries[i].data[1] = 0xffffffffu;
            continue;
        }
        const float* pm = states[i].position_mass;
        uint32_t qx = (uint32_t)quantize_f32(pm[0], meta->bounds[0], meta->bounds[2]);
        uint32_t qy = (uint32_t)quantize_f32(pm[1], meta->bounds[1], meta->bounds[2]);
        entries[i].data[0] = morton2(qx, qy);
        entries[i].data[1] = i;
        memcpy(bodies[i].position_mass, pm, sizeof bodies[i].position_mass);
        memset(bodies[i].meta, 0, sizeof bodies[i].meta);
    }
    /* one-invocation bitonic sort */
    for (uint32_t k = 2; k <= padded; k <<= 1) {
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t i = 0; i < padded; ++i) {
                uint32_t ixj = i ^ j;
                if (ixj > i && ixj < padded) {
                    int ascending = (i & k) == 0;
                    cref_gpu_entry a = entries[i], b = entries[ixj];
                    int swap = ascending ? entry_after(&a, &b) : entry_before(&a, &b);
                    if (swap) {
                        entries[i] = b;
                        entries[ixj] = a;
                    }
                }
            }
        }
    }
    for (uint32_t p = 0; p < n; ++p) bodies[entries[p].data[1]].meta[0] = p;
    /* topology */
    uint32_t md = max_depth < 16u ? max_depth : 16u;
    cells[0] = empty_cell_f32(meta->bounds[0], meta->bounds[1], meta->bounds[2], 0, n, 0);
    uint32_t next_free = 1, level_start = 0, level_end = 1, leaf_count = 0, max_seen = 0;
    while (level_start < level_end) {
        for (uint32_t pi = level_start; pi < level_end; ++pi) {
            cref_gpu_cell parent = cells[pi];
            uint32_t start = parent.range_depth[0], end = parent.range_depth[1];
            uint32_t depth = parent.range_depth[2];
            if (depth > max_seen) max_seen = depth;
            if (end - start <= bucket || depth >= md) {
                leaf_count += 1;
                continue;
            }
            uint32_t shift = 30u - 2u * depth;
            uint32_t cursor = start;
            while (cursor < end) {
                uint32_t quadrant = (entries[cursor].data[0] >> shift) & 3u;
                uint32_t group_start = cursor;
                cursor += 1;
                while (cursor < end && ((entries[cursor].data[0] >> shift) & 3u) == quadrant) cursor += 1;
                if (next_free >= capacity) {
                    meta->data[0] = next_free;
                    meta->data[1] = leaf_count;
                    meta->data[2] = max_seen;
                    meta->data[3] = 1;
                    return 0;
                }
                float h = 0.5f * parent.center_half_mass[2];
                float ox = (quadrant & 1u) ? h : -h;
                float oy = (quadrant & 2u) ? h : -h;
                cells[next_free] = empty_cell_f32(parent.center_half_mass[0] + ox,
                                                  parent.center_half_mass[1] + oy, h, group_start,
                                                  cursor, depth + 1u);
                parent.children[quadrant] = next_free;
                next_free += 1;
            }
            cells[pi] = parent;
        }
        if (next_free == level_end) break;
        level_start = level_end;
        level_end = next_free;
    }
    meta->data[0] = next_free;
    meta->data[1] = leaf_count;
    meta->data[2] = max_seen;
    meta->data[3] = 0;
    /* aggregates, reverse cell order */
    for (uint32_t r = next_free; r > 0; --r) {
        cref_gpu_cell* cell = &cells[r - 1];
        int leaf = cell->children[0] == CREF_NO_CHILD && cell->children[1] == CREF_NO_CHILD &&
                   cell->children[2] == CREF_NO_CHILD && cell->children[3] == CREF_NO_CHILD;
        float mass = 0.0f, wx = 0.0f, wy = 0.0f;
        if (leaf) {
            for (uint32_t p = cell->range_depth[0]; p < cell->range_depth[1]; ++p) {
                const float* b = bodies[entries[p].data[1]].position_mass;
                mass = mass + b[2];
                wx = wx + b[2] * b[0];
                wy = wy + b[2] * b[1];
            }
        } else {
            for (int q = 0; q < 4; ++q) {
                uint32_t ci = cell->children[q];
                if (ci != CREF_NO_CHILD) {
                    float cm = cells[ci].center_half_mass[3];
                    mass = mass + cm;
                    wx = wx + cm * cells[ci].com[0];
                    wy = wy + cm * cells[ci].com[1];
                }
            }
        }
        cell->center_half_mass[3] = mass;
        cell->com[0] = wx / mass;
        cell->com[1] = wy / mass;
        cell->com[2] = 0.0f;
        cell->com[3] = 0.0f;
    }
    return 0;
}
