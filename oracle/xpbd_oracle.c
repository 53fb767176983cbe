/*
 * CPU ORACLE for the XPBD cloth substep -- TEST INFRASTRUCTURE ONLY.
 *
 * This file restates, in plain scalar C, the algorithm of the reference
 * cloth_core substep so that the HIP product can be checked against it.
 * It is never linked into the product library; only tests/, the smoke check
 * and the CPU-baseline leg of bench.py load it (through oracle/oracle.py).
 *
 * It is compiled twice (oracle/Makefile):
 *   liborc_f32.so  with ORC_REAL=float   -- same arithmetic type as the
 *                                           reference shaders (fp32)
 *   liborc_f64.so  with ORC_REAL=double  -- the FP64 twin used to state the
 *                                           numerical error of fp32 results
 * The host-side mesh build (constraint generation, colouring, topology) is
 * always fp32/integer, as the reference computes it on the host in f32.
 *
 * Reference functions restated here (paths relative to the reference root):
 *   mesh build ............ crates/cloth_core/src/mesh.rs:407-1377
 *   inverse masses ........ crates/cloth_core/src/mesh.rs:237-312
 *   Welsh-Powell colouring  crates/cloth_core/src/coloring.rs:6-173
 *   virtual sample count .. crates/cloth_core/src/virtual_mesh.rs:75-200
 *   hash cell/table rule .. crates/cloth_core/src/simulation/pipelines.rs:899-912
 *   substep loop .......... crates/cloth_core/src/simulation/dispatch.rs:7-354
 *   predict ............... crates/cloth_core/src/shaders/predict.wgsl:33-55
 *   distance .............. crates/cloth_core/src/shaders/distance.wgsl:60-124
 *   bending ............... crates/cloth_core/src/shaders/bending.wgsl:59-126
 *   collider resolve ...... crates/cloth_core/src/shaders/collision.wgsl:83-108, 273-652
 *   spatial hash .......... crates/cloth_core/src/shaders/spatial_grid_sort.wgsl:25-195
 *   edge hash ............. crates/cloth_core/src/shaders/spatial_grid_sort_edges.wgsl:51-199
 *   self collision V-V/V-T  crates/cloth_core/src/shaders/self_collision.wgsl:71-681
 *   self collision E-E .... crates/cloth_core/src/shaders/self_collision_ee_direct.wgsl:66-776
 *   self collision apply .. crates/cloth_core/src/shaders/self_collision_apply.wgsl:57-177
 *   update_vel ............ crates/cloth_core/src/shaders/update_vel.wgsl:33-53
 *
 * Where the parity of each function is pinned (golden fixtures, closed
 * forms, brute force) is listed in the test files under tests/.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* Built with -fopenmp (liborc_f32_omp.so, the bench's multithreaded CPU baseline) the per-vertex
 * passes and each colour's constraints are split over threads.  Constraints of one colour share no
 * vertex and the per-vertex passes write only their own vertex, so every thread count gives the
 * serial result bit for bit.  Self-collision stays serial (shared counters). */
#ifdef _OPENMP
#define ORC_PARALLEL_FOR _Pragma("omp parallel for schedule(static)")
#else
#define ORC_PARALLEL_FOR
#endif

#ifndef ORC_REAL
#define ORC_REAL float
#endif
typedef ORC_REAL real;

/* Correctly rounded sqrt/floor in the working type.  For float, sqrt in
 * double then rounding to float is correctly rounded (53 >= 2*24+2). */
#define RSQRT(x) ((real)sqrt((double)(x)))
#define RFLOOR(x) ((real)floor((double)(x)))

typedef struct { real x, y, z; } v3;

static inline v3 mk(real x, real y, real z) { v3 r = {x, y, z}; return r; }
static inline v3 add(v3 a, v3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 sub(v3 a, v3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 mul(v3 a, real s) { return mk(a.x * s, a.y * s, a.z * s); }
static inline v3 dvs(v3 a, real s) { return mk(a.x / s, a.y / s, a.z / s); }
static inline v3 neg(v3 a) { return mk(-a.x, -a.y, -a.z); }
static inline real dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline v3 cross(v3 a, v3 b) {
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline real len3(v3 a) { return RSQRT(dot(a, a)); }
static inline real rmax(real a, real b) { return a > b ? a : b; }
static inline real rmin(real a, real b) { return a < b ? a : b; }
static inline real rclamp(real x, real lo, real hi) { return rmin(rmax(x, lo), hi); }
/* WGSL mix(a, b, t) = a * (1 - t) + b * t */
static inline real rmix(real a, real b, real t) { return a * ((real)1 - t) + b * t; }
static inline v3 vmix(v3 a, v3 b, real t) {
    return mk(rmix(a.x, b.x, t), rmix(a.y, b.y, t), rmix(a.z, b.z, t));
}
static inline v3 vmin(v3 a, v3 b) { return mk(rmin(a.x, b.x), rmin(a.y, b.y), rmin(a.z, b.z)); }
static inline v3 vmax(v3 a, v3 b) { return mk(rmax(a.x, b.x), rmax(a.y, b.y), rmax(a.z, b.z)); }

/* ------------------------------------------------------------------ */
/* small growable arrays                                               */
/* ------------------------------------------------------------------ */
typedef struct { uint32_t *d; size_t n, cap; } vu32;
typedef struct { float *d; size_t n, cap; } vf32;
static void vu_push(vu32 *v, uint32_t x) {
    if (v->n == v->cap) { v->cap = v->cap ? v->cap * 2 : 16; v->d = (uint32_t *)realloc(v->d, v->cap * 4); }
    v->d[v->n++] = x;
}
static void vf_push(vf32 *v, float x) {
    if (v->n == v->cap) { v->cap = v->cap ? v->cap * 2 : 16; v->d = (float *)realloc(v->d, v->cap * 4); }
    v->d[v->n++] = x;
}

/* open-addressing set of (u32,u32) keys */
typedef struct { uint64_t *k; size_t cap, n; } pset;
#define PSET_EMPTY 0xFFFFFFFFFFFFFFFFull
static void pset_init(pset *s, size_t expect) {
    size_t cap = 64;
    while (cap < expect * 2 + 8) cap <<= 1;
    s->cap = cap; s->n = 0;
    s->k = (uint64_t *)malloc(cap * 8);
    for (size_t i = 0; i < cap; i++) s->k[i] = PSET_EMPTY;
}
static uint64_t mix64(uint64_t x) {
    x ^= x >> 33; x *= 0xff51afd7ed558ccdull; x ^= x >> 33; x *= 0xc4ceb9fe1a85ec53ull; x ^= x >> 33;
    return x;
}
static int pset_insert(pset *s, uint32_t a, uint32_t b); /* returns 1 if newly inserted */
static void pset_grow(pset *s) {
    pset t; pset_init(&t, s->cap);
    for (size_t i = 0; i < s->cap; i++)
        if (s->k[i] != PSET_EMPTY) pset_insert(&t, (uint32_t)(s->k[i] >> 32), (uint32_t)s->k[i]);
    free(s->k); *s = t;
}
static int pset_insert(pset *s, uint32_t a, uint32_t b) {
    if ((s->n + 1) * 2 > s->cap) pset_grow(s);
    uint64_t key = ((uint64_t)a << 32) | b;
    size_t m = s->cap - 1, i = (size_t)mix64(key) & m;
    while (s->k[i] != PSET_EMPTY) {
        if (s->k[i] == key) return 0;
        i = (i + 1) & m;
    }
    s->k[i] = key; s->n++;
    return 1;
}
static int pset_has(const pset *s, uint32_t a, uint32_t b) {
    uint64_t key = ((uint64_t)a << 32) | b;
    size_t m = s->cap - 1, i = (size_t)mix64(key) & m;
    while (s->k[i] != PSET_EMPTY) {
        if (s->k[i] == key) return 1;
        i = (i + 1) & m;
    }
    return 0;
}
static void pset_free(pset *s) { free(s->k); s->k = NULL; }

static int cmpu64(const void *a, const void *b) {
    uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
    return x < y ? -1 : x > y;
}
static int cmp_u32(const void *a, const void *b) {
    uint32_t x = *(const uint32_t *)a, y = *(const uint32_t *)b;
    return x < y ? -1 : x > y;
}
/* sort + dedup, returns new length */
static size_t sort_unique(uint32_t *a, size_t n) {
    if (!n) return 0;
    qsort(a, n, 4, cmp_u32);
    size_t w = 1;
    for (size_t i = 1; i < n; i++) if (a[i] != a[w - 1]) a[w++] = a[i];
    return w;
}
static int bsearch_u32(const uint32_t *a, size_t n, uint32_t x) {
    size_t lo = 0, hi = n;
    while (lo < hi) { size_t mid = lo + (hi - lo) / 2; if (a[mid] == x) return 1; if (a[mid] < x) lo = mid + 1; else hi = mid; }
    return 0;
}

static float fdist(const float *a, const float *b) { /* (b - a) as in virtual_mesh.rs:202-207 */
    float dx = b[0] - a[0], dy = b[1] - a[1], dz = b[2] - a[2];
    return sqrtf(dx * dx + dy * dy + dz * dz);
}
static float frest(const float *p0, const float *p1) { /* (p0 - p1) as in mesh.rs:480-485 */
    float dx = p0[0] - p1[0], dy = p0[1] - p1[1], dz = p0[2] - p1[2];
    return sqrtf(dx * dx + dy * dy + dz * dz);
}

static float compliance_of(float k) { /* mesh.rs:450-472 */
    if (k >= 5000.0f) return 0.0f;
    if (k > 0.0f) return 1.0f / k;
    return 1e10f;
}

/* ------------------------------------------------------------------ */
/* mesh                                                                */
/* ------------------------------------------------------------------ */
typedef struct {
    uint32_t nv;
    float *pos0;      /* 3*nv initial positions */
    float *inv_mass;  /* nv */
    float *thick;     /* nv */
    uint32_t *layer;  /* nv */

    uint32_t nd;      /* distance constraints, colour-sorted */
    uint32_t *d_v0, *d_v1, *d_type;
    float *d_rest, *d_tc, *d_cc;
    uint32_t nd_col, *d_col_off, *d_col_cnt;
    uint32_t *d_remap;          /* generated index -> sorted index */
    uint32_t ne_in, *edge_to_constraint; /* input edge -> sorted distance index (0xFFFFFFFF invalid) */

    uint32_t nb;      /* bending constraints, colour-sorted */
    uint32_t *b_v0, *b_v1, *b_v2, *b_v3;
    float *b_rest, *b_comp, *b_pad0, *b_pad1;
    uint32_t nb_col, *b_col_off, *b_col_cnt;

    uint32_t ns;      /* sewing constraints, colour-sorted (mesh.rs:909-947) */
    uint32_t *s_v0, *s_v1;
    float *s_rest;    /* initial (current) rest lengths; the target is 0 */
    uint32_t ns_col, *s_col_off, *s_col_cnt;
    float sew_speed;

    uint32_t *adj_off, *adj_idx;
    uint32_t *th_off, *th_idx; float *th_rest;
    float *local_len;
    uint32_t *star_off, *star_v0, *star_v1;
    uint32_t ntri, *tri;
    uint32_t ne, *e_v0, *e_v1, *e_bbits;  /* Morton-sorted unique edges */
    uint32_t *island;
    float mean_edge, max_edge;
    uint32_t n_virtual;
    float cell_size;
    uint32_t table_size;
    int has_multiple_layers;
} orc_mesh;

/* Welsh-Powell, coloring.rs:6-79 (distance) and 82-173 (bending).
 * cverts: K vertex slots per constraint (0xFFFFFFFF = unused slot).
 * reg: which slots are registered in vertex->constraint lists;
 * deg/used use the same registered slots (see the bending special cases). */
typedef struct { uint32_t deg; uint32_t idx; } wp_key;
static int cmp_wp(const void *a, const void *b) {
    const wp_key *x = (const wp_key *)a, *y = (const wp_key *)b;
    if (x->deg != y->deg) return x->deg > y->deg ? -1 : 1; /* Reverse(deg) */
    return x->idx < y->idx ? -1 : x->idx > y->idx;
}

static void welsh_powell(uint32_t nv, uint32_t nc,
                         /* registration lists: for each constraint, verts to register */
                         const uint32_t *reg_v, const uint32_t *reg_n, uint32_t reg_stride,
                         /* degree verts */
                         const uint32_t *deg_v, const uint32_t *deg_n, uint32_t deg_stride,
                         /* verts whose lists give used colours */
                         const uint32_t *use_v, const uint32_t *use_n, uint32_t use_stride,
                         uint32_t *colors_out, uint32_t *ncolors_out) {
    uint32_t *cnt = (uint32_t *)calloc(nv + 1, 4);
    for (uint32_t i = 0; i < nc; i++)
        for (uint32_t s = 0; s < reg_n[i]; s++) cnt[reg_v[i * reg_stride + s]]++;
    uint32_t *off = (uint32_t *)malloc((nv + 1) * 4);
    off[0] = 0;
    for (uint32_t v = 0; v < nv; v++) off[v + 1] = off[v] + cnt[v];
    uint32_t *lst = (uint32_t *)malloc((size_t)(off[nv] ? off[nv] : 1) * 4);
    uint32_t *fill = (uint32_t *)calloc(nv + 1, 4);
    for (uint32_t i = 0; i < nc; i++)
        for (uint32_t s = 0; s < reg_n[i]; s++) {
            uint32_t v = reg_v[i * reg_stride + s];
            lst[off[v] + fill[v]++] = i;
        }
    wp_key *ord = (wp_key *)malloc((size_t)(nc ? nc : 1) * sizeof(wp_key));
    for (uint32_t i = 0; i < nc; i++) {
        uint32_t d = 0;
        for (uint32_t s = 0; s < deg_n[i]; s++) d += cnt[deg_v[i * deg_stride + s]];
        ord[i].deg = d; ord[i].idx = i;
    }
    qsort(ord, nc, sizeof(wp_key), cmp_wp);
    for (uint32_t i = 0; i < nc; i++) colors_out[i] = 0xFFFFFFFFu;
    uint32_t ncol = 0;
    size_t stamp_cap = 64;
    uint32_t *stamp = (uint32_t *)calloc(stamp_cap, 4);
    for (uint32_t o = 0; o < nc; o++) {
        uint32_t i = ord[o].idx;
        for (uint32_t s = 0; s < use_n[i]; s++) {
            uint32_t v = use_v[i * use_stride + s];
            for (uint32_t k = off[v]; k < off[v + 1]; k++) {
                uint32_t a = lst[k];
                if (a != i && colors_out[a] != 0xFFFFFFFFu) {
                    uint32_t c = colors_out[a];
                    if (c >= stamp_cap) {
                        size_t nc2 = stamp_cap; while (nc2 <= c) nc2 *= 2;
                        stamp = (uint32_t *)realloc(stamp, nc2 * 4);
                        memset(stamp + stamp_cap, 0, (nc2 - stamp_cap) * 4);
                        stamp_cap = nc2;
                    }
                    stamp[c] = o + 1;
                }
            }
        }
        uint32_t c = 0;
        while (c < stamp_cap && stamp[c] == o + 1) c++;
        colors_out[i] = c;
        if (c + 1 > ncol) ncol = c + 1;
    }
    *ncolors_out = ncol;
    free(stamp); free(ord); free(fill); free(lst); free(off); free(cnt);
}

/* colour-bucket reorder (stable within colour) */
static void bucket_order(uint32_t nc, const uint32_t *col, uint32_t ncol,
                         uint32_t *perm_sorted_to_orig, uint32_t *remap_orig_to_sorted,
                         uint32_t *col_off, uint32_t *col_cnt) {
    for (uint32_t c = 0; c < ncol; c++) col_cnt[c] = 0;
    for (uint32_t i = 0; i < nc; i++) col_cnt[col[i]]++;
    uint32_t acc = 0;
    for (uint32_t c = 0; c < ncol; c++) { col_off[c] = acc; acc += col_cnt[c]; }
    uint32_t *fill = (uint32_t *)calloc(ncol ? ncol : 1, 4);
    for (uint32_t i = 0; i < nc; i++) {
        uint32_t s = col_off[col[i]] + fill[col[i]]++;
        perm_sorted_to_orig[s] = i;
        if (remap_orig_to_sorted) remap_orig_to_sorted[i] = s;
    }
    free(fill);
}

typedef struct { uint64_t key; uint32_t opp; uint32_t seq; } eopp;
static int cmp_eopp(const void *a, const void *b) {
    const eopp *x = (const eopp *)a, *y = (const eopp *)b;
    if (x->key != y->key) return x->key < y->key ? -1 : 1;
    return x->seq < y->seq ? -1 : x->seq > y->seq;
}

static uint32_t expand_bits_10(uint32_t v) { /* mesh.rs:314-320 */
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}
static uint32_t f2u_sat(float x) { /* Rust `as u32` (saturating, NaN -> 0) */
    if (!(x > 0.0f)) return 0;
    if (x >= 4294967296.0f) return 0xFFFFFFFFu;
    return (uint32_t)x;
}
static uint32_t morton_code_3d(const float *p, const float *bmin, const float *inv) { /* mesh.rs:322-327 */
    uint32_t x = f2u_sat(fminf(fmaxf((p[0] - bmin[0]) * inv[0], 0.0f), 1.0f) * 1023.0f);
    uint32_t y = f2u_sat(fminf(fmaxf((p[1] - bmin[1]) * inv[1], 0.0f), 1.0f) * 1023.0f);
    uint32_t z = f2u_sat(fminf(fmaxf((p[2] - bmin[2]) * inv[2], 0.0f), 1.0f) * 1023.0f);
    return (expand_bits_10(x) << 2) | (expand_bits_10(y) << 1) | expand_bits_10(z);
}
typedef struct { uint32_t code; uint32_t idx; } mkey;
static int cmp_mkey(const void *a, const void *b) {
    const mkey *x = (const mkey *)a, *y = (const mkey *)b;
    if (x->code != y->code) return x->code < y->code ? -1 : 1;
    return x->idx < y->idx ? -1 : x->idx > y->idx; /* sort_by_cached_key is stable */
}

static uint32_t next_pow2(uint32_t x) { uint32_t p = 1; while (p < x) p <<= 1; return p; }

/* virtual_mesh.rs:75-200 -- only the number of samples is needed (it sizes the hash table) */
static uint32_t virtual_sample_count(const orc_mesh *m, const float *pos) {
    if (m->nv == 0 || m->ntri == 0) return 0;
    float thickness = m->thick[0];
    float spacing = thickness * 1.5f;
    if (spacing <= 1e-6f) return 0;
    /* edge -> first face containing it */
    pset seen; pset_init(&seen, (size_t)m->ntri * 3);
    for (uint32_t f = 0; f < m->ntri; f++) {
        uint32_t a = m->tri[3 * f], b = m->tri[3 * f + 1], c = m->tri[3 * f + 2];
        pset_insert(&seen, a < b ? a : b, a < b ? b : a);
        pset_insert(&seen, b < c ? b : c, b < c ? c : b);
        pset_insert(&seen, c < a ? c : a, c < a ? a : c);
    }
    uint64_t total = 0;
    const uint64_t CAP = 150000;
    for (uint32_t k = 0; k < m->nd; k++) {
        uint32_t u = m->d_v0[k], v = m->d_v1[k];
        if (!pset_has(&seen, u < v ? u : v, u < v ? v : u)) continue;
        float len = fdist(pos + 3 * u, pos + 3 * v);
        uint32_t ke = (uint32_t)ceilf(len / spacing);
        if (ke > 16) ke = 16;
        if (ke > 1) total += ke - 1;
    }
    for (uint32_t f = 0; f < m->ntri; f++) {
        const float *pa = pos + 3 * m->tri[3 * f], *pb = pos + 3 * m->tri[3 * f + 1], *pc = pos + 3 * m->tri[3 * f + 2];
        float lab = fdist(pa, pb), lbc = fdist(pb, pc), lca = fdist(pc, pa);
        float lmax = fmaxf(fmaxf(lab, lbc), lca);
        uint32_t k = (uint32_t)ceilf(lmax / spacing);
        if (k > 16) k = 16;
        if (k >= 3) total += (uint64_t)(k - 1) * (k - 2) / 2;
    }
    pset_free(&seen);
    return (uint32_t)(total < CAP ? total : CAP);
}

/* Sewing groups: union-find over the valid sewing pairs, each vertex labelled by the smallest
 * vertex of its group (mesh.rs:709-739 and 998-1030 build the same partition). */
static void sew_reps(uint32_t nv, const uint32_t *sew, uint32_t nsew, uint32_t *rep) {
    for (uint32_t i = 0; i < nv; i++) rep[i] = i;
    for (uint32_t k = 0; k < nsew; k++) {
        uint32_t u = sew[2 * k], v = sew[2 * k + 1];
        if (u >= nv || v >= nv || u == v) continue;
        while (rep[u] != u) u = rep[u];
        while (rep[v] != v) v = rep[v];
        if (u != v) rep[u > v ? u : v] = u < v ? u : v;
    }
    for (uint32_t i = 0; i < nv; i++) {
        uint32_t r = i;
        while (rep[r] != r) r = rep[r];
        rep[i] = r;
    }
}

/* Unfolded distance of q (over edge e0-e1) and r (over f0-f1) about a shared seam axis:
 * sqrt((s_q - s_r)^2 + (h_q + h_r)^2) with s the along-edge and h the perpendicular offset from
 * each edge's first end; -1 when either edge is not longer than 1e-6 (mesh.rs:789-822, 1436-1458). */
static float seam_unfold(const float *pos, uint32_t e0, uint32_t e1, uint32_t q, uint32_t f0, uint32_t f1, uint32_t r) {
    const float *A0 = pos + 3 * e0, *A1 = pos + 3 * e1, *B0 = pos + 3 * f0, *B1 = pos + 3 * f1;
    const float *Q = pos + 3 * q, *R = pos + 3 * r;
    float ea[3] = {A1[0] - A0[0], A1[1] - A0[1], A1[2] - A0[2]}, eb[3] = {B1[0] - B0[0], B1[1] - B0[1], B1[2] - B0[2]};
    float la = sqrtf(ea[0] * ea[0] + ea[1] * ea[1] + ea[2] * ea[2]);
    float lb = sqrtf(eb[0] * eb[0] + eb[1] * eb[1] + eb[2] * eb[2]);
    if (!(la > 1e-6f && lb > 1e-6f)) return -1.0f;
    float ua[3] = {ea[0] / la, ea[1] / la, ea[2] / la}, ub[3] = {eb[0] / lb, eb[1] / lb, eb[2] / lb};
    float da[3] = {Q[0] - A0[0], Q[1] - A0[1], Q[2] - A0[2]}, db[3] = {R[0] - B0[0], R[1] - B0[1], R[2] - B0[2]};
    float sa = da[0] * ua[0] + da[1] * ua[1] + da[2] * ua[2];
    float sb = db[0] * ub[0] + db[1] * ub[1] + db[2] * ub[2];
    float pa[3] = {da[0] - sa * ua[0], da[1] - sa * ua[1], da[2] - sa * ua[2]};
    float pb[3] = {db[0] - sb * ub[0], db[1] - sb * ub[1], db[2] - sb * ub[2]};
    float ha = sqrtf(pa[0] * pa[0] + pa[1] * pa[1] + pa[2] * pa[2]);
    float hb = sqrtf(pb[0] * pb[0] + pb[1] * pb[1] + pb[2] * pb[2]);
    float ds = sa - sb, dh = ha + hb;
    return sqrtf(ds * ds + dh * dh);
}

/* Rest length of a two-hop pair joined only through sewing (mesh.rs:1382-1490): the smallest of,
 * over original neighbours m1 of i and m2 of nbr in one sewing group (not i's or nbr's group), the
 * direct distance when m1 == m2, the unfolded distance over each seam edge m1-k1 / m2-k2 whose far
 * ends share a group, or |i - m1| + |nbr - m2| when no seam edge gives a smaller value; 0 unless
 * above 1e-5.  orig_off / orig: the original adjacency. */
static float seam_two_hop(uint32_t i, uint32_t nbr, const float *pos, const uint32_t *orig_off, const uint32_t *orig,
                          const uint32_t *rep) {
    uint32_t ri = rep[i], rn = rep[nbr];
    float best = 3.4028235e38f;
    for (uint32_t a = orig_off[i]; a < orig_off[i + 1]; a++) {
        uint32_t m1 = orig[a], rm = rep[m1];
        if (rm == ri || rm == rn) continue;
        for (uint32_t b = orig_off[nbr]; b < orig_off[nbr + 1]; b++) {
            uint32_t m2 = orig[b];
            if (rep[m2] != rm) continue;
            if (m1 == m2) {
                float d = frest(pos + 3 * i, pos + 3 * nbr);
                if (d < best) best = d;
                continue;
            }
            int found = 0;
            for (uint32_t c = orig_off[m1]; c < orig_off[m1 + 1]; c++) {
                uint32_t k1 = orig[c], rk = rep[k1];
                if (rk == rm) continue;
                for (uint32_t d = orig_off[m2]; d < orig_off[m2 + 1]; d++) {
                    uint32_t k2 = orig[d];
                    if (rep[k2] != rk) continue;
                    float fl = seam_unfold(pos, m1, k1, i, m2, k2, nbr);
                    if (fl >= 0.0f && fl < best) { best = fl; found = 1; }
                }
            }
            if (!found) {
                float slit = frest(pos + 3 * i, pos + 3 * m1) + frest(pos + 3 * nbr, pos + 3 * m2);
                if (slit < best) best = slit;
            }
        }
    }
    return (best < 3.4028235e38f && isfinite(best) && best > 1e-5f) ? best : 0.0f;
}

typedef struct { uint64_t ckey; uint32_t u, v, opp, ord; } seament;
static int cmp_seament(const void *a, const void *b) {
    const seament *x = (const seament *)a, *y = (const seament *)b;
    if (x->ckey != y->ckey) return x->ckey < y->ckey ? -1 : 1;
    return x->ord < y->ord ? -1 : x->ord > y->ord;
}

orc_mesh *orc_mesh_build_sewn(const float *pos, uint32_t nv, const uint32_t *edges, uint32_t ne_in,
                              const uint32_t *faces, uint32_t nf, const float *inv_masses,
                              const float *thicknesses, const uint32_t *layer_ids, uint32_t default_layer,
                              float default_thickness, float tension_k, float compression_k, float shear_k,
                              float bending_k, const uint32_t *sew, uint32_t nsew, float sew_speed);
orc_mesh *orc_mesh_build(const float *pos, uint32_t nv, const uint32_t *edges, uint32_t ne_in,
                         const uint32_t *faces, uint32_t nf, const float *inv_masses,
                         const float *thicknesses, const uint32_t *layer_ids, uint32_t default_layer,
                         float default_thickness, float tension_k, float compression_k, float shear_k,
                         float bending_k) {
    return orc_mesh_build_sewn(pos, nv, edges, ne_in, faces, nf, inv_masses, thicknesses, layer_ids, default_layer,
                               default_thickness, tension_k, compression_k, shear_k, bending_k, NULL, 0, 1.0f);
}

orc_mesh *orc_mesh_build_sewn(const float *pos, uint32_t nv, const uint32_t *edges, uint32_t ne_in,
                              const uint32_t *faces, uint32_t nf, const float *inv_masses,
                              const float *thicknesses, const uint32_t *layer_ids, uint32_t default_layer,
                              float default_thickness, float tension_k, float compression_k, float shear_k,
                              float bending_k, const uint32_t *sew, uint32_t nsew, float sew_speed) {
    const int sewing = sew && nsew;
    uint32_t *rep = (uint32_t *)malloc((size_t)(nv ? nv : 1) * 4);
    sew_reps(nv, sewing ? sew : NULL, sewing ? nsew : 0, rep);
    orc_mesh *m = (orc_mesh *)calloc(1, sizeof(orc_mesh));
    m->nv = nv;
    m->pos0 = (float *)malloc((size_t)(nv ? nv : 1) * 12);
    memcpy(m->pos0, pos, (size_t)nv * 12);
    m->inv_mass = (float *)malloc((size_t)(nv ? nv : 1) * 4);
    m->thick = (float *)malloc((size_t)(nv ? nv : 1) * 4);
    m->layer = (uint32_t *)malloc((size_t)(nv ? nv : 1) * 4);
    for (uint32_t i = 0; i < nv; i++) {
        m->inv_mass[i] = inv_masses ? inv_masses[i] : 1.0f;
        m->thick[i] = thicknesses ? thicknesses[i] : default_thickness;
        m->layer[i] = layer_ids ? layer_ids[i] : default_layer;
    }
    float tc = compliance_of(tension_k), cc = compliance_of(compression_k), sc = compliance_of(shear_k);

    /* 1. stretch (mesh.rs:474-499) */
    vu32 dv0 = {0}, dv1 = {0}, dty = {0};
    vf32 drest = {0}, dtc = {0}, dcc = {0};
    uint32_t *raw_edge_to_dist = (uint32_t *)malloc((size_t)(ne_in ? ne_in : 1) * 4);
    pset original; pset_init(&original, ne_in);
    for (uint32_t e = 0; e < ne_in; e++) {
        uint32_t a = edges[2 * e], b = edges[2 * e + 1];
        if (a < nv && b < nv && a != b) {
            raw_edge_to_dist[e] = (uint32_t)dv0.n;
            vu_push(&dv0, a); vu_push(&dv1, b); vu_push(&dty, 0);
            vf_push(&drest, frest(pos + 3 * a, pos + 3 * b)); vf_push(&dtc, tc); vf_push(&dcc, cc);
            pset_insert(&original, a < b ? a : b, a < b ? b : a);
        } else {
            raw_edge_to_dist[e] = 0xFFFFFFFFu;
        }
    }

    /* edge -> opposite vertices (face order), mesh.rs:514-527 */
    eopp *eo = NULL; size_t neo = 0;
    if (faces && nf) {
        eo = (eopp *)malloc((size_t)nf * 3 * sizeof(eopp));
        uint32_t seq = 0;
        for (uint32_t f = 0; f < nf; f++) {
            uint32_t f0 = faces[3 * f], f1 = faces[3 * f + 1], f2 = faces[3 * f + 2];
            if (f0 >= nv || f1 >= nv || f2 >= nv) continue;
            uint32_t es[3][3] = {{f0 < f1 ? f0 : f1, f0 < f1 ? f1 : f0, f2},
                                 {f1 < f2 ? f1 : f2, f1 < f2 ? f2 : f1, f0},
                                 {f2 < f0 ? f2 : f0, f2 < f0 ? f0 : f2, f1}};
            for (int k = 0; k < 3; k++) {
                eo[neo].key = ((uint64_t)es[k][0] << 32) | es[k][1];
                eo[neo].opp = es[k][2];
                eo[neo].seq = seq++;
                neo++;
            }
        }
        qsort(eo, neo, sizeof(eopp), cmp_eopp);
    }

    /* 2. shear (mesh.rs:508-560) */
    if (neo) {
        pset existing; pset_init(&existing, ne_in + neo);
        for (uint32_t e = 0; e < ne_in; e++) {
            uint32_t a = edges[2 * e], b = edges[2 * e + 1];
            if (a < nv && b < nv && a != b) pset_insert(&existing, a < b ? a : b, a < b ? b : a);
        }
        for (size_t g = 0; g < neo;) {
            size_t h = g;
            while (h < neo && eo[h].key == eo[g].key) h++;
            uint32_t ka = (uint32_t)(eo[g].key >> 32), kb = (uint32_t)eo[g].key;
            if (!pset_has(&original, ka, kb) && h - g == 2) {
                uint32_t o0 = eo[g].opp, o1 = eo[g + 1].opp;
                uint32_t pairs[2][2] = {{ka, kb}, {o0 < o1 ? o0 : o1, o0 < o1 ? o1 : o0}};
                for (int k = 0; k < 2; k++) {
                    uint32_t a = pairs[k][0], b = pairs[k][1];
                    if (a != b && pset_insert(&existing, a, b)) {
                        vu_push(&dv0, a); vu_push(&dv1, b); vu_push(&dty, 1);
                        vf_push(&drest, frest(pos + 3 * a, pos + 3 * b)); vf_push(&dtc, sc); vf_push(&dcc, sc);
                    }
                }
            }
            g = h;
        }
        pset_free(&existing);
    }

    /* colour distance constraints (coloring.rs:6-79) */
    uint32_t nd = (uint32_t)dv0.n;
    m->nd = nd;
    {
        uint32_t *vv = (uint32_t *)malloc((size_t)(nd ? nd : 1) * 8);
        uint32_t *two = (uint32_t *)malloc((size_t)(nd ? nd : 1) * 4);
        for (uint32_t i = 0; i < nd; i++) { vv[2 * i] = dv0.d[i]; vv[2 * i + 1] = dv1.d[i]; two[i] = 2; }
        uint32_t *col = (uint32_t *)malloc((size_t)(nd ? nd : 1) * 4);
        uint32_t ncol = 0;
        welsh_powell(nv, nd, vv, two, 2, vv, two, 2, vv, two, 2, col, &ncol);
        m->nd_col = ncol;
        m->d_col_off = (uint32_t *)malloc((size_t)(ncol ? ncol : 1) * 4);
        m->d_col_cnt = (uint32_t *)malloc((size_t)(ncol ? ncol : 1) * 4);
        uint32_t *perm = (uint32_t *)malloc((size_t)(nd ? nd : 1) * 4);
        m->d_remap = (uint32_t *)malloc((size_t)(nd ? nd : 1) * 4);
        bucket_order(nd, col, ncol, perm, m->d_remap, m->d_col_off, m->d_col_cnt);
        m->d_v0 = (uint32_t *)malloc((size_t)(nd ? nd : 1) * 4);
        m->d_v1 = (uint32_t *)malloc((size_t)(nd ? nd : 1) * 4);
        m->d_type = (uint32_t *)malloc((size_t)(nd ? nd : 1) * 4);
        m->d_rest = (float *)malloc((size_t)(nd ? nd : 1) * 4);
        m->d_tc = (float *)malloc((size_t)(nd ? nd : 1) * 4);
        m->d_cc = (float *)malloc((size_t)(nd ? nd : 1) * 4);
        for (uint32_t s = 0; s < nd; s++) {
            uint32_t o = perm[s];
            m->d_v0[s] = dv0.d[o]; m->d_v1[s] = dv1.d[o]; m->d_type[s] = dty.d[o];
            m->d_rest[s] = drest.d[o]; m->d_tc[s] = dtc.d[o]; m->d_cc[s] = dcc.d[o];
        }
        free(perm); free(col); free(two); free(vv);
    }
    m->ne_in = ne_in;
    m->edge_to_constraint = (uint32_t *)malloc((size_t)(ne_in ? ne_in : 1) * 4);
    for (uint32_t e = 0; e < ne_in; e++)
        m->edge_to_constraint[e] = raw_edge_to_dist[e] == 0xFFFFFFFFu ? 0xFFFFFFFFu : m->d_remap[raw_edge_to_dist[e]];
    free(raw_edge_to_dist);

    /* 3. bending (mesh.rs:648-705), strand bending (mesh.rs:833-892) */
    vu32 bv0 = {0}, bv1 = {0}, bv2 = {0}, bv3 = {0};
    vf32 brest = {0}, bcomp = {0}, bpad0 = {0}, bpad1 = {0};
    float bend_c = bending_k > 0.0f ? 1.0f / bending_k : 1e10f;
    for (size_t g = 0; g < neo;) {
        size_t h = g;
        while (h < neo && eo[h].key == eo[g].key) h++;
        uint32_t ka = (uint32_t)(eo[g].key >> 32), kb = (uint32_t)eo[g].key;
        if (pset_has(&original, ka, kb) && h - g == 2 && eo[g].opp != eo[g + 1].opp) {
            uint32_t v2 = eo[g].opp, v3 = eo[g + 1].opp;
            vu_push(&bv0, ka); vu_push(&bv1, kb); vu_push(&bv2, v2); vu_push(&bv3, v3);
            vf_push(&brest, frest(pos + 3 * v2, pos + 3 * v3)); vf_push(&bcomp, bend_c);
            vf_push(&bpad0, 0.0f); vf_push(&bpad1, 0.0f);
        }
        g = h;
    }
    /* seam-gated bending (mesh.rs:706-831): boundary edges (one opposite vertex) with both ends
     * sewn, grouped by their pair of sewing groups; a pair of exactly two such edges bends about the
     * seam, at the unfolded distance of its opposite vertices, gated at half of it (>= 1 cm) */
    if (sewing && neo) {
        uint8_t *has_sew = (uint8_t *)calloc(nv + 1, 1);
        for (uint32_t k = 0; k < nsew; k++) {
            uint32_t u = sew[2 * k], v = sew[2 * k + 1];
            if (u < nv && v < nv && u != v) { has_sew[u] = 1; has_sew[v] = 1; }
        }
        seament *se = (seament *)malloc(neo * sizeof(seament));
        size_t nse = 0;
        for (size_t g = 0; g < neo;) {
            size_t h = g;
            while (h < neo && eo[h].key == eo[g].key) h++;
            uint32_t ea = (uint32_t)(eo[g].key >> 32), eb = (uint32_t)eo[g].key;
            if (h - g == 1 && has_sew[ea] && has_sew[eb] && rep[ea] != rep[eb]) {
                uint32_t ra = rep[ea], rb = rep[eb];
                seament x;
                x.ckey = ra < rb ? ((uint64_t)ra << 32) | rb : ((uint64_t)rb << 32) | ra;
                x.u = ra < rb ? ea : eb; x.v = ra < rb ? eb : ea; x.opp = eo[g].opp; x.ord = (uint32_t)nse;
                se[nse++] = x;
            }
            g = h;
        }
        qsort(se, nse, sizeof(seament), cmp_seament);
        for (size_t g = 0; g < nse;) {
            size_t h = g;
            while (h < nse && se[h].ckey == se[g].ckey) h++;
            if (h - g == 2) {
                uint32_t a0 = se[g].u, a1 = se[g].v, v2 = se[g].opp, b0 = se[g + 1].u, b1 = se[g + 1].v, v3 = se[g + 1].opp;
                if (!(a0 == b0 || a1 == b1 || v2 == v3 || rep[v2] == rep[v3])) {
                    float flat = seam_unfold(pos, a0, a1, v2, b0, b1, v3);
                    if (flat >= 0.0f && flat > 1e-5f) {
                        vu_push(&bv0, a0); vu_push(&bv1, b0); vu_push(&bv2, v2); vu_push(&bv3, v3);
                        vf_push(&brest, flat); vf_push(&bcomp, bend_c);
                        vf_push(&bpad0, 1.0f); vf_push(&bpad1, fmaxf(flat * 0.5f, 0.01f));
                    }
                }
            }
            g = h;
        }
        free(se); free(has_sew);
    }
    {
        pset face_edges; pset_init(&face_edges, (size_t)nf * 3 + 1);
        if (faces) for (uint32_t f = 0; f < nf; f++) {
            uint32_t f0 = faces[3 * f], f1 = faces[3 * f + 1], f2 = faces[3 * f + 2];
            pset_insert(&face_edges, f0 < f1 ? f0 : f1, f0 < f1 ? f1 : f0);
            pset_insert(&face_edges, f1 < f2 ? f1 : f2, f1 < f2 ? f2 : f1);
            pset_insert(&face_edges, f2 < f0 ? f2 : f0, f2 < f0 ? f0 : f2);
        }
        /* strand adjacency with insertion order */
        uint32_t *scnt = (uint32_t *)calloc(nv + 1, 4);
        for (uint32_t e = 0; e < ne_in; e++) {
            uint32_t a = edges[2 * e], b = edges[2 * e + 1];
            if (a < nv && b < nv && a != b && !pset_has(&face_edges, a < b ? a : b, a < b ? b : a)) { scnt[a]++; scnt[b]++; }
        }
        uint32_t *soff = (uint32_t *)malloc((nv + 1) * 4);
        soff[0] = 0;
        for (uint32_t v = 0; v < nv; v++) soff[v + 1] = soff[v] + scnt[v];
        uint32_t *sl = (uint32_t *)malloc((size_t)(soff[nv] ? soff[nv] : 1) * 4);
        uint32_t *sf = (uint32_t *)calloc(nv + 1, 4);
        for (uint32_t e = 0; e < ne_in; e++) {
            uint32_t a = edges[2 * e], b = edges[2 * e + 1];
            if (a < nv && b < nv && a != b && !pset_has(&face_edges, a < b ? a : b, a < b ? b : a)) {
                sl[soff[a] + sf[a]++] = b;
                sl[soff[b] + sf[b]++] = a;
            }
        }
        if (soff[nv] > 0) {
            pset bpairs; pset_init(&bpairs, bv0.n + soff[nv]);
            for (size_t i = 0; i < bv0.n; i++) {
                uint32_t a = bv2.d[i], b = bv3.d[i];
                pset_insert(&bpairs, a < b ? a : b, a < b ? b : a);
            }
            float strand_c = fmaxf(bend_c * 10.0f, 0.5f);
            for (uint32_t u = 0; u < nv; u++) {
                uint32_t n = scnt[u];
                if (n < 2) continue;
                for (uint32_t i = 0; i < n; i++)
                    for (uint32_t j = i + 1; j < n; j++) {
                        uint32_t a = sl[soff[u] + i], b = sl[soff[u] + j];
                        if (a != b && pset_insert(&bpairs, a < b ? a : b, a < b ? b : a)) {
                            vu_push(&bv0, u); vu_push(&bv1, u); vu_push(&bv2, a); vu_push(&bv3, b);
                            vf_push(&brest, frest(pos + 3 * a, pos + 3 * b)); vf_push(&bcomp, strand_c);
                            vf_push(&bpad0, 2.0f); vf_push(&bpad1, 0.0f);
                        }
                    }
            }
            pset_free(&bpairs);
        }
        free(sf); free(sl); free(soff); free(scnt);
        pset_free(&face_edges);
    }
    /* colour bending (coloring.rs:82-173) */
    uint32_t nb = (uint32_t)bv0.n;
    m->nb = nb;
    {
        uint32_t *reg = (uint32_t *)malloc((size_t)(nb ? nb : 1) * 16);
        uint32_t *regn = (uint32_t *)malloc((size_t)(nb ? nb : 1) * 4);
        uint32_t *dgv = (uint32_t *)malloc((size_t)(nb ? nb : 1) * 16);
        uint32_t *dgn = (uint32_t *)malloc((size_t)(nb ? nb : 1) * 4);
        uint32_t *usv = (uint32_t *)malloc((size_t)(nb ? nb : 1) * 16);
        uint32_t *usn = (uint32_t *)malloc((size_t)(nb ? nb : 1) * 4);
        for (uint32_t i = 0; i < nb; i++) {
            uint32_t v0 = bv0.d[i], v1 = bv1.d[i], v2 = bv2.d[i], v3 = bv3.d[i];
            int seam = bpad0.d[i] > 0.5f;
            uint32_t r = 0;
            reg[4 * i + r++] = v2; reg[4 * i + r++] = v3;
            if (seam) {
                if (v0 < nv && v0 != v2 && v0 != v3) reg[4 * i + r++] = v0;
                if (v1 < nv && v1 != v2 && v1 != v3 && v1 != v0) reg[4 * i + r++] = v1;
            }
            regn[i] = r;
            uint32_t d = 0;
            dgv[4 * i + d++] = v2; dgv[4 * i + d++] = v3;
            if (seam) {
                if (v0 < nv) dgv[4 * i + d++] = v0;
                if (v1 < nv) dgv[4 * i + d++] = v1;
            }
            dgn[i] = d;
            uint32_t u = 0;
            if (seam) {
                if (v0 < nv) usv[4 * i + u++] = v0;
                if (v1 < nv) usv[4 * i + u++] = v1;
            }
            if (v2 < nv) usv[4 * i + u++] = v2;
            if (v3 < nv) usv[4 * i + u++] = v3;
            usn[i] = u;
        }
        uint32_t *col = (uint32_t *)malloc((size_t)(nb ? nb : 1) * 4);
        uint32_t ncol = 0;
        welsh_powell(nv, nb, reg, regn, 4, dgv, dgn, 4, usv, usn, 4, col, &ncol);
        m->nb_col = ncol;
        m->b_col_off = (uint32_t *)malloc((size_t)(ncol ? ncol : 1) * 4);
        m->b_col_cnt = (uint32_t *)malloc((size_t)(ncol ? ncol : 1) * 4);
        uint32_t *perm = (uint32_t *)malloc((size_t)(nb ? nb : 1) * 4);
        bucket_order(nb, col, ncol, perm, NULL, m->b_col_off, m->b_col_cnt);
#define BALLOC(p, T) p = (T *)malloc((size_t)(nb ? nb : 1) * 4)
        BALLOC(m->b_v0, uint32_t); BALLOC(m->b_v1, uint32_t); BALLOC(m->b_v2, uint32_t); BALLOC(m->b_v3, uint32_t);
        BALLOC(m->b_rest, float); BALLOC(m->b_comp, float); BALLOC(m->b_pad0, float); BALLOC(m->b_pad1, float);
#undef BALLOC
        for (uint32_t s = 0; s < nb; s++) {
            uint32_t o = perm[s];
            m->b_v0[s] = bv0.d[o]; m->b_v1[s] = bv1.d[o]; m->b_v2[s] = bv2.d[o]; m->b_v3[s] = bv3.d[o];
            m->b_rest[s] = brest.d[o]; m->b_comp[s] = bcomp.d[o]; m->b_pad0[s] = bpad0.d[o]; m->b_pad1[s] = bpad1.d[o];
        }
        free(perm); free(col); free(reg); free(regn); free(dgv); free(dgn); free(usv); free(usn);
    }

    /* sewing constraints in input order (mesh.rs:909-944), coloured like the distance pairs
     * (coloring.rs:236-300) */
    m->sew_speed = sew_speed > 0.0f ? sew_speed : 0.0f;
    {
        vu32 s0 = {0}, s1 = {0}; vf32 sr = {0};
        if (sewing)
            for (uint32_t k = 0; k < nsew; k++) {
                uint32_t a = sew[2 * k], b = sew[2 * k + 1];
                if (a < nv && b < nv && a != b) { vu_push(&s0, a); vu_push(&s1, b); vf_push(&sr, frest(pos + 3 * a, pos + 3 * b)); }
            }
        uint32_t ns = (uint32_t)s0.n;
        m->ns = ns;
        uint32_t *vv = (uint32_t *)malloc((size_t)(ns ? ns : 1) * 8);
        uint32_t *two = (uint32_t *)malloc((size_t)(ns ? ns : 1) * 4);
        for (uint32_t i = 0; i < ns; i++) { vv[2 * i] = s0.d[i]; vv[2 * i + 1] = s1.d[i]; two[i] = 2; }
        uint32_t *col = (uint32_t *)malloc((size_t)(ns ? ns : 1) * 4);
        uint32_t ncol = 0;
        welsh_powell(nv, ns, vv, two, 2, vv, two, 2, vv, two, 2, col, &ncol);
        m->ns_col = ncol;
        m->s_col_off = (uint32_t *)malloc((size_t)(ncol ? ncol : 1) * 4);
        m->s_col_cnt = (uint32_t *)malloc((size_t)(ncol ? ncol : 1) * 4);
        uint32_t *perm = (uint32_t *)malloc((size_t)(ns ? ns : 1) * 4);
        bucket_order(ns, col, ncol, perm, NULL, m->s_col_off, m->s_col_cnt);
        m->s_v0 = (uint32_t *)malloc((size_t)(ns ? ns : 1) * 4);
        m->s_v1 = (uint32_t *)malloc((size_t)(ns ? ns : 1) * 4);
        m->s_rest = (float *)malloc((size_t)(ns ? ns : 1) * 4);
        for (uint32_t q = 0; q < ns; q++) { m->s_v0[q] = s0.d[perm[q]]; m->s_v1[q] = s1.d[perm[q]]; m->s_rest[q] = sr.d[perm[q]]; }
        free(perm); free(col); free(two); free(vv); free(s0.d); free(s1.d); free(sr.d);
    }

    /* 4. adjacency + local edge lengths (mesh.rs:949-1075) */
    uint32_t *acnt = (uint32_t *)calloc(nv + 1, 4);
    float *lsum = (float *)calloc(nv + 1, 4);
    uint32_t *lcnt = (uint32_t *)calloc(nv + 1, 4);
    float total_len = 0.0f, max_len = 0.0f;
    uint32_t valid_edges = 0;
    for (uint32_t e = 0; e < ne_in; e++) {
        uint32_t a = edges[2 * e], b = edges[2 * e + 1];
        if (a < nv && b < nv && a != b) {
            float len = frest(pos + 3 * a, pos + 3 * b);
            total_len += len;
            if (len > max_len) max_len = len;
            valid_edges++;
            acnt[a]++; acnt[b]++;
            lsum[a] += len; lcnt[a]++;
            lsum[b] += len; lcnt[b]++;
        }
    }
    m->mean_edge = valid_edges ? total_len / (float)valid_edges : 0.0f;
    m->max_edge = max_len;
    m->local_len = (float *)malloc((size_t)(nv ? nv : 1) * 4);
    for (uint32_t i = 0; i < nv; i++)
        m->local_len[i] = lcnt[i] > 0 ? lsum[i] / (float)lcnt[i] : fmaxf(default_thickness, 0.01f);
    uint32_t *aoff_raw = (uint32_t *)malloc((nv + 1) * 4);
    aoff_raw[0] = 0;
    for (uint32_t v = 0; v < nv; v++) aoff_raw[v + 1] = aoff_raw[v] + acnt[v];
    uint32_t *araw = (uint32_t *)malloc((size_t)(aoff_raw[nv] ? aoff_raw[nv] : 1) * 4);
    uint32_t *afill = (uint32_t *)calloc(nv + 1, 4);
    for (uint32_t e = 0; e < ne_in; e++) {
        uint32_t a = edges[2 * e], b = edges[2 * e + 1];
        if (a < nv && b < nv && a != b) { araw[aoff_raw[a] + afill[a]++] = b; araw[aoff_raw[b] + afill[b]++] = a; }
    }
    /* the original adjacency, sorted and unique (orig_off / orig) */
    uint32_t *orig_off = (uint32_t *)malloc((nv + 1) * 4);
    vu32 origv = {0};
    for (uint32_t v = 0; v < nv; v++) {
        orig_off[v] = (uint32_t)origv.n;
        size_t n = sort_unique(araw + aoff_raw[v], acnt[v]);
        for (size_t k = 0; k < n; k++) vu_push(&origv, araw[aoff_raw[v] + k]);
    }
    orig_off[nv] = (uint32_t)origv.n;
    uint32_t *orig = origv.d ? origv.d : (uint32_t *)calloc(1, 4);
    /* with sewing, a group's vertices neighbour each other and share every outside neighbour of the
     * group, both ways (mesh.rs:1031-1063); the adjacency CSR is then sorted and unique */
    vu32 *cl = (vu32 *)calloc(nv + 1, sizeof(vu32));
    for (uint32_t v = 0; v < nv; v++)
        for (uint32_t k = orig_off[v]; k < orig_off[v + 1]; k++) vu_push(&cl[v], orig[k]);
    if (sewing) {
        uint32_t *gcnt = (uint32_t *)calloc(nv + 1, 4), *goff = (uint32_t *)malloc((nv + 1) * 4);
        for (uint32_t i = 0; i < nv; i++) gcnt[rep[i]]++;
        goff[0] = 0;
        for (uint32_t r = 0; r < nv; r++) goff[r + 1] = goff[r] + gcnt[r];
        uint32_t *gm = (uint32_t *)malloc((size_t)(nv ? nv : 1) * 4), *gf = (uint32_t *)calloc(nv + 1, 4);
        for (uint32_t i = 0; i < nv; i++) gm[goff[rep[i]] + gf[rep[i]]++] = i;
        for (uint32_t r = 0; r < nv; r++) {
            if (gcnt[r] < 2) continue;
            const uint32_t *g = gm + goff[r];
            for (uint32_t a = 0; a < gcnt[r]; a++)
                for (uint32_t b = 0; b < gcnt[r]; b++)
                    if (a != b) vu_push(&cl[g[a]], g[b]);
            vu32 shared = {0};
            for (uint32_t a = 0; a < gcnt[r]; a++)
                for (uint32_t k = orig_off[g[a]]; k < orig_off[g[a] + 1]; k++)
                    if (rep[orig[k]] != r) vu_push(&shared, orig[k]);
            size_t nsh = sort_unique(shared.d, shared.n);
            for (uint32_t a = 0; a < gcnt[r]; a++)
                for (size_t k = 0; k < nsh; k++) { vu_push(&cl[g[a]], shared.d[k]); vu_push(&cl[shared.d[k]], g[a]); }
            free(shared.d);
        }
        free(gcnt); free(goff); free(gm); free(gf);
    }
    m->adj_off = (uint32_t *)malloc((nv + 1) * 4);
    vu32 aidx = {0};
    for (uint32_t v = 0; v < nv; v++) {
        m->adj_off[v] = (uint32_t)aidx.n;
        size_t n = sort_unique(cl[v].d, cl[v].n);
        for (size_t k = 0; k < n; k++) vu_push(&aidx, cl[v].d[k]);
        free(cl[v].d);
    }
    free(cl);
    m->adj_off[nv] = (uint32_t)aidx.n;
    m->adj_idx = aidx.d ? aidx.d : (uint32_t *)calloc(1, 4);

    /* two-hop CSR (mesh.rs:1087-1191) over the adjacency above.  A true two-hop pair is of face
     * vertices in different sewing groups, joined by no original edge, whose groups are not
     * neighbours in the contracted graph (c1, in group labels); its rest length is the distance when
     * the pair is two hops apart in the original mesh too, else the seam rest length. */
    vu32 *c1 = (vu32 *)calloc(nv + 1, sizeof(vu32));
    if (sewing) {
        for (uint32_t e = 0; e < ne_in; e++) {
            uint32_t a = edges[2 * e], b = edges[2 * e + 1];
            if (a < nv && b < nv && a != b && rep[a] != rep[b]) { vu_push(&c1[rep[a]], rep[b]); vu_push(&c1[rep[b]], rep[a]); }
        }
        for (uint32_t v = 0; v < nv; v++) c1[v].n = sort_unique(c1[v].d, c1[v].n);
    }
    vu32 onb = {0};
    uint8_t *is_face = (uint8_t *)calloc(nv + 1, 1);
    if (faces) for (uint32_t f = 0; f < nf; f++)
        for (int k = 0; k < 3; k++) if (faces[3 * f + k] < nv) is_face[faces[3 * f + k]] = 1;
    m->th_off = (uint32_t *)malloc((nv + 1) * 4);
    vu32 thi = {0}; vf32 thr = {0};
    vu32 nb_tmp = {0};
    for (uint32_t i = 0; i < nv; i++) {
        nb_tmp.n = 0;
        vu_push(&nb_tmp, i);
        for (uint32_t k = m->adj_off[i]; k < m->adj_off[i + 1]; k++) {
            uint32_t u = m->adj_idx[k];
            vu_push(&nb_tmp, u);
            for (uint32_t ku = m->adj_off[u]; ku < m->adj_off[u + 1]; ku++) vu_push(&nb_tmp, m->adj_idx[ku]);
        }
        size_t n = sort_unique(nb_tmp.d, nb_tmp.n);
        m->th_off[i] = (uint32_t)thi.n;
        const uint32_t *ai = m->adj_idx + m->adj_off[i];
        size_t an = m->adj_off[i + 1] - m->adj_off[i];
        size_t on = 0;
        if (sewing) {
            onb.n = 0;
            vu_push(&onb, i);
            for (uint32_t k = orig_off[i]; k < orig_off[i + 1]; k++) {
                uint32_t u = orig[k];
                vu_push(&onb, u);
                for (uint32_t ku = orig_off[u]; ku < orig_off[u + 1]; ku++) vu_push(&onb, orig[ku]);
            }
            on = sort_unique(onb.d, onb.n);
        }
        for (size_t k = 0; k < n; k++) {
            uint32_t nbr = nb_tmp.d[k];
            vu_push(&thi, nbr);
            float rest = 0.0f;
            if (!sewing) {
                int true2 = nbr != i && is_face[i] && is_face[nbr] && !bsearch_u32(ai, an, nbr);
                if (true2) rest = frest(pos + 3 * i, pos + 3 * nbr);
            } else {
                int true2 = nbr != i && is_face[i] && is_face[nbr] && rep[i] != rep[nbr] &&
                            !bsearch_u32(orig + orig_off[i], orig_off[i + 1] - orig_off[i], nbr) &&
                            !bsearch_u32(c1[rep[i]].d, c1[rep[i]].n, rep[nbr]);
                if (true2)
                    rest = bsearch_u32(onb.d, on, nbr) ? frest(pos + 3 * i, pos + 3 * nbr)
                                                       : seam_two_hop(i, nbr, pos, orig_off, orig, rep);
            }
            vf_push(&thr, rest);
        }
    }
    m->th_off[nv] = (uint32_t)thi.n;
    m->th_idx = thi.d ? thi.d : (uint32_t *)calloc(1, 4);
    m->th_rest = thr.d ? thr.d : (float *)calloc(1, 4);
    free(nb_tmp.d); free(is_face); free(onb.d);
    for (uint32_t v = 0; v < nv; v++) free(c1[v].d);
    free(c1);

    /* star lists (mesh.rs:1193-1212), triangles */
    uint32_t *scount = (uint32_t *)calloc(nv + 1, 4);
    uint32_t ntri = 0;
    if (faces) for (uint32_t f = 0; f < nf; f++) {
        uint32_t f0 = faces[3 * f], f1 = faces[3 * f + 1], f2 = faces[3 * f + 2];
        if (f0 < nv && f1 < nv && f2 < nv) { scount[f0]++; scount[f1]++; scount[f2]++; ntri++; }
    }
    m->star_off = (uint32_t *)malloc((nv + 1) * 4);
    m->star_off[0] = 0;
    for (uint32_t v = 0; v < nv; v++) m->star_off[v + 1] = m->star_off[v] + scount[v];
    uint32_t ns = m->star_off[nv];
    m->star_v0 = (uint32_t *)malloc((size_t)(ns ? ns : 1) * 4);
    m->star_v1 = (uint32_t *)malloc((size_t)(ns ? ns : 1) * 4);
    m->ntri = ntri;
    m->tri = (uint32_t *)malloc((size_t)(ntri ? ntri : 1) * 12);
    memset(scount, 0, (nv + 1) * 4);
    ntri = 0;
    if (faces) for (uint32_t f = 0; f < nf; f++) {
        uint32_t f0 = faces[3 * f], f1 = faces[3 * f + 1], f2 = faces[3 * f + 2];
        if (f0 < nv && f1 < nv && f2 < nv) {
            uint32_t s;
            s = m->star_off[f0] + scount[f0]++; m->star_v0[s] = f1; m->star_v1[s] = f2;
            s = m->star_off[f1] + scount[f1]++; m->star_v0[s] = f2; m->star_v1[s] = f0;
            s = m->star_off[f2] + scount[f2]++; m->star_v0[s] = f0; m->star_v1[s] = f1;
            m->tri[3 * ntri] = f0; m->tri[3 * ntri + 1] = f1; m->tri[3 * ntri + 2] = f2; ntri++;
        }
    }
    free(scount);

    /* islands (mesh.rs:1228-1255) */
    m->island = (uint32_t *)malloc((size_t)(nv ? nv : 1) * 4);
    for (uint32_t i = 0; i < nv; i++) m->island[i] = 0xFFFFFFFFu;
    uint32_t *queue = (uint32_t *)malloc((size_t)(nv ? nv : 1) * 4);
    uint32_t cur_island = 0;
    for (uint32_t i = 0; i < nv; i++) {
        if (m->island[i] != 0xFFFFFFFFu) continue;
        size_t qh = 0, qt = 0;
        m->island[i] = cur_island; queue[qt++] = i;
        while (qh < qt) {
            uint32_t u = queue[qh++];
            for (uint32_t k = m->adj_off[u]; k < m->adj_off[u + 1]; k++) {
                uint32_t v = m->adj_idx[k];
                if (v < nv && m->island[v] == 0xFFFFFFFFu) { m->island[v] = cur_island; queue[qt++] = v; }
            }
        }
        cur_island++;
    }
    free(queue);

    /* Morton-sorted unique edges (mesh.rs:1257-1289) + boundary bits (1291-1344) */
    {
        pset es; pset_init(&es, ne_in);
        vu32 e0 = {0}, e1 = {0};
        for (uint32_t e = 0; e < ne_in; e++) {
            uint32_t a = edges[2 * e], b = edges[2 * e + 1];
            if (a < nv && b < nv && a != b) {
                uint32_t x = a < b ? a : b, y = a < b ? b : a;
                if (pset_insert(&es, x, y)) { vu_push(&e0, x); vu_push(&e1, y); }
            }
        }
        pset_free(&es);
        uint32_t ne = (uint32_t)e0.n;
        m->ne = ne;
        m->e_v0 = (uint32_t *)malloc((size_t)(ne ? ne : 1) * 4);
        m->e_v1 = (uint32_t *)malloc((size_t)(ne ? ne : 1) * 4);
        if (ne && nv) {
            float bmin[3] = {3.4028235e38f, 3.4028235e38f, 3.4028235e38f};
            float bmax[3] = {-3.4028235e38f, -3.4028235e38f, -3.4028235e38f};
            for (uint32_t i = 0; i < nv; i++)
                for (int a = 0; a < 3; a++) { bmin[a] = fminf(bmin[a], pos[3 * i + a]); bmax[a] = fmaxf(bmax[a], pos[3 * i + a]); }
            float inv[3];
            for (int a = 0; a < 3; a++) inv[a] = 1.0f / fmaxf(bmax[a] - bmin[a], 1e-5f);
            mkey *mk_ = (mkey *)malloc(ne * sizeof(mkey));
            for (uint32_t k = 0; k < ne; k++) {
                const float *p0 = pos + 3 * e0.d[k], *p1 = pos + 3 * e1.d[k];
                float mid[3] = {(p0[0] + p1[0]) * 0.5f, (p0[1] + p1[1]) * 0.5f, (p0[2] + p1[2]) * 0.5f};
                mk_[k].code = morton_code_3d(mid, bmin, inv); mk_[k].idx = k;
            }
            qsort(mk_, ne, sizeof(mkey), cmp_mkey);
            for (uint32_t k = 0; k < ne; k++) { m->e_v0[k] = e0.d[mk_[k].idx]; m->e_v1[k] = e1.d[mk_[k].idx]; }
            free(mk_);
        }
        free(e0.d); free(e1.d);
        uint32_t words = (ne + 31) / 32;
        m->e_bbits = (uint32_t *)calloc(words ? words : 1, 4);
        if (faces && nf) {
            /* triangle count per edge key; sorted key array + counting via eo-like list */
            uint64_t *keys = (uint64_t *)malloc((size_t)nf * 3 * 8);
            size_t nk = 0;
            for (uint32_t f = 0; f < nf; f++) {
                uint32_t f0 = faces[3 * f], f1 = faces[3 * f + 1], f2 = faces[3 * f + 2];
                uint32_t te[3][2] = {{f0 < f1 ? f0 : f1, f0 < f1 ? f1 : f0}, {f1 < f2 ? f1 : f2, f1 < f2 ? f2 : f1},
                                     {f2 < f0 ? f2 : f0, f2 < f0 ? f0 : f2}};
                for (int k = 0; k < 3; k++)
                    if (te[k][0] < nv && te[k][1] < nv && te[k][0] != te[k][1])
                        keys[nk++] = ((uint64_t)te[k][0] << 32) | te[k][1];
            }
            qsort(keys, nk, 8, cmpu64);
            /* the same triangle edges in sewing-group labels (pairs within a group dropped) */
            uint64_t *ckeys = (uint64_t *)malloc((nk ? nk : 1) * 8);
            size_t nck = 0;
            for (size_t k = 0; k < nk; k++) {
                uint32_t ra = rep[(uint32_t)(keys[k] >> 32)], rb = rep[(uint32_t)keys[k]];
                if (ra != rb) ckeys[nck++] = ra < rb ? ((uint64_t)ra << 32) | rb : ((uint64_t)rb << 32) | ra;
            }
            qsort(ckeys, nck, 8, cmpu64);
            for (uint32_t k = 0; k < ne; k++) {
                uint32_t ra = rep[m->e_v0[k]], rb = rep[m->e_v1[k]];
                if (ra == rb) continue;  /* an edge inside a sewing group is never open */
                uint64_t key = ((uint64_t)m->e_v0[k] << 32) | m->e_v1[k];
                size_t lo = 0, hi = nk;
                while (lo < hi) { size_t mid = lo + (hi - lo) / 2; if (keys[mid] < key) lo = mid + 1; else hi = mid; }
                size_t c = 0;
                while (lo + c < nk && keys[lo + c] == key) c++;
                int open = c == 0;
                if (c == 1) {  /* open only if its contracted edge is in one triangle too (mesh.rs:1323-1330) */
                    uint64_t ck = ra < rb ? ((uint64_t)ra << 32) | rb : ((uint64_t)rb << 32) | ra;
                    lo = 0; hi = nck;
                    while (lo < hi) { size_t mid = lo + (hi - lo) / 2; if (ckeys[mid] < ck) lo = mid + 1; else hi = mid; }
                    size_t cc = 0;
                    while (lo + cc < nck && ckeys[lo + cc] == ck) cc++;
                    open = cc == 1;
                }
                if (open) m->e_bbits[k >> 5] |= 1u << (k & 31);
            }
            free(keys); free(ckeys);
        } else {
            for (uint32_t k = 0; k < ne; k++)
                if (rep[m->e_v0[k]] != rep[m->e_v1[k]]) m->e_bbits[k >> 5] |= 1u << (k & 31);
        }
    }

    /* island layer assignment (pipelines.rs:131-166) */
    if (nv) {
        uint32_t first = m->layer[0];
        int all_same = 1;
        for (uint32_t i = 0; i < nv; i++) if (m->layer[i] != first) { all_same = 0; break; }
        uint32_t max_island = 0;
        for (uint32_t i = 0; i < nv; i++) if (m->island[i] > max_island) max_island = m->island[i];
        if (all_same && max_island > 0) {
            float *zs = (float *)calloc(max_island + 1, 4);
            uint32_t *zc = (uint32_t *)calloc(max_island + 1, 4);
            for (uint32_t i = 0; i < nv; i++) { zs[m->island[i]] += pos[3 * i + 2]; zc[m->island[i]]++; }
            float *mz = (float *)malloc((max_island + 1) * 4);
            uint32_t *ord = (uint32_t *)malloc((max_island + 1) * 4);
            for (uint32_t k = 0; k <= max_island; k++) { mz[k] = zs[k] / (float)(zc[k] ? zc[k] : 1); ord[k] = k; }
            for (uint32_t a = 1; a <= max_island; a++) { /* stable insertion sort by mean z */
                uint32_t x = ord[a]; uint32_t b = a;
                while (b > 0 && mz[ord[b - 1]] > mz[x]) { ord[b] = ord[b - 1]; b--; }
                ord[b] = x;
            }
            uint32_t *to_layer = (uint32_t *)malloc((max_island + 1) * 4);
            for (uint32_t r = 0; r <= max_island; r++) to_layer[ord[r]] = r;
            for (uint32_t i = 0; i < nv; i++) m->layer[i] = to_layer[m->island[i]];
            free(zs); free(zc); free(mz); free(ord); free(to_layer);
        }
        m->has_multiple_layers = 0;
        for (uint32_t i = 0; i < nv; i++) if (m->layer[i] != m->layer[0]) { m->has_multiple_layers = 1; break; }
    }

    /* hash parameters (pipelines.rs:899-912) */
    m->n_virtual = virtual_sample_count(m, pos);
    float th = nv ? m->thick[0] : 0.005f;
    float edge_based = fminf(fmaxf(m->max_edge * 0.85f, m->mean_edge * 1.0f), 0.025f);
    m->cell_size = fmaxf(fmaxf(th * 4.0f, edge_based), 0.01f);
    uint32_t cap = nv + m->n_virtual;
    uint32_t ts = next_pow2(cap * 4u);
    if (ts < 1024) ts = 1024;
    if (ts > 65536) ts = 65536;
    m->table_size = ts;

    free(aoff_raw); free(araw); free(afill); free(acnt); free(lsum); free(lcnt);
    free(orig_off); free(orig); free(rep);
    free(eo);
    pset_free(&original);
    free(dv0.d); free(dv1.d); free(dty.d); free(drest.d); free(dtc.d); free(dcc.d);
    free(bv0.d); free(bv1.d); free(bv2.d); free(bv3.d); free(brest.d); free(bcomp.d); free(bpad0.d); free(bpad1.d);
    return m;
}

void orc_mesh_free(orc_mesh *m) {
    if (!m) return;
    void *ps[] = {m->pos0, m->inv_mass, m->thick, m->layer, m->d_v0, m->d_v1, m->d_type, m->d_rest, m->d_tc, m->d_cc,
                  m->d_col_off, m->d_col_cnt, m->d_remap, m->edge_to_constraint, m->b_v0, m->b_v1, m->b_v2, m->b_v3,
                  m->b_rest, m->b_comp, m->b_pad0, m->b_pad1, m->b_col_off, m->b_col_cnt, m->adj_off, m->adj_idx,
                  m->th_off, m->th_idx, m->th_rest, m->local_len, m->star_off, m->star_v0, m->star_v1, m->tri,
                  m->e_v0, m->e_v1, m->e_bbits, m->island, m->s_v0, m->s_v1, m->s_rest, m->s_col_off,
                  m->s_col_cnt};
    for (size_t i = 0; i < sizeof(ps) / sizeof(ps[0]); i++) free(ps[i]);
    free(m);
}

/* mesh accessors for tests: what = field id, returns element count; out may be NULL */
enum {
    OM_DIST_V0 = 0, OM_DIST_V1, OM_DIST_TYPE, OM_DIST_REST, OM_DIST_TC, OM_DIST_CC, OM_DIST_COL_OFF, OM_DIST_COL_CNT,
    OM_BEND_V0, OM_BEND_V1, OM_BEND_V2, OM_BEND_V3, OM_BEND_REST, OM_BEND_COMP, OM_BEND_PAD0, OM_BEND_COL_OFF,
    OM_BEND_COL_CNT, OM_ADJ_OFF, OM_ADJ_IDX, OM_TH_OFF, OM_TH_IDX, OM_TH_REST, OM_LOCAL_LEN, OM_STAR_OFF,
    OM_STAR_V0, OM_STAR_V1, OM_EDGE_V0, OM_EDGE_V1, OM_EDGE_BBITS, OM_ISLAND, OM_SCALARS, OM_EDGE_TO_CONSTRAINT,
    OM_LAYER, OM_BEND_PAD1, OM_SEW_V0, OM_SEW_V1, OM_SEW_REST, OM_SEW_COL_OFF, OM_SEW_COL_CNT
};
uint32_t orc_mesh_get(const orc_mesh *m, int what, void *out) {
#define CP(ptr, n) do { if (out && (n)) memcpy(out, ptr, (size_t)(n) * 4); return (uint32_t)(n); } while (0)
    switch (what) {
    case OM_DIST_V0: CP(m->d_v0, m->nd);
    case OM_DIST_V1: CP(m->d_v1, m->nd);
    case OM_DIST_TYPE: CP(m->d_type, m->nd);
    case OM_DIST_REST: CP(m->d_rest, m->nd);
    case OM_DIST_TC: CP(m->d_tc, m->nd);
    case OM_DIST_CC: CP(m->d_cc, m->nd);
    case OM_DIST_COL_OFF: CP(m->d_col_off, m->nd_col);
    case OM_DIST_COL_CNT: CP(m->d_col_cnt, m->nd_col);
    case OM_BEND_V0: CP(m->b_v0, m->nb);
    case OM_BEND_V1: CP(m->b_v1, m->nb);
    case OM_BEND_V2: CP(m->b_v2, m->nb);
    case OM_BEND_V3: CP(m->b_v3, m->nb);
    case OM_BEND_REST: CP(m->b_rest, m->nb);
    case OM_BEND_COMP: CP(m->b_comp, m->nb);
    case OM_BEND_PAD0: CP(m->b_pad0, m->nb);
    case OM_BEND_COL_OFF: CP(m->b_col_off, m->nb_col);
    case OM_BEND_COL_CNT: CP(m->b_col_cnt, m->nb_col);
    case OM_ADJ_OFF: CP(m->adj_off, m->nv + 1);
    case OM_ADJ_IDX: CP(m->adj_idx, m->adj_off[m->nv]);
    case OM_TH_OFF: CP(m->th_off, m->nv + 1);
    case OM_TH_IDX: CP(m->th_idx, m->th_off[m->nv]);
    case OM_TH_REST: CP(m->th_rest, m->th_off[m->nv]);
    case OM_LOCAL_LEN: CP(m->local_len, m->nv);
    case OM_STAR_OFF: CP(m->star_off, m->nv + 1);
    case OM_STAR_V0: CP(m->star_v0, m->star_off[m->nv]);
    case OM_STAR_V1: CP(m->star_v1, m->star_off[m->nv]);
    case OM_EDGE_V0: CP(m->e_v0, m->ne);
    case OM_EDGE_V1: CP(m->e_v1, m->ne);
    case OM_EDGE_BBITS: CP(m->e_bbits, (m->ne + 31) / 32);
    case OM_ISLAND: CP(m->island, m->nv);
    case OM_EDGE_TO_CONSTRAINT: CP(m->edge_to_constraint, m->ne_in);
    case OM_LAYER: CP(m->layer, m->nv);
    case OM_BEND_PAD1: CP(m->b_pad1, m->nb);
    case OM_SEW_V0: CP(m->s_v0, m->ns);
    case OM_SEW_V1: CP(m->s_v1, m->ns);
    case OM_SEW_REST: CP(m->s_rest, m->ns);
    case OM_SEW_COL_OFF: CP(m->s_col_off, m->ns_col);
    case OM_SEW_COL_CNT: CP(m->s_col_cnt, m->ns_col);
    case OM_SCALARS: {
        /* [nv, nd, nb, nd_col, nb_col, ne, ntri, n_virtual, table_size, has_multi, bits(cell), bits(mean), bits(max)] */
        uint32_t s[13];
        s[0] = m->nv; s[1] = m->nd; s[2] = m->nb; s[3] = m->nd_col; s[4] = m->nb_col; s[5] = m->ne; s[6] = m->ntri;
        s[7] = m->n_virtual; s[8] = m->table_size; s[9] = (uint32_t)m->has_multiple_layers;
        memcpy(&s[10], &m->cell_size, 4); memcpy(&s[11], &m->mean_edge, 4); memcpy(&s[12], &m->max_edge, 4);
        CP(s, 13);
    }
    default: return 0;
    }
#undef CP
}

/* mesh.rs:237-312 (areal part; no strands, no pins) */
void orc_areal_inv_masses(const float *pos, uint32_t nv, const uint32_t *faces, uint32_t nf, float density,
                          const float *pin_weights, float *out) {
    float *mass = (float *)calloc(nv ? nv : 1, 4);
    if (density > 0.0f)
        for (uint32_t f = 0; f < nf; f++) {
            uint32_t a = faces[3 * f], b = faces[3 * f + 1], c = faces[3 * f + 2];
            if (a >= nv || b >= nv || c >= nv || a == b || b == c || c == a) continue;
            const float *pa = pos + 3 * a, *pb = pos + 3 * b, *pc = pos + 3 * c;
            float ab[3] = {pb[0] - pa[0], pb[1] - pa[1], pb[2] - pa[2]};
            float ac[3] = {pc[0] - pa[0], pc[1] - pa[1], pc[2] - pa[2]};
            float cr[3] = {ab[1] * ac[2] - ab[2] * ac[1], ab[2] * ac[0] - ab[0] * ac[2], ab[0] * ac[1] - ab[1] * ac[0]};
            float area = 0.5f * sqrtf(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]);
            float share = area * density / 3.0f;
            mass[a] += share; mass[b] += share; mass[c] += share;
        }
    for (uint32_t i = 0; i < nv; i++) {
        float w = pin_weights ? fminf(fmaxf(pin_weights[i], 0.0f), 1.0f) : 0.0f;
        out[i] = w >= 1.0f ? 0.0f : (1.0f - w) / fmaxf(mass[i], 1e-9f);
    }
    free(mass);
}

/* coloring.rs:180-... generic edge-pair colouring (returns colour per pair) */
void orc_color_edge_pairs(uint32_t nv, const uint32_t *pairs, uint32_t np, uint32_t *colors_out) {
    uint32_t *reg = (uint32_t *)calloc((size_t)(np ? np : 1) * 2, 4);
    uint32_t *regn = (uint32_t *)calloc((size_t)(np ? np : 1), 4);
    for (uint32_t i = 0; i < np; i++) {
        uint32_t a = pairs[2 * i], b = pairs[2 * i + 1];
        if (a < nv && b < nv) { reg[2 * i] = a; reg[2 * i + 1] = b; regn[i] = 2; }
        else regn[i] = 0;
    }
    uint32_t ncol = 0;
    welsh_powell(nv, np, reg, regn, 2, reg, regn, 2, reg, regn, 2, colors_out, &ncol);
    for (uint32_t i = 0; i < np; i++) if (regn[i] == 0) colors_out[i] = 0;
    free(reg); free(regn);
}

/* ------------------------------------------------------------------ */
/* spatial hash (spatial_grid_sort.wgsl:25-195)                         */
/* ------------------------------------------------------------------ */
static inline int32_t f2i_sat(real x) { /* WGSL i32(f32): truncation, saturating */
    if (x != x) return 0;
    if (x >= (real)2147483520.0) return 2147483647;
    if (x <= (real)-2147483648.0) return (int32_t)0x80000000u;
    return (int32_t)x;
}
static inline void cell_of(v3 p, real cell, int32_t *c) {
    c[0] = f2i_sat(RFLOOR(p.x / cell));
    c[1] = f2i_sat(RFLOOR(p.y / cell));
    c[2] = f2i_sat(RFLOOR(p.z / cell));
}
static inline uint32_t hash_coords(const int32_t *c, uint32_t table) {
    uint32_t n = ((uint32_t)c[0] * 73856093u) ^ ((uint32_t)c[1] * 19349663u) ^ ((uint32_t)c[2] * 83492791u);
    return n % table;
}

/* Counting sort; within a bucket elements are in ascending index order. */
static void hash_build(const v3 *pts, uint32_t n, real cell, uint32_t table, uint32_t *keys, uint32_t *starts,
                       uint32_t *sorted) {
    memset(starts, 0, (size_t)(table + 1) * 4);
    for (uint32_t i = 0; i < n; i++) {
        int32_t c[3]; cell_of(pts[i], cell, c);
        keys[i] = hash_coords(c, table);
        starts[keys[i]]++;
    }
    uint32_t acc = 0;
    for (uint32_t h = 0; h < table; h++) { uint32_t c = starts[h]; starts[h] = acc; acc += c; }
    starts[table] = acc;
    uint32_t *fill = (uint32_t *)calloc(table, 4);
    for (uint32_t i = 0; i < n; i++) sorted[starts[keys[i]] + fill[keys[i]]++] = i;
    free(fill);
}

/* standalone: keys/starts/sorted for a point cloud (test entry) */
void orc_hash_points(const float *pts3, uint32_t n, float cell, uint32_t table, uint32_t *keys, uint32_t *starts,
                     uint32_t *sorted) {
    v3 *p = (v3 *)malloc((size_t)(n ? n : 1) * sizeof(v3));
    for (uint32_t i = 0; i < n; i++) p[i] = mk((real)pts3[3 * i], (real)pts3[3 * i + 1], (real)pts3[3 * i + 2]);
    hash_build(p, n, (real)cell, table, keys, starts, sorted);
    free(p);
}

/* ------------------------------------------------------------------ */
/* simulator                                                           */
/* ------------------------------------------------------------------ */
typedef struct {
    uint32_t type; real friction, restitution; v3 a; real radius; v3 b;
} collider;

typedef struct {
    float p0[3], friction, p1[3], thickness, p2[3], restitution;
    uint32_t flags;
} mtri;

typedef struct {
    int32_t dx, dy, dz; uint32_t count;
    int32_t pdx, pdy, pdz; uint32_t pcount;
    int32_t cpx, cpy, cpz, cnx, cny, cnz;
    uint32_t untangle, ee_count;
} accum_t;

typedef struct {
    orc_mesh *m;
    uint32_t nv;
    v3 *x, *p, *v;
    real *w;
    v3 *x0;             /* initial positions (reset) */
    real *w0;
    real *lam_d, *lam_b;
    /* per-constraint compliances in the working type (raw overrides possible) */
    real *d_tc, *d_cc, *b_comp;
    real gravity[3];
    real damping;
    real tension_damping, compression_damping, shear_damping, bending_damping;
    uint32_t iterations;
    int self_collision;
    collider *cols; uint32_t ncols;
    mtri *tris; float *tri_bounds; uint32_t ntris;   /* Morton-sorted, bounds = (cx, cy, cz, r) */
    int single_sided_recovery;
    real cell_size; uint32_t table_size;
    accum_t *acc;
    uint32_t *hkeys, *hstarts, *hsorted;     /* vertex hash */
    uint32_t *ekeys, *estarts, *esorted;     /* edge hash */
    uint32_t max_search_iterations;
    real relief_factor, max_disp_ratio;
    int enable_relief, exclude_neighbors;
    uint64_t stats[6]; /* contact counters: vv, vt_ccd, vt_prox, ee_ccd, ee_prox, hash_pairs */
    uint32_t *sdf_tex; float *sdf_info, *sdf_xf;     /* SDF colliders (sdf_oracle.c) */
    uint32_t sdf_w, sdf_h, sdf_d, sdf_nb;
    uint32_t npins, *pin_v; v3 *pin_t; real *pin_w;  /* dynamic pin table (distinct vertices) */
    uint32_t solver_mode;                            /* 1: fixed-point Jacobi distance pass */
    /* pair cache (self_collision_algorithm 1): vertex-major slots of 4 indices, counters */
    int pair_cache, pair_final_fallback, pair_brute;
    uint32_t pair_quota_vt, pair_quota_ee, pair_margin_mode;
    real pair_safety, pair_hscale, pair_hmax, pair_dt_frame;
    uint32_t *pair_vt, *pair_ee, pair_counters[4];
    /* collider-edge collision: unique collider edges {p0, thickness, p1, friction, n, flags} and
     * their grid (32768 buckets) */
    int edge_collision; real edge_scale, edge_offset;
    /* wrinkle field: parameters (origin4, axis4, normal4, binormal4, z/r ranges, windows) and texture */
    int wrinkle_on; float wrinkle[22]; uint8_t *wtex; uint32_t ww, wh;
    float *cedges; uint32_t *cedge_flags, nce, *cstarts, *csorted;
    int32_t *jac;                                    /* per-vertex sums {dx, dy, dz, count} */
    /* sewing: current and target rest lengths, lambdas; the simulator's stiffness, lock and lock
     * distance (mod.rs:637-640) and shrink speed */
    real *s_cur, *s_tgt, *lam_s;
    float sew_stiffness, sew_lock_dist, sew_speed;
    int sew_lock;
    /* dynamic bone SDF (setup_dynamic_bone_sdf): rest mesh with four skin influences per vertex,
     * triangle sources {i0, i1, i2, bone} with corner weights and open-edge masks, bind matrices,
     * bake parameters (16 words per bone, GpuBakeParams layout) and the device parameter buffer's
     * contents; counters follow dispatch_dynamic_bone_sdf_update */
    int dyn_ready, dyn_enabled, dyn_baked, dyn_dirty_set;
    uint32_t dyn_interval, dyn_counter, dyn_res, dyn_nb, dyn_nv, dyn_nt, dyn_ndirty;
    float *dyn_rest, *dyn_w, *dyn_tw, *dyn_bind, *dyn_tris, *dyn_skinned;
    uint32_t *dyn_bi, *dyn_src, *dyn_bmask, *dyn_params, *dyn_params_dev, *dyn_dirty;
    /* multi-layer normal untangling: the user flag and the vertex normals of the substep's start */
    int normal_untangling;
    v3 *normals;
} orc_sim;

orc_sim *orc_sim_create_sewn(const float *pos, uint32_t nv, const uint32_t *edges, uint32_t ne,
                             const uint32_t *faces, uint32_t nf, const float *inv_masses, const float *thicknesses,
                             const uint32_t *layer_ids, uint32_t layer_id, float thickness, float stiffness,
                             float bending_stiffness, float compression_stiffness, float shear_stiffness,
                             const uint32_t *sew, uint32_t nsew, float sew_speed, float sew_lock_distance);
orc_sim *orc_sim_create(const float *pos, uint32_t nv, const uint32_t *edges, uint32_t ne, const uint32_t *faces,
                        uint32_t nf, const float *inv_masses, const float *thicknesses, const uint32_t *layer_ids,
                        uint32_t layer_id, float thickness, float stiffness, float bending_stiffness,
                        float compression_stiffness, float shear_stiffness) {
    return orc_sim_create_sewn(pos, nv, edges, ne, faces, nf, inv_masses, thicknesses, layer_ids, layer_id, thickness,
                               stiffness, bending_stiffness, compression_stiffness, shear_stiffness, NULL, 0, 1.0f, -1.0f);
}

/* ClothSimulator::new with sewing_springs (lib.rs:1643-1751); sew_lock_distance < 0: not given. */
orc_sim *orc_sim_create_sewn(const float *pos, uint32_t nv, const uint32_t *edges, uint32_t ne,
                             const uint32_t *faces, uint32_t nf, const float *inv_masses, const float *thicknesses,
                             const uint32_t *layer_ids, uint32_t layer_id, float thickness, float stiffness,
                             float bending_stiffness, float compression_stiffness, float shear_stiffness,
                             const uint32_t *sew, uint32_t nsew, float sew_speed, float sew_lock_distance) {
    orc_sim *s = (orc_sim *)calloc(1, sizeof(orc_sim));
    s->m = orc_mesh_build_sewn(pos, nv, edges, ne, faces, nf, inv_masses, thicknesses, layer_ids, layer_id, thickness,
                               stiffness, compression_stiffness, shear_stiffness, bending_stiffness, sew, nsew,
                               sew_speed);
    orc_mesh *m = s->m;
    s->s_cur = (real *)malloc((m->ns ? m->ns : 1) * sizeof(real));
    s->s_tgt = (real *)calloc(m->ns ? m->ns : 1, sizeof(real));
    s->lam_s = (real *)calloc(m->ns ? m->ns : 1, sizeof(real));
    for (uint32_t k = 0; k < m->ns; k++) s->s_cur[k] = (real)m->s_rest[k];
    s->sew_stiffness = 10000.0f; s->sew_lock = 1;
    s->sew_lock_dist = sew_lock_distance >= 0.0f ? sew_lock_distance : 0.02f;
    s->sew_speed = m->sew_speed;
    s->nv = nv;
    size_t N = nv ? nv : 1;
    s->x = (v3 *)malloc(N * sizeof(v3)); s->p = (v3 *)malloc(N * sizeof(v3)); s->v = (v3 *)calloc(N, sizeof(v3));
    s->x0 = (v3 *)malloc(N * sizeof(v3));
    s->w = (real *)malloc(N * sizeof(real)); s->w0 = (real *)malloc(N * sizeof(real));
    for (uint32_t i = 0; i < nv; i++) {
        s->x0[i] = mk(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
        s->x[i] = s->x0[i]; s->p[i] = s->x0[i];
        s->w[i] = s->w0[i] = m->inv_mass[i];
    }
    s->lam_d = (real *)calloc(m->nd ? m->nd : 1, sizeof(real));
    s->lam_b = (real *)calloc(m->nb ? m->nb : 1, sizeof(real));
    s->d_tc = (real *)malloc((m->nd ? m->nd : 1) * sizeof(real));
    s->d_cc = (real *)malloc((m->nd ? m->nd : 1) * sizeof(real));
    s->b_comp = (real *)malloc((m->nb ? m->nb : 1) * sizeof(real));
    for (uint32_t k = 0; k < m->nd; k++) { s->d_tc[k] = m->d_tc[k]; s->d_cc[k] = m->d_cc[k]; }
    for (uint32_t k = 0; k < m->nb; k++) s->b_comp[k] = m->b_comp[k];
    s->gravity[0] = 0; s->gravity[1] = 0; s->gravity[2] = (real)-9.81f;
    s->damping = (real)1.0f;
    s->iterations = 2;
    s->self_collision = 1;
    s->cell_size = m->cell_size; s->table_size = m->table_size;
    s->acc = (accum_t *)calloc(N, sizeof(accum_t));
    s->jac = (int32_t *)calloc(4 * N, sizeof(int32_t));
    s->pair_vt = (uint32_t *)malloc(N * 8 * 4 * sizeof(uint32_t));
    s->pair_ee = (uint32_t *)malloc(N * 8 * 4 * sizeof(uint32_t));
    for (size_t k = 0; k < N * 8; k++) { s->pair_vt[4 * k] = 0xFFFFFFFFu; s->pair_ee[4 * k] = 0xFFFFFFFFu; }
    {   /* simulation/mod.rs:583-597 */
        uint64_t b = (uint64_t)nv * 8u;
        uint32_t budget = (uint32_t)(b < 8192 ? 8192 : b > 262144 ? 262144 : b);
        uint32_t q = nv ? budget / nv : 1;
        s->pair_quota_vt = s->pair_quota_ee = q < 1 ? 1 : q > 8 ? 8 : q;
    }
    s->pair_final_fallback = 1; s->pair_margin_mode = 1;
    s->pair_safety = (real)0.005f; s->pair_hscale = (real)1.3f; s->pair_hmax = (real)0.02f;
    s->pair_dt_frame = (real)(1.0f / 60.0f);
    s->edge_scale = 1; s->edge_offset = 0;
    s->hkeys = (uint32_t *)malloc(N * 4); s->hsorted = (uint32_t *)malloc(N * 4);
    s->hstarts = (uint32_t *)malloc((size_t)(s->table_size + 1) * 4);
    size_t NE = m->ne ? m->ne : 1;
    s->ekeys = (uint32_t *)malloc(NE * 4); s->esorted = (uint32_t *)malloc(NE * 4);
    s->estarts = (uint32_t *)malloc((size_t)(s->table_size + 1) * 4);
    s->max_search_iterations = 256;
    s->single_sided_recovery = 1;
    s->relief_factor = (real)0.2f; s->max_disp_ratio = (real)0.2f;
    s->enable_relief = 1; s->exclude_neighbors = 1;
    s->normal_untangling = 1;
    s->normals = (v3 *)calloc(nv ? nv : 1, sizeof(v3));
    return s;
}

void orc_sim_free(orc_sim *s) {
    if (!s) return;
    orc_mesh_free(s->m);
    free(s->x); free(s->p); free(s->v); free(s->x0); free(s->w); free(s->w0); free(s->lam_d); free(s->lam_b);
    free(s->d_tc); free(s->d_cc); free(s->b_comp); free(s->cols); free(s->acc); free(s->tris); free(s->tri_bounds);
    free(s->hkeys); free(s->hstarts); free(s->hsorted); free(s->ekeys); free(s->estarts); free(s->esorted);
    free(s->sdf_tex); free(s->sdf_info); free(s->sdf_xf);
    free(s->pin_v); free(s->pin_t); free(s->pin_w); free(s->jac); free(s->pair_vt); free(s->pair_ee);
    free(s->cedges); free(s->cedge_flags); free(s->cstarts); free(s->csorted); free(s->wtex);
    free(s->s_cur); free(s->s_tgt); free(s->lam_s);
    free(s->dyn_rest); free(s->dyn_w); free(s->dyn_tw); free(s->dyn_bind); free(s->dyn_tris); free(s->dyn_skinned);
    free(s->dyn_bi); free(s->dyn_src); free(s->dyn_bmask); free(s->dyn_params); free(s->dyn_params_dev);
    free(s->dyn_dirty); free(s->normals);
    free(s);
}

const orc_mesh *orc_sim_mesh(const orc_sim *s) { return s->m; }
void orc_sim_set_gravity(orc_sim *s, float gx, float gy, float gz) { s->gravity[0] = gx; s->gravity[1] = gy; s->gravity[2] = gz; }
void orc_sim_set_damping(orc_sim *s, float d) { s->damping = d > 0.0f ? d : 0.0f; }
void orc_sim_set_damping_all(orc_sim *s, float t, float c, float sh, float b) {
    s->tension_damping = t > 0 ? t : 0; s->compression_damping = c > 0 ? c : 0;
    s->shear_damping = sh > 0 ? sh : 0; s->bending_damping = b > 0 ? b : 0;
}
void orc_sim_set_iterations(orc_sim *s, uint32_t it) { s->iterations = it; }
void orc_sim_set_self_collision(orc_sim *s, int on) { s->self_collision = on; }
/* mod.rs:920-953 */
void orc_sim_set_self_collision_options(orc_sim *s, float relief, float ratio, int exclude_neighbors,
                                        int normal_untangling, uint32_t max_iterations) {
    s->relief_factor = relief; s->max_disp_ratio = ratio;
    s->enable_relief = relief < 0.999f;
    s->exclude_neighbors = exclude_neighbors;
    s->normal_untangling = normal_untangling != 0;
    s->max_search_iterations = max_iterations;
}
/* effective_normal_untangling (simulation/mod.rs:885-890) */
static int untangle_on(const orc_sim *s) { return s->normal_untangling && s->m->has_multiple_layers; }
void orc_sim_set_hash_cell(orc_sim *s, float cell) { if (cell > 0.0f) s->cell_size = cell; }
/* raw compliance override per constraint kind (no reference counterpart: the
 * reference maps stiffness >= 5000 N/m to zero compliance) */
void orc_sim_set_compliances_raw(orc_sim *s, float stretch, float shear, float bend) {
    for (uint32_t k = 0; k < s->m->nd; k++) {
        if (s->m->d_type[k] == 1) { s->d_tc[k] = shear; s->d_cc[k] = shear; }
        else { s->d_tc[k] = stretch; s->d_cc[k] = stretch; }
    }
    for (uint32_t k = 0; k < s->m->nb; k++) s->b_comp[k] = bend;
}
static void add_col(orc_sim *s, collider c) {
    s->cols = (collider *)realloc(s->cols, (s->ncols + 1) * sizeof(collider));
    s->cols[s->ncols++] = c;
}
void orc_sim_add_sphere(orc_sim *s, const float *c, float r, float fr, float rest) {
    collider k = {0, fr, rest, mk(c[0], c[1], c[2]), r, mk(0, 0, 0)}; add_col(s, k);
}
void orc_sim_add_capsule(orc_sim *s, const float *a, const float *b, float r, float fr, float rest) {
    collider k = {1, fr, rest, mk(a[0], a[1], a[2]), r, mk(b[0], b[1], b[2])}; add_col(s, k);
}
void orc_sim_add_plane(orc_sim *s, const float *pt, const float *n, float fr, float rest) {
    collider k = {2, fr, rest, mk(pt[0], pt[1], pt[2]), 0, mk(n[0], n[1], n[2])}; add_col(s, k);
}
static void build_collider_edges(orc_sim *s);
void orc_sim_clear_colliders(orc_sim *s) { s->ncols = 0; s->ntris = 0; build_collider_edges(s); }

/* set_bone_sdf_colliders (lib.rs:2235-2285): texels as pack2x16float words, Z, Y, X order; 20
 * floats of collider info each; transforms start as identity (simulation/mod.rs:1700-1719). */
void orc_sim_set_sdf(orc_sim *s, uint32_t w, uint32_t h, uint32_t d, const uint32_t *tex, const float *info, uint32_t nb) {
    free(s->sdf_tex); free(s->sdf_info); free(s->sdf_xf);
    s->sdf_tex = NULL; s->sdf_info = NULL; s->sdf_xf = NULL; s->sdf_nb = 0;
    if (!nb) return;
    const size_t nvox = (size_t)w * h * d;
    s->sdf_tex = (uint32_t *)malloc(4 * nvox);
    memcpy(s->sdf_tex, tex, 4 * nvox);
    s->sdf_info = (float *)malloc(sizeof(float) * 20 * nb);
    memcpy(s->sdf_info, info, sizeof(float) * 20 * nb);
    s->sdf_xf = (float *)calloc((size_t)32 * nb, sizeof(float));
    for (uint32_t b = 0; b < nb; b++)
        for (int m = 0; m < 2; m++)
            for (int k = 0; k < 4; k++) s->sdf_xf[(size_t)b * 32 + m * 16 + k * 5] = 1.0f;
    s->sdf_w = w; s->sdf_h = h; s->sdf_d = d; s->sdf_nb = nb;
}
/* update_bone_transforms (lib.rs:2287-2330 -> simulation/mod.rs:1865-1938): world and inverse-world
 * 4x4 matrices, uploaded as given (WGSL reads each given row as a column); each collider's world
 * bounding sphere goes to info slots 3, 7, 11 (centre) and 15 (radius): the centre of the local box
 * (slots 0-2, 4-6) through the world matrix, radius = half the box diagonal times the largest
 * column length, plus 0.1 m. */
void orc_sim_set_bone_transforms(orc_sim *s, const float *world, const float *inv, uint32_t n) {
    for (uint32_t b = 0; b < n && b < s->sdf_nb; b++) {
        memcpy(&s->sdf_xf[(size_t)b * 32], world + (size_t)b * 16, 64);
        memcpy(&s->sdf_xf[(size_t)b * 32 + 16], inv + (size_t)b * 16, 64);
        float *I = &s->sdf_info[(size_t)b * 20];
        const float *M = world + (size_t)b * 16;   /* M[4r + c] = m[r][c] */
        const float cl[3] = {(I[0] + I[4]) * 0.5f, (I[1] + I[5]) * 0.5f, (I[2] + I[6]) * 0.5f};
        const float dx = I[4] - I[0], dy = I[5] - I[1], dz = I[6] - I[2];
        const float rl = 0.5f * sqrtf(dx * dx + dy * dy + dz * dz);
        float cw[3], sc[3];
        for (int k = 0; k < 3; k++) cw[k] = M[k] * cl[0] + M[4 + k] * cl[1] + M[8 + k] * cl[2] + M[12 + k];
        for (int r = 0; r < 3; r++) sc[r] = sqrtf(M[4 * r] * M[4 * r] + M[4 * r + 1] * M[4 * r + 1] + M[4 * r + 2] * M[4 * r + 2]);
        float ms = sc[0];
        if (sc[1] > ms) ms = sc[1];
        if (sc[2] > ms) ms = sc[2];
        I[3] = cw[0]; I[7] = cw[1]; I[11] = cw[2]; I[15] = rl * ms + 0.1f;
    }
}

void orc_bone_bake_tiles(const uint32_t *params, uint32_t nbake, const float *tris, uint32_t *tex);

/* clear_bone_sdf_colliders (simulation/mod.rs:1941-1948) */
void orc_sim_clear_bone_sdf(orc_sim *s) {
    orc_sim_set_sdf(s, 0, 0, 0, NULL, NULL, 0);
    s->dyn_ready = s->dyn_enabled = 0;
}

static int cmp_u64(const void *a, const void *b) {
    const uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
    return x < y ? -1 : x > y;
}
static int has_u64(const uint64_t *k, size_t n, uint64_t key) {  /* binary search in sorted keys */
    size_t lo = 0, hi = n;
    while (lo < hi) { size_t mid = (lo + hi) / 2; if (k[mid] < key) lo = mid + 1; else hi = mid; }
    return lo < n && k[lo] == key;
}

/* setup_dynamic_bone_sdf (lib.rs:2362-2588 -> simulation/mod.rs:1998-2312).  Returns 0, or 1 for the
 * inputs the build rejects (zero sizes, too few bind matrices, indices out of range, tiles outside the
 * atlas).  rest: nv x 3; bi / bw: nv x 4; src: nt x 4 (i0, i1, i2, bone); tw: nt x 3; bind: nbind x 16. */
int orc_sim_setup_dynamic_bone_sdf(orc_sim *s, uint32_t w, uint32_t h, uint32_t d, uint32_t res, const float *infos,
                                   uint32_t nb, const float *rest, uint32_t nv, const uint32_t *bi, const float *bw,
                                   const uint32_t *src, const float *tw, uint32_t nt, const float *bind,
                                   uint32_t nbind, uint32_t interval) {
    if (!res || !w || !h || !d || nbind < nb) return 1;
    for (size_t i = 0; i < (size_t)nv * 4; i++) if (bw[i] > 0.0f && bi[i] >= nb) return 1;
    for (uint32_t t = 0; t < nt; t++)
        if (src[4 * t] >= nv || src[4 * t + 1] >= nv || src[4 * t + 2] >= nv || src[4 * t + 3] >= nb) return 1;
    if (!nv || !nt || !nb) { s->dyn_interval = interval ? interval : 1; return 0; }
    const uint32_t cols = w / res ? w / res : 1, rows = h / res ? h / res : 1;
    for (uint32_t b = 0; b < nb; b++) {
        const uint32_t l = b / (cols * rows), rem = b % (cols * rows), r = rem / cols, c = rem % cols;
        if ((uint64_t)(c + 1) * res > w || (uint64_t)(r + 1) * res > h || (uint64_t)(l + 1) * res > d) return 1;
    }
    if (w > 2048 || h > 2048 || d > 2048) return 1;
    /* undirected edges used by exactly one distinct triangle (the sources deduplicated by their sorted
     * vertex triple, collect_open_boundary_undirected_edges, sdf_baker.rs:203-240) */
    uint64_t *tk = (uint64_t *)malloc(sizeof(uint64_t) * nt);
    uint32_t *first = (uint32_t *)malloc(sizeof(uint32_t) * nt);
    uint64_t *ek = (uint64_t *)malloc(sizeof(uint64_t) * 3 * (size_t)nt);
    size_t ne = 0;
    for (uint32_t t = 0; t < nt; t++) {  /* first occurrence of each sorted triple (n^2 for clarity) */
        uint32_t a = src[4 * t], b2 = src[4 * t + 1], c = src[4 * t + 2], x;
        if (a > b2) { x = a; a = b2; b2 = x; }
        if (b2 > c) { x = b2; b2 = c; c = x; }
        if (a > b2) { x = a; a = b2; b2 = x; }
        tk[t] = ((uint64_t)a << 42) ^ ((uint64_t)b2 << 21) ^ (uint64_t)c;
        first[t] = 1;
        for (uint32_t u = 0; u < t; u++) {
            uint32_t p = src[4 * u], q = src[4 * u + 1], r2 = src[4 * u + 2], y;
            if (p > q) { y = p; p = q; q = y; }
            if (q > r2) { y = q; q = r2; r2 = y; }
            if (p > q) { y = p; p = q; q = y; }
            if (p == a && q == b2 && r2 == c) { first[t] = 0; break; }
        }
        const uint32_t i0 = src[4 * t], i1 = src[4 * t + 1], i2 = src[4 * t + 2];
        if (!first[t] || i0 > 0x7FFFFFFFu || i1 > 0x7FFFFFFFu || i2 > 0x7FFFFFFFu || i0 == i1 || i1 == i2 || i2 == i0)
            continue;
        const uint32_t e[3][2] = {{i0, i1}, {i1, i2}, {i2, i0}};
        for (int k = 0; k < 3; k++) {
            const uint32_t lo = e[k][0] < e[k][1] ? e[k][0] : e[k][1], hi = e[k][0] < e[k][1] ? e[k][1] : e[k][0];
            ek[ne++] = ((uint64_t)lo << 32) | hi;
        }
    }
    qsort(ek, ne, sizeof(uint64_t), cmp_u64);
    uint64_t *open = (uint64_t *)malloc(sizeof(uint64_t) * (ne ? ne : 1));
    size_t no = 0;
    for (size_t i = 0; i < ne;) {
        size_t j = i;
        while (j < ne && ek[j] == ek[i]) j++;
        if (j - i == 1) open[no++] = ek[i];
        i = j;
    }
    free(tk); free(first); free(ek);
    free(s->dyn_rest); free(s->dyn_w); free(s->dyn_tw); free(s->dyn_bind); free(s->dyn_tris); free(s->dyn_skinned);
    free(s->dyn_bi); free(s->dyn_src); free(s->dyn_bmask); free(s->dyn_params); free(s->dyn_params_dev);
    s->dyn_rest = (float *)malloc(12 * (size_t)nv); memcpy(s->dyn_rest, rest, 12 * (size_t)nv);
    s->dyn_bi = (uint32_t *)malloc(16 * (size_t)nv); memcpy(s->dyn_bi, bi, 16 * (size_t)nv);
    s->dyn_w = (float *)malloc(16 * (size_t)nv); memcpy(s->dyn_w, bw, 16 * (size_t)nv);
    s->dyn_src = (uint32_t *)malloc(16 * (size_t)nt); memcpy(s->dyn_src, src, 16 * (size_t)nt);
    s->dyn_tw = (float *)malloc(12 * (size_t)nt); memcpy(s->dyn_tw, tw, 12 * (size_t)nt);
    s->dyn_bind = (float *)malloc(64 * (size_t)nb); memcpy(s->dyn_bind, bind, 64 * (size_t)nb);
    s->dyn_bmask = (uint32_t *)malloc(4 * (size_t)nt);
    for (uint32_t t = 0; t < nt; t++) {
        const uint32_t i0 = src[4 * t], i1 = src[4 * t + 1], i2 = src[4 * t + 2];
        uint32_t m = 0;
        if (i0 <= 0x7FFFFFFFu && i1 <= 0x7FFFFFFFu && i2 <= 0x7FFFFFFFu && i0 != i1 && i1 != i2 && i2 != i0) {
            const uint32_t e[3][2] = {{i0, i1}, {i1, i2}, {i2, i0}};
            for (int k = 0; k < 3; k++) {
                const uint32_t lo = e[k][0] < e[k][1] ? e[k][0] : e[k][1], hi = e[k][0] < e[k][1] ? e[k][1] : e[k][0];
                if (has_u64(open, no, ((uint64_t)lo << 32) | hi)) m |= 1u << k;
            }
        }
        s->dyn_bmask[t] = m;
    }
    free(open);
    /* bake parameters: the bone's local box, its triangle range (first index, count) and tile */
    s->dyn_params = (uint32_t *)calloc(16 * (size_t)nb, 4);
    for (uint32_t b = 0; b < nb; b++) {
        uint32_t *q = s->dyn_params + 16 * (size_t)b;
        const float *I = infos + 20 * (size_t)b;
        uint32_t start = 0, count = 0;
        for (uint32_t t = 0; t < nt; t++)
            if (src[4 * t + 3] == b) { if (!count) start = t; count++; }
        memcpy(q, I, 12); q[3] = start; memcpy(q + 4, I + 4, 12); q[7] = count;
        const uint32_t l = b / (cols * rows), rem = b % (cols * rows);
        q[8] = rem % cols; q[9] = rem / cols; q[10] = l; q[11] = res; q[12] = w; q[13] = h; q[14] = d; q[15] = 0;
    }
    s->dyn_params_dev = (uint32_t *)malloc(64 * (size_t)nb);
    memcpy(s->dyn_params_dev, s->dyn_params, 64 * (size_t)nb);
    s->dyn_tris = (float *)calloc(16 * (size_t)nt, 4);
    s->dyn_skinned = (float *)calloc(3 * (size_t)nv, 4);
    /* the texture (zero until the first bake), the infos, identity transforms appended */
    const uint32_t have = s->sdf_nb;
    float *xf = (float *)calloc(32 * (size_t)nb, 4);
    for (uint32_t b = 0; b < nb; b++) {
        if (b < have) memcpy(xf + 32 * (size_t)b, s->sdf_xf + 32 * (size_t)b, 128);
        else for (int m = 0; m < 2; m++) for (int k = 0; k < 4; k++) xf[32 * (size_t)b + 16 * m + 5 * k] = 1.0f;
    }
    uint32_t *zero = (uint32_t *)calloc((size_t)w * h * d, 4);
    orc_sim_set_sdf(s, w, h, d, zero, infos, nb);
    memcpy(s->sdf_xf, xf, 128 * (size_t)nb);
    free(zero); free(xf);
    s->dyn_nb = nb; s->dyn_nv = nv; s->dyn_nt = nt; s->dyn_res = res;
    s->dyn_ready = 1; s->dyn_enabled = 1; s->dyn_baked = 0; s->dyn_dirty_set = 0; s->dyn_ndirty = 0;
    s->dyn_interval = interval ? interval : 1;
    return 0;
}
void orc_sim_set_dynamic_dirty(orc_sim *s, const uint32_t *bones, uint32_t n, int has_list) {
    free(s->dyn_dirty);
    s->dyn_dirty = NULL;
    s->dyn_dirty_set = has_list != 0;
    s->dyn_ndirty = has_list ? n : 0;
    if (s->dyn_ndirty) { s->dyn_dirty = (uint32_t *)malloc(4 * (size_t)n); memcpy(s->dyn_dirty, bones, 4 * (size_t)n); }
}
void orc_sim_set_dynamic_enabled(orc_sim *s, int on) { s->dyn_enabled = on && s->dyn_ready; }
void orc_sim_set_dynamic_interval(orc_sim *s, uint32_t k) { s->dyn_interval = k ? k : 1; }
/* the skinned positions (nv x 3) and bake triangles (nt x 16 floats) of the last update */
void orc_sim_dynamic_arrays(const orc_sim *s, float *skinned, float *tris) {
    if (skinned) memcpy(skinned, s->dyn_skinned, 12 * (size_t)s->dyn_nv);
    if (tris) memcpy(tris, s->dyn_tris, 64 * (size_t)s->dyn_nt);
}
void orc_sim_sdf_arrays(const orc_sim *s, uint32_t *tex, float *info) {
    if (tex && s->sdf_nb) memcpy(tex, s->sdf_tex, 4 * (size_t)s->sdf_w * s->sdf_h * s->sdf_d);
    if (info && s->sdf_nb) memcpy(info, s->sdf_info, 80 * (size_t)s->sdf_nb);
}

/* (m * (x, y, z, 1)).r for a matrix stored as WGSL reads it (column c at m[4c .. 4c+3]) */
static void wgsl_apply(const float *m, const real *v, real *o) {
    for (int r = 0; r < 3; r++) o[r] = (real)m[r] * v[0] + (real)m[4 + r] * v[1] + (real)m[8 + r] * v[2] + (real)m[12 + r];
}

/* dispatch_dynamic_bone_sdf_update (simulation/mod.rs:2320-2457): at the start of a step frame */
static void dynamic_sdf_update(orc_sim *s) {
    if (!s->dyn_enabled) return;
    s->dyn_counter += 1u;
    if ((s->dyn_counter - 1u) % s->dyn_interval != 0u) return;
    const uint32_t nb = s->dyn_nb;
    if (s->sdf_nb < nb) return;
    /* skin matrices S = W B in the reference's host f32 arithmetic, laid out for WGSL (transposed) */
    float *skin = (float *)malloc(64 * (size_t)nb);
    for (uint32_t b = 0; b < nb; b++) {
        const float *W = s->sdf_xf + 32 * (size_t)b, *B = s->dyn_bind + 16 * (size_t)b;
        for (int i = 0; i < 4; i++)
            for (int j = 0; j < 4; j++) {  /* w_row[i][k] = W[4k + i]; s_row[i][j]; stored at [4j + i] */
                const float v = W[i] * B[j] + W[4 + i] * B[4 + j] + W[8 + i] * B[8 + j] + W[12 + i] * B[12 + j];
                skin[16 * (size_t)b + 4 * j + i] = v;
            }
    }
    /* skinning.wgsl:17-43 */
    for (uint32_t i = 0; i < s->dyn_nv; i++) {
        const real pos[3] = {s->dyn_rest[3 * i], s->dyn_rest[3 * i + 1], s->dyn_rest[3 * i + 2]};
        real p[3] = {0, 0, 0}, ws = 0;
        for (int k = 0; k < 4; k++) {
            const real wk = s->dyn_w[4 * i + k];
            if (wk > 0) {
                real t[3];
                wgsl_apply(skin + 16 * (size_t)s->dyn_bi[4 * i + k], pos, t);
                for (int c = 0; c < 3; c++) p[c] = p[c] + wk * t[c];
                ws = ws + wk;
            }
        }
        if (ws <= (real)1e-6f) for (int c = 0; c < 3; c++) p[c] = pos[c];
        for (int c = 0; c < 3; c++) s->dyn_skinned[3 * i + c] = (float)p[c];
    }
    /* prep_bone_triangles.wgsl:31-67 */
    for (uint32_t t = 0; t < s->dyn_nt; t++) {
        const uint32_t *q = s->dyn_src + 4 * (size_t)t;
        const float *inv = s->sdf_xf + 32 * (size_t)q[3] + 16;
        real P[3][3];
        for (int c = 0; c < 3; c++) {
            const real v[3] = {s->dyn_skinned[3 * q[c]], s->dyn_skinned[3 * q[c] + 1], s->dyn_skinned[3 * q[c] + 2]};
            wgsl_apply(inv, v, P[c]);
        }
        const real e1[3] = {P[1][0] - P[0][0], P[1][1] - P[0][1], P[1][2] - P[0][2]};
        const real e2[3] = {P[2][0] - P[0][0], P[2][1] - P[0][1], P[2][2] - P[0][2]};
        const real cr[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
        const real ls = cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2];
        real n[3] = {0, 1, 0};
        if (ls > (real)1e-12f) { const real il = 1 / (real)sqrt((double)ls); for (int c = 0; c < 3; c++) n[c] = cr[c] * il; }
        float *o = s->dyn_tris + 16 * (size_t)t;
        for (int c = 0; c < 3; c++) {
            for (int k = 0; k < 3; k++) o[4 * c + k] = (float)P[c][k];
            o[4 * c + 3] = s->dyn_tw[3 * t + c];
        }
        for (int k = 0; k < 3; k++) o[12 + k] = (float)n[k];
        o[15] = (float)s->dyn_bmask[t];
    }
    free(skin);
    uint32_t count = nb;
    if (s->dyn_baked && s->dyn_dirty_set) {
        if (!s->dyn_ndirty) return;
        count = s->dyn_ndirty;
        uint32_t j = 0;
        for (uint32_t k = 0; k < s->dyn_ndirty; k++)
            if (s->dyn_dirty[k] < nb && j < nb) memcpy(s->dyn_params_dev + 16 * (size_t)j++, s->dyn_params + 16 * (size_t)s->dyn_dirty[k], 64);
    } else {
        memcpy(s->dyn_params_dev, s->dyn_params, 64 * (size_t)nb);
    }
    orc_bone_bake_tiles(s->dyn_params_dev, count < nb ? count : nb, s->dyn_tris, s->sdf_tex);
    s->dyn_baked = 1;
}

typedef struct { uint32_t code, idx; } tkey;
static int cmp_tkey(const void *a, const void *b) {
    const tkey *x = (const tkey *)a, *y = (const tkey *)b;
    if (x->code != y->code) return x->code < y->code ? -1 : 1;
    return x->idx < y->idx ? -1 : x->idx > y->idx;
}
/* set_mesh_collider_triangles (lib.rs:2175-2233) + upload_colliders Morton sort and
 * bounding spheres (simulation/mod.rs:2582-2667). attrs: NULL or 4 floats per triangle. */
void orc_sim_set_mesh_triangles(orc_sim *s, const float *t9, uint32_t nt, const float *attrs, float friction,
                                float thickness, float restitution, int single_sided) {
    mtri *T = (mtri *)malloc((nt ? nt : 1) * sizeof(mtri));
    for (uint32_t t = 0; t < nt; t++) {
        for (int a = 0; a < 3; a++) { T[t].p0[a] = t9[9 * t + a]; T[t].p1[a] = t9[9 * t + 3 + a]; T[t].p2[a] = t9[9 * t + 6 + a]; }
        if (attrs) {
            T[t].friction = attrs[4 * t]; T[t].thickness = attrs[4 * t + 1]; T[t].restitution = attrs[4 * t + 2];
            float r = roundf(attrs[4 * t + 3]);
            T[t].flags = r > 0.0f ? (r >= 4294967296.0f ? 0xFFFFFFFFu : (uint32_t)r) : 0u;
        } else {
            T[t].friction = friction; T[t].thickness = thickness; T[t].restitution = restitution;
            T[t].flags = single_sided ? 1u : 0u;
        }
    }
    float bmin[3] = {3.4028235e38f, 3.4028235e38f, 3.4028235e38f}, bmax[3] = {-3.4028235e38f, -3.4028235e38f, -3.4028235e38f};
    float *cen = (float *)malloc((nt ? nt : 1) * 12);
    for (uint32_t t = 0; t < nt; t++)
        for (int a = 0; a < 3; a++) {
            cen[3 * t + a] = (T[t].p0[a] + T[t].p1[a] + T[t].p2[a]) / 3.0f;
            bmin[a] = fminf(bmin[a], cen[3 * t + a]); bmax[a] = fmaxf(bmax[a], cen[3 * t + a]);
        }
    float inv[3];
    for (int a = 0; a < 3; a++) inv[a] = 1.0f / fmaxf(bmax[a] - bmin[a], 1e-5f);
    tkey *k = (tkey *)malloc((nt ? nt : 1) * sizeof(tkey));
    for (uint32_t t = 0; t < nt; t++) { k[t].code = morton_code_3d(cen + 3 * t, bmin, inv); k[t].idx = t; }
    qsort(k, nt, sizeof(tkey), cmp_tkey);
    free(s->tris); free(s->tri_bounds);
    s->tris = (mtri *)malloc((nt ? nt : 1) * sizeof(mtri));
    s->tri_bounds = (float *)malloc((nt ? nt : 1) * 16);
    for (uint32_t t = 0; t < nt; t++) {
        mtri m = T[k[t].idx];
        s->tris[t] = m;
        float cx = (m.p0[0] + m.p1[0] + m.p2[0]) / 3.0f, cy = (m.p0[1] + m.p1[1] + m.p2[1]) / 3.0f,
              cz = (m.p0[2] + m.p1[2] + m.p2[2]) / 3.0f;
        float d0 = (m.p0[0] - cx) * (m.p0[0] - cx) + (m.p0[1] - cy) * (m.p0[1] - cy) + (m.p0[2] - cz) * (m.p0[2] - cz);
        float d1 = (m.p1[0] - cx) * (m.p1[0] - cx) + (m.p1[1] - cy) * (m.p1[1] - cy) + (m.p1[2] - cz) * (m.p1[2] - cz);
        float d2 = (m.p2[0] - cx) * (m.p2[0] - cx) + (m.p2[1] - cy) * (m.p2[1] - cy) + (m.p2[2] - cz) * (m.p2[2] - cz);
        s->tri_bounds[4 * t] = cx; s->tri_bounds[4 * t + 1] = cy; s->tri_bounds[4 * t + 2] = cz;
        s->tri_bounds[4 * t + 3] = sqrtf(fmaxf(fmaxf(d0, d1), d2)) + m.thickness;
    }
    s->ntris = nt;
    free(T); free(cen); free(k);
    build_collider_edges(s);
}

/* extract_collider_edges (mesh.rs:1492-1567) over the Morton-ordered triangles: the unique edges
 * keyed by the bits of their end points (smaller triple first), in order of first appearance;
 * thickness = max, friction averaged pairwise in triangle order, normal = normalised sum of the
 * triangle normals, single-sided only if all are.  Then build_collider_edge_grid
 * (spatial_hash.rs:473-529): each edge in every bucket of its half-cell samples. */
typedef struct { uint32_t key[6]; uint32_t seq; } ekey_t;
static int cmp_ekey(const void *a, const void *b) {
    const ekey_t *x = (const ekey_t *)a, *y = (const ekey_t *)b;
    for (int i = 0; i < 6; i++) if (x->key[i] != y->key[i]) return x->key[i] < y->key[i] ? -1 : 1;
    return x->seq < y->seq ? -1 : x->seq > y->seq;
}
static int cmp_u32pair(const void *a, const void *b) {
    const uint32_t *x = (const uint32_t *)a, *y = (const uint32_t *)b;
    return x[0] < y[0] ? -1 : x[0] > y[0];
}
static uint32_t fbits(float x) { uint32_t u; memcpy(&u, &x, 4); return u; }
static int32_t f2i_sat_f(float x) {
    if (x != x) return 0;
    if (x >= 2147483520.0f) return 2147483647;
    if (x <= -2147483648.0f) return INT32_MIN;
    return (int32_t)x;
}
static void build_collider_edges(orc_sim *s) {
    uint32_t nt = s->ntris, n3 = 3 * nt;
    free(s->cedges); free(s->cedge_flags); free(s->cstarts); free(s->csorted);
    ekey_t *K = (ekey_t *)malloc((n3 ? n3 : 1) * sizeof(ekey_t));
    float (*tn)[3] = (float (*)[3])malloc((nt ? nt : 1) * 12);
    for (uint32_t t = 0; t < nt; t++) {
        const mtri *T = &s->tris[t];
        float e1[3] = {T->p1[0] - T->p0[0], T->p1[1] - T->p0[1], T->p1[2] - T->p0[2]};
        float e2[3] = {T->p2[0] - T->p0[0], T->p2[1] - T->p0[1], T->p2[2] - T->p0[2]};
        float cr[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
        float cl = sqrtf(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]);
        for (int a = 0; a < 3; a++) tn[t][a] = cl > 1e-12f ? cr[a] / cl : 0.0f;
        const float *P3[3] = {T->p0, T->p1, T->p2};
        for (int k = 0; k < 3; k++) {
            const float *a = P3[k], *b = P3[(k + 1) % 3];
            uint32_t ab[3] = {fbits(a[0]), fbits(a[1]), fbits(a[2])}, bb[3] = {fbits(b[0]), fbits(b[1]), fbits(b[2])};
            int a_first = ab[0] != bb[0] ? ab[0] < bb[0] : ab[1] != bb[1] ? ab[1] < bb[1] : ab[2] < bb[2];
            ekey_t *e = &K[3 * t + k];
            memcpy(e->key, a_first ? ab : bb, 12); memcpy(e->key + 3, a_first ? bb : ab, 12);
            e->seq = 3 * t + k;
        }
    }
    qsort(K, n3, sizeof(ekey_t), cmp_ekey);
    /* groups in order of first appearance */
    uint32_t *first = (uint32_t *)malloc(2 * (n3 ? n3 : 1) * 4);  /* (first seq, group start) */
    uint32_t ng = 0;
    for (uint32_t i = 0; i < n3; i++)
        if (i == 0 || memcmp(K[i].key, K[i - 1].key, 24) != 0) { first[2 * ng] = K[i].seq; first[2 * ng + 1] = i; ng++; }
    qsort(first, ng, 8, cmp_u32pair);
    s->cedges = (float *)malloc((ng ? ng : 1) * 11 * sizeof(float));
    s->cedge_flags = (uint32_t *)malloc((ng ? ng : 1) * 4);
    for (uint32_t g = 0; g < ng; g++) {
        uint32_t i0 = first[2 * g + 1];
        float *E = s->cedges + 11 * (size_t)g;
        float ns[3] = {0, 0, 0};
        uint32_t flags = 0;
        for (uint32_t i = i0; i < n3 && memcmp(K[i].key, K[i0].key, 24) == 0; i++) {
            uint32_t t = K[i].seq / 3, k = K[i].seq % 3;
            const mtri *T = &s->tris[t];
            if (i == i0) {
                const float *P3[3] = {T->p0, T->p1, T->p2};
                for (int a = 0; a < 3; a++) { E[a] = P3[k][a]; E[4 + a] = P3[(k + 1) % 3][a]; }
                E[3] = T->thickness; E[7] = T->friction;
                ns[0] = tn[t][0]; ns[1] = tn[t][1]; ns[2] = tn[t][2];
                flags = T->flags & 1u;
            } else {
                E[3] = fmaxf(E[3], T->thickness);
                E[7] = (E[7] + T->friction) * 0.5f;
                ns[0] += tn[t][0]; ns[1] += tn[t][1]; ns[2] += tn[t][2];
                flags &= T->flags & 1u;
            }
        }
        float nl = sqrtf(ns[0] * ns[0] + ns[1] * ns[1] + ns[2] * ns[2]);
        for (int a = 0; a < 3; a++) E[8 + a] = nl > 1e-6f ? ns[a] / nl : 0.0f;
        s->cedge_flags[g] = flags;
    }
    s->nce = ng;
    /* grid */
    const uint32_t table = 32768;
    float cell = (float)s->m->cell_size, safe = fmaxf(cell, 1e-4f), pitch = safe * 0.5f;
    s->cstarts = (uint32_t *)calloc(table + 1, 4);
    uint32_t *counts = (uint32_t *)calloc(table, 4);
    size_t cap = 64, ne = 0;
    uint32_t *ent = (uint32_t *)malloc(cap * 8);
    uint32_t *hs = (uint32_t *)malloc(4098 * 4);
    for (uint32_t g = 0; g < ng; g++) {
        const float *E = s->cedges + 11 * (size_t)g;
        float dx = E[4] - E[0], dy = E[5] - E[1], dz = E[6] - E[2];
        float len = sqrtf(dx * dx + dy * dy + dz * dz);
        float cf = ceilf(len / pitch);
        size_t steps = (cf != cf || cf <= 0.0f) ? 0 : (cf >= 1.8446744e19f ? (size_t)-1 : (size_t)cf);
        if (steps < 1) steps = 1;
        if (steps > 4096) steps = 4096;
        uint32_t nh = 0;
        for (size_t k = 0; k <= steps; k++) {
            float t = (float)k / (float)steps;
            float px = E[0] + dx * t, py = E[1] + dy * t, pz = E[2] + dz * t;
            uint32_t c0 = (uint32_t)f2i_sat_f(floorf(px / safe)), c1 = (uint32_t)f2i_sat_f(floorf(py / safe)),
                     c2 = (uint32_t)f2i_sat_f(floorf(pz / safe));
            hs[nh++] = ((c0 * 73856093u) ^ (c1 * 19349663u) ^ (c2 * 83492791u)) % table;
        }
        qsort(hs, nh, 4, cmp_u32);
        for (uint32_t q = 0; q < nh; q++) {
            if (q > 0 && hs[q] == hs[q - 1]) continue;
            if (ne == cap) { cap *= 2; ent = (uint32_t *)realloc(ent, cap * 8); }
            ent[2 * ne] = hs[q]; ent[2 * ne + 1] = g; ne++;
            counts[hs[q]]++;
        }
    }
    uint32_t sum = 0;
    for (uint32_t h = 0; h < table; h++) { s->cstarts[h] = sum; sum += counts[h]; }
    s->cstarts[table] = sum;
    s->csorted = (uint32_t *)malloc((ne ? ne : 1) * 4);
    memcpy(counts, s->cstarts, table * 4);
    for (size_t i = 0; i < ne; i++) s->csorted[counts[ent[2 * i]]++] = ent[2 * i + 1];
    free(K); free(tn); free(first); free(counts); free(ent); free(hs);
}
void orc_sim_set_edge_collision(orc_sim *s, int on, float scale, float offset) {
    s->edge_collision = on;
    s->edge_scale = scale < 1.0f ? 1.0f : scale;
    s->edge_offset = offset < 0.0f ? 0.0f : offset;
}
uint32_t orc_sim_collider_edges(const orc_sim *s, float *out, uint32_t *flags) {
    if (out) memcpy(out, s->cedges, (size_t)s->nce * 11 * sizeof(float));
    if (flags) memcpy(flags, s->cedge_flags, (size_t)s->nce * 4);
    return s->nce;
}
void orc_sim_reset(orc_sim *s) {
    for (uint32_t i = 0; i < s->nv; i++) { s->x[i] = s->x0[i]; s->p[i] = s->x0[i]; s->v[i] = mk(0, 0, 0); s->w[i] = s->w0[i]; }
}
void orc_sim_set_state(orc_sim *s, const double *x, const double *p, const double *v) {
    for (uint32_t i = 0; i < s->nv; i++) {
        if (x) s->x[i] = mk((real)x[3 * i], (real)x[3 * i + 1], (real)x[3 * i + 2]);
        if (p) s->p[i] = mk((real)p[3 * i], (real)p[3 * i + 1], (real)p[3 * i + 2]);
        if (v) s->v[i] = mk((real)v[3 * i], (real)v[3 * i + 1], (real)v[3 * i + 2]);
    }
}
/* which: 0 = position x, 1 = predicted p, 2 = velocity */
void orc_sim_get(const orc_sim *s, int which, double *out) {
    const v3 *a = which == 0 ? s->x : which == 1 ? s->p : s->v;
    for (uint32_t i = 0; i < s->nv; i++) { out[3 * i] = a[i].x; out[3 * i + 1] = a[i].y; out[3 * i + 2] = a[i].z; }
}
void orc_sim_get_lambdas(const orc_sim *s, double *ld, double *lb) {
    for (uint32_t k = 0; k < s->m->nd; k++) ld[k] = s->lam_d[k];
    for (uint32_t k = 0; k < s->m->nb; k++) lb[k] = s->lam_b[k];
}
/* Test hook: overwrite the lambdas (distance in colour order, then bending); either may be NULL. */
void orc_sim_set_lambdas(orc_sim *s, const double *ld, const double *lb) {
    for (uint32_t k = 0; ld && k < s->m->nd; k++) s->lam_d[k] = (real)ld[k];
    for (uint32_t k = 0; lb && k < s->m->nb; k++) s->lam_b[k] = (real)lb[k];
}

/* The pin table (replaced whole; the host policy of simulation/mod.rs:2768-2875 that builds it,
 * and the inverse masses it implies, is restated in oracle.py) and the inverse masses. */
void orc_sim_set_pin_table(orc_sim *s, const uint32_t *v, const float *t, const float *w, uint32_t n) {
    free(s->pin_v); free(s->pin_t); free(s->pin_w);
    s->npins = n;
    s->pin_v = (uint32_t *)malloc((n ? n : 1) * sizeof(uint32_t));
    s->pin_t = (v3 *)malloc((n ? n : 1) * sizeof(v3));
    s->pin_w = (real *)malloc((n ? n : 1) * sizeof(real));
    for (uint32_t k = 0; k < n; k++) {
        s->pin_v[k] = v[k];
        s->pin_t[k] = mk(t[3 * k], t[3 * k + 1], t[3 * k + 2]);
        s->pin_w[k] = w[k];
    }
}
void orc_sim_set_inv_masses(orc_sim *s, const uint32_t *idx, const float *val, uint32_t n) {
    for (uint32_t k = 0; k < n; k++)
        if (idx[k] < s->nv) s->w[idx[k]] = val[k];
}
/* which: 0 = the working inverse masses, 1 = those of the mesh at creation */
void orc_sim_get_inv_masses(const orc_sim *s, int which, float *out) {
    for (uint32_t i = 0; i < s->nv; i++) out[i] = (float)(which ? s->w0[i] : s->w[i]);
}

/* pin.wgsl:55-92, every pin of the table */
static void k_pins(orc_sim *s, real dt_sub) {
    for (uint32_t k = 0; k < s->npins; k++) {
        real w = rclamp(s->pin_w[k], (real)0, (real)1); /* NaN -> 0 */
        if (w <= 0) continue;
        uint32_t v = s->pin_v[k];
        real inv_m = s->w[v];
        if (inv_m <= 0) { s->p[v] = s->pin_t[k]; continue; }
        v3 d = sub(s->p[v], s->pin_t[k]);
        real dist = len3(d);
        if (!(dist > (real)1e-7f)) continue;
        v3 n = dvs(d, dist);
        real dt = rmax(dt_sub, (real)1e-8f);
        real compliance = (real)1e-3f * ((real)1 - w) / rmax(w, (real)1e-6f);
        real alpha = compliance / (dt * dt);
        real corr = -dist / (inv_m + alpha);
        s->p[v] = add(s->p[v], mul(n, inv_m * corr));
    }
}

/* The wrinkle field's parameters (on, 22 floats as tc_sim_get_wrinkle_field_params) and its RGBA8
 * texture (w x h, rows along v; NULL or a zero size: none). */
void orc_sim_set_wrinkle(orc_sim *s, int on, const float *params, const uint8_t *tex, uint32_t w, uint32_t h) {
    s->wrinkle_on = on;
    memcpy(s->wrinkle, params, sizeof(s->wrinkle));
    free(s->wtex); s->wtex = NULL; s->ww = s->wh = 0;
    if (tex && w && h) {
        s->wtex = (uint8_t *)malloc((size_t)w * h * 4);
        memcpy(s->wtex, tex, (size_t)w * h * 4);
        s->ww = w; s->wh = h;
    }
}
/* atan2 by octant reduction and the odd Taylor series of atan to a^15 on |a| <= tan(pi/8) (the
 * product's atan2_det, operation for operation) */
static float atan2_det(float y, float x) {
    float ax = fabsf(x), ay = fabsf(y);
    float mx = fmaxf(ax, ay), mn = fminf(ax, ay);
    if (!(mx > 0.0f)) return 0.0f;
    float a = mn / mx, off = 0.0f;
    if (a > 0.41421356f) { a = (a - 1.0f) / (a + 1.0f); off = 0.78539816f; }
    float sq = a * a;
    float p = -1.0f / 15.0f;
    p = p * sq + 1.0f / 13.0f; p = p * sq - 1.0f / 11.0f; p = p * sq + 1.0f / 9.0f; p = p * sq - 1.0f / 7.0f;
    p = p * sq + 1.0f / 5.0f; p = p * sq - 1.0f / 3.0f; p = p * sq + 1.0f;
    float r = off + a * p;
    if (ay > ax) r = 1.5707964f - r;
    if (x < 0.0f) r = 3.1415927f - r;
    if (y < 0.0f) r = -r;
    return r;
}
float orc_atan2_det(float y, float x) { return atan2_det(y, x); }
/* wrinkle_field.wgsl:58-154 with a bilinear sample (x = u W - 0.5, u repeating, v clamped) */
static void k_wrinkle(orc_sim *s, real dt_sub) {
    if (!s->wrinkle_on || !s->wtex) return;
    const float *W = s->wrinkle;
    const real EPS = (real)1e-7f, TWO_PI = (real)6.2831855f;
    v3 org = mk(W[0], W[1], W[2]), axis = mk(W[4], W[5], W[6]), nrm = mk(W[8], W[9], W[10]), bin = mk(W[12], W[13], W[14]);
    real infl = W[3], bone_r = W[7], stiff = W[11], blend = W[15];
    real z_min = W[16], z_max = W[17], r_min = W[18], r_max = W[19], vwin = W[20], cwin = W[21];
    int TW = (int)s->ww, TH = (int)s->wh;
    for (uint32_t i = 0; i < s->nv; i++) {
        if (s->w[i] <= 0) continue;
        v3 diff = sub(s->p[i], org);
        real z = dot(diff, axis);
        if (z < z_min || z > z_max) continue;
        v3 rv = sub(diff, mul(axis, z));
        real r = len3(rv);
        if (r <= (real)1e-5f) continue;
        if (infl > 0 && r_max > 0 && r > r_max + infl) continue;
        real theta = (real)atan2_det((float)dot(rv, bin), (float)dot(rv, nrm));
        if (theta < 0) theta += TWO_PI;
        real uq = theta / TWO_PI;
        real u = uq - (real)trunc((double)uq);
        real zl = rmax(z_max - z_min, EPS);
        real vv = rclamp((z - z_min) / zl, 0, 1);
        real fx = u * (real)TW - (real)0.5f, fy = vv * (real)TH - (real)0.5f;
        real x0f = (real)floor((double)fx), y0f = (real)floor((double)fy);
        real tx = fx - x0f, ty = fy - y0f;
        int x0 = (int)x0f, y0 = (int)y0f;
        int xa = ((x0 % TW) + TW) % TW, xb = (((x0 + 1) % TW) + TW) % TW;
        int ya = y0 < 0 ? 0 : (y0 > TH - 1 ? TH - 1 : y0), yb = y0 + 1 < 0 ? 0 : (y0 + 1 > TH - 1 ? TH - 1 : y0 + 1);
        real t[4];
        for (int c = 0; c < 4; c++) {
            real t00 = (real)((float)s->wtex[4 * ((size_t)ya * TW + xa) + c] / 255.0f);
            real t10 = (real)((float)s->wtex[4 * ((size_t)ya * TW + xb) + c] / 255.0f);
            real t01 = (real)((float)s->wtex[4 * ((size_t)yb * TW + xa) + c] / 255.0f);
            real t11 = (real)((float)s->wtex[4 * ((size_t)yb * TW + xb) + c] / 255.0f);
            t[c] = (t00 * (1 - tx) + t10 * tx) * (1 - ty) + (t01 * (1 - tx) + t11 * tx) * ty;
        }
        real rl = rmax(r_max - r_min, EPS);
        real tv_r = r_min + t[2] * rl, tc_r = r_min + t[3] * rl;
        real dt = rmax(dt_sub, (real)1e-8f);
        real a_pull = 0, a_push = 0;
        if (t[0] > 0) {
            real vr = rmax(tv_r, bone_r + (real)0.002f);
            real margin = rmax(0, r - vr);
            a_pull = t[0] * blend * stiff * rclamp(margin / rmax(vwin, (real)0.001f), 0, 1);
        }
        if (t[1] > 0) {
            real ex = rmax(0, tc_r - r);
            a_push = t[1] * blend * stiff * rclamp(ex / rmax(cwin, (real)0.001f), 0, 1);
        }
        real a = a_push - a_pull;
        if (!((real)fabs((double)a) > EPS)) continue;
        v3 dir = r > EPS ? dvs(rv, r) : nrm;
        v3 dv = mul(dir, a * dt);
        s->v[i] = add(s->v[i], dv);
        s->p[i] = add(s->p[i], mul(dv, dt));
    }
}

/* predict.wgsl:33-55 */
static void k_predict(orc_sim *s, real dt_sub) {
    real dt = rmax(dt_sub, (real)1e-8f);
    v3 g = mk(s->gravity[0], s->gravity[1], s->gravity[2]);
    ORC_PARALLEL_FOR
    for (uint32_t i = 0; i < s->nv; i++) {
        if (s->w[i] <= 0) { s->v[i] = mk(0, 0, 0); s->p[i] = s->x[i]; }
        else {
            s->v[i] = add(mul(s->v[i], rmax((real)0, (real)1 - s->damping * dt)), mul(g, dt));
            s->p[i] = add(s->x[i], mul(s->v[i], dt));
        }
    }
}

void orc_sim_set_solver_mode(orc_sim *s, uint32_t mode) { s->solver_mode = mode; }

/* distance.wgsl:60-124 for one constraint */
static void k_distance_one(orc_sim *s, uint32_t k, real dt_sub) {
    const orc_mesh *m = s->m;
    uint32_t a = m->d_v0[k], b = m->d_v1[k];
    real w0 = s->w[a], w1 = s->w[b], wsum = w0 + w1;
    if (wsum <= (real)1e-7f) return;
    v3 p0 = s->p[a], p1 = s->p[b];
    v3 delta = sub(p0, p1);
    real dist = len3(delta);
    if (dist <= (real)1e-7f) return;
    real c = dist - (real)m->d_rest[k];
    real comp = c >= 0 ? s->d_tc[k] : s->d_cc[k];
    if (comp >= (real)1e9f) return;
    v3 dir = dvs(delta, dist);
    real dt = rmax(dt_sub, (real)1e-8f);
    real alpha = comp / (dt * dt);
    real damp = m->d_type[k] == 1 ? s->shear_damping : (c >= 0 ? s->tension_damping : s->compression_damping);
    real tau = damp * (real)1e-5f;
    real gamma = rmin(tau / dt, (real)1e4f);
    v3 rel = sub(sub(p0, s->x[a]), sub(p1, s->x[b]));
    real vproj = dot(dir, rel);
    real lam = s->lam_d[k];
    real dl = (-c - alpha * lam - gamma * vproj) / (((real)1 + gamma) * wsum + alpha);
    s->lam_d[k] = lam + dl;
    v3 c0 = mul(dir, w0 * dl);
    v3 c1 = mul(dir, -w1 * dl);
    s->p[a] = add(s->p[a], c0);
    s->p[b] = add(s->p[b], c1);
}

/* bending.wgsl:59-126 */
static void k_bending_one(orc_sim *s, uint32_t k, real dt_sub) {
    const orc_mesh *m = s->m;
    real comp = s->b_comp[k];
    if (comp >= (real)1e9f) return;
    uint32_t v2 = m->b_v2[k], v3_ = m->b_v3[k];
    real w2 = s->w[v2], w3 = s->w[v3_], wsum = w2 + w3;
    if (wsum <= (real)1e-7f) return;
    v3 p2 = s->p[v2], p3 = s->p[v3_];
    v3 delta = sub(p2, p3);
    real dist = len3(delta);
    if (dist <= (real)1e-7f) return;
    float pad0 = m->b_pad0[k];
    if (pad0 > 0.5f && pad0 < 1.5f) {
        if (dist >= (real)m->b_rest[k]) return;
        real gap = len3(sub(s->p[m->b_v0[k]], s->p[m->b_v1[k]]));
        if (gap > (real)m->b_pad1[k]) return;
    }
    v3 dir = dvs(delta, dist);
    real dt = rmax(dt_sub, (real)1e-8f);
    real alpha = comp / (dt * dt);
    real tau = s->bending_damping * (real)1e-5f;
    real gamma = rmin(tau / dt, (real)1e4f);
    v3 rel = sub(sub(p2, s->x[v2]), sub(p3, s->x[v3_]));
    real vproj = dot(dir, rel);
    real c = dist - (real)m->b_rest[k];
    real lam = s->lam_b[k];
    real dl = (-c - alpha * lam - gamma * vproj) / (((real)1 + gamma) * wsum + alpha);
    s->lam_b[k] = lam + dl;
    s->p[v2] = add(s->p[v2], mul(dir, w2 * dl));
    s->p[v3_] = add(s->p[v3_], mul(dir, -w3 * dl));
}

/* distance_atomic.wgsl:78-158 (solve_distance_atomic) for one constraint: the XPBD step with the
 * Rayleigh term always present, a quarter of it kept in lambda, the corrections added to the
 * per-vertex sums in 1e-6 m fixed point (saturated at +-5e7, truncated toward zero, wrapping adds) */
static inline int32_t jac_fixed(real x) {
    return f2i_sat(rclamp(x * (real)1000000.0f, (real)-50000000.0f, (real)50000000.0f));
}
static inline void jac_add(int32_t *a, int32_t v) { *a = (int32_t)((uint32_t)*a + (uint32_t)v); }
static void k_distance_jacobi_one(orc_sim *s, uint32_t k, real dt_sub) {
    const orc_mesh *m = s->m;
    uint32_t a = m->d_v0[k], b = m->d_v1[k];
    real w0 = s->w[a], w1 = s->w[b], ws = w0 + w1;
    if (ws <= (real)1e-7f) return;
    v3 p0 = s->p[a], p1 = s->p[b];
    v3 delta = sub(p0, p1);
    real dist = len3(delta);
    if (dist <= (real)1e-7f) return;
    real c = dist - (real)m->d_rest[k];
    real comp = c >= 0 ? s->d_tc[k] : s->d_cc[k];
    if (comp >= (real)1e9f) return;
    v3 dir = dvs(delta, dist);
    real dt = rmax(dt_sub, (real)1e-8f);
    real alpha = comp / (dt * dt);
    real dc = m->d_type[k] == 1 ? s->shear_damping : (c >= 0 ? s->tension_damping : s->compression_damping);
    real gamma = rmin(dc * (real)1e-5f / dt, (real)1e4f);
    v3 rel = sub(sub(p0, s->x[a]), sub(p1, s->x[b]));
    real vproj = dot(dir, rel);
    real l = s->lam_d[k];
    real dl = (-c - alpha * l - gamma * vproj) / (((real)1 + gamma) * ws + alpha);
    s->lam_d[k] = l + dl * (real)0.25f;
    v3 c0 = mul(dir, w0 * dl), c1 = mul(dir, -w1 * dl);
    int32_t *A = s->jac + 4 * (size_t)a, *B = s->jac + 4 * (size_t)b;
    jac_add(&A[0], jac_fixed(c0.x)); jac_add(&A[1], jac_fixed(c0.y)); jac_add(&A[2], jac_fixed(c0.z)); A[3] += 1;
    jac_add(&B[0], jac_fixed(c1.x)); jac_add(&B[1], jac_fixed(c1.y)); jac_add(&B[2], jac_fixed(c1.z)); B[3] += 1;
}
/* distance_atomic.wgsl:162-182 (apply_atomic_accum): the mean correction times 1.25, sums cleared */
static void k_jacobi_apply(orc_sim *s) {
    for (uint32_t v = 0; v < s->nv; v++) {
        int32_t *A = s->jac + 4 * (size_t)v;
        uint32_t count = (uint32_t)A[3];
        if (count == 0) continue;
        real scale = (real)(1.0f / 1000000.0f) / (real)(float)count;
        s->p[v] = add(s->p[v], mk((real)(float)A[0] * scale * (real)1.25f, (real)(float)A[1] * scale * (real)1.25f,
                                  (real)(float)A[2] * scale * (real)1.25f));
        A[0] = A[1] = A[2] = A[3] = 0;
    }
}
/* sewing_shrink.wgsl:40-49 (once per substep) and the sewing lambdas' clear (dispatch.rs:117-119) */
static void k_sew_shrink(orc_sim *s, real dt_sub) {
    real dt = rmax(dt_sub, (real)1e-8f);
    for (uint32_t k = 0; k < s->m->ns; k++) {
        s->s_cur[k] = rmax(s->s_tgt[k], s->s_cur[k] - (real)s->sew_speed * dt);
        s->lam_s[k] = 0;
    }
}

/* sewing.wgsl:56-110: XPBD at the simulator's sewing compliance; a pair within
 * max(thickness sum, lock distance) with the lock on projects at zero compliance to the target
 * length and keeps that as its current rest length */
static void k_sewing_one(orc_sim *s, uint32_t k, real dt_sub) {
    const orc_mesh *m = s->m;
    uint32_t a = m->s_v0[k], b = m->s_v1[k];
    real w0 = s->w[a], w1 = s->w[b], wsum = w0 + w1;
    if (wsum <= (real)1e-7f) return;
    v3 delta = sub(s->p[a], s->p[b]);
    real dist = len3(delta);
    if (dist <= (real)1e-7f) return;
    real dt = rmax(dt_sub, (real)1e-8f);
    v3 dir = dvs(delta, dist);
    float compf = s->sew_stiffness >= 5000.0f ? 0.0f : 1.0f / fmaxf(s->sew_stiffness, 1.0f);
    real target = s->s_cur[k], comp = (real)compf;
    real thresh = rmax((real)(m->thick[a] + m->thick[b]), (real)s->sew_lock_dist);
    if (s->sew_lock && dist <= thresh) {
        comp = 0;
        target = s->s_tgt[k];
        s->s_cur[k] = s->s_tgt[k];
    }
    real alpha = comp / (dt * dt);
    real c = dist - target;
    real lam = s->lam_s[k];
    real dl = (-c - alpha * lam) / (wsum + alpha);
    s->lam_s[k] = lam + dl;
    s->p[a] = add(s->p[a], mul(dir, w0 * dl));
    s->p[b] = sub(s->p[b], mul(dir, w1 * dl));
}

static void sewing_pass(orc_sim *s, real dt_sub) {
    const orc_mesh *m = s->m;
    for (uint32_t c = 0; c < m->ns_col; c++)
        for (uint32_t k = m->s_col_off[c]; k < m->s_col_off[c] + m->s_col_cnt[c]; k++) k_sewing_one(s, k, dt_sub);
}

void orc_sim_set_sewing(orc_sim *s, float stiffness, int lock, float lock_distance, float speed) {
    s->sew_stiffness = fmaxf(stiffness, 1.0f);
    s->sew_lock = lock;
    s->sew_lock_dist = fmaxf(lock_distance, 0.0f);
    s->sew_speed = isfinite(speed) ? fmaxf(speed, 0.0f) : 1.0f;
}
void orc_sim_sewing_rest_lengths(orc_sim *s, double *out, const double *in) {
    for (uint32_t k = 0; k < s->m->ns; k++) {
        if (out) out[k] = (double)s->s_cur[k];
        if (in) s->s_cur[k] = (real)in[k];
    }
}

static void distance_pass(orc_sim *s, real dt_sub) {
    const orc_mesh *m = s->m;
    if (s->solver_mode == 1) {
        for (uint32_t k = 0; k < m->nd; k++) k_distance_jacobi_one(s, k, dt_sub);
        k_jacobi_apply(s);
        return;
    }
    for (uint32_t c = 0; c < m->nd_col; c++) {
        ORC_PARALLEL_FOR
        for (uint32_t k = m->d_col_off[c]; k < m->d_col_off[c] + m->d_col_cnt[c]; k++) k_distance_one(s, k, dt_sub);
    }
}

static int seg_tri(v3 po, v3 pn, v3 a, v3 b, v3 c, real *t_out, real *u_out, real *v_out);
/* collision.wgsl:364-374 */
static real seg_sphere_dist_sq(v3 a, v3 b, v3 c) {
    v3 ab = sub(b, a);
    real ab2 = dot(ab, ab);
    if (ab2 <= (real)1e-12f) { v3 d = sub(a, c); return dot(d, d); }
    real t = rclamp(dot(sub(c, a), ab) / ab2, 0, 1);
    v3 d = sub(add(a, mul(ab, t)), c);
    return dot(d, d);
}
/* collision.wgsl:273-319 closest point with face-interior flag */
static v3 closest_face(v3 p, v3 a, v3 b, v3 c, int *is_face) {
    *is_face = 0;
    v3 ab = sub(b, a), ac = sub(c, a), ap = sub(p, a);
    real d1 = dot(ab, ap), d2 = dot(ac, ap);
    if (d1 <= 0 && d2 <= 0) return a;
    v3 bp = sub(p, b);
    real d3 = dot(ab, bp), d4 = dot(ac, bp);
    if (d3 >= 0 && d4 <= d3) return b;
    real vc = d1 * d4 - d3 * d2;
    if (vc <= 0 && d1 >= 0 && d3 <= 0) return add(a, mul(ab, d1 / (d1 - d3)));
    v3 cp = sub(p, c);
    real d5 = dot(ab, cp), d6 = dot(ac, cp);
    if (d6 >= 0 && d5 <= d6) return c;
    real vb = d5 * d2 - d1 * d6;
    if (vb <= 0 && d2 >= 0 && d6 <= 0) return add(a, mul(ac, d2 / (d2 - d6)));
    real va = d3 * d6 - d5 * d4;
    if (va <= 0 && (d4 - d3) >= 0 && (d5 - d6) >= 0) return add(b, mul(sub(c, b), (d4 - d3) / ((d4 - d3) + (d5 - d6))));
    real denom = (real)1 / (va + vb + vc);
    *is_face = 1;
    return add(add(a, mul(ab, vb * denom)), mul(ac, vc * denom));
}

/* collision.wgsl:83-108 */
static void contact_response(v3 target, v3 n, real fr, real rest, real dt, v3 p_init, v3 *xn, v3 *p) {
    v3 total = sub(target, *xn);
    v3 tang = sub(total, mul(n, dot(total, n)));
    *p = sub(target, mul(tang, rclamp(fr, 0, 1)));
    real edt = rmax(dt, (real)1e-6f);
    real vin = dot(dvs(sub(p_init, *xn), edt), n);
    real vout = 0;
    if (vin < 0) vout = -rclamp(rest, 0, 1) * vin;
    real dxn = dot(sub(target, *xn), n) - vout * edt;
    *xn = add(*xn, mul(n, dxn));
}

int orc_sdf_contact(const uint32_t *tex, uint32_t w, uint32_t h, uint32_t d, const float *info, const float *xf,
                    uint32_t nb, uint32_t ntris_mesh, const real *p3, real vth, real *target3, real *normal3,
                    real *fr, real *rest);

/* collision.wgsl:377-652: primitives, mesh triangles, then SDF colliders (sdf_oracle.c) */
static void k_collision(orc_sim *s, real dt) {
    ORC_PARALLEL_FOR
    for (uint32_t i = 0; i < s->nv; i++) {
        if (s->w[i] <= 0) continue;
        v3 p = s->p[i], pinit = p, xn = s->x[i];
        real th = (real)s->m->thick[i];
        for (uint32_t c = 0; c < s->ncols; c++) {
            const collider *col = &s->cols[c];
            if (col->type == 0) {
                real tr = col->radius + th;
                v3 d = sub(p, col->a);
                real dist = len3(d);
                if (dist < tr) {
                    v3 n = mk(0, 0, 1);
                    if (dist > (real)1e-7f) n = dvs(d, dist);
                    contact_response(add(col->a, mul(n, tr)), n, col->friction, col->restitution, dt, pinit, &xn, &p);
                }
            } else if (col->type == 1) {
                real tr = col->radius + th;
                v3 ab = sub(col->b, col->a);
                real l2 = dot(ab, ab);
                real t = 0;
                if (l2 > (real)1e-7f) t = rclamp(dot(sub(p, col->a), ab) / l2, 0, 1);
                v3 cl = add(col->a, mul(ab, t));
                v3 d = sub(p, cl);
                real dist = len3(d);
                if (dist < tr) {
                    v3 n = mk(0, 0, 1);
                    if (dist > (real)1e-7f) n = dvs(d, dist);
                    contact_response(add(cl, mul(n, tr)), n, col->friction, col->restitution, dt, pinit, &xn, &p);
                }
            } else if (col->type == 2) {
                v3 n = dvs(col->b, len3(col->b));
                real dist = dot(sub(p, col->a), n);
                if (dist < th) {
                    contact_response(add(p, mul(n, th - dist)), n, col->friction, col->restitution, dt, pinit, &xn, &p);
                }
            }
        }
        /* triangle mesh, direct scan in Morton order (collision.wgsl:447-542) */
        int front = 0, has_rec = 0;
        real best_depth = (real)999999.0f;
        v3 best_pos = p, best_n = mk(0, 0, 0);
        real best_fr = 0, best_rest = 0;
        v3 x_old = s->x[i];
        for (uint32_t k = 0; k < s->ntris; k++) {
            const float *b = s->tri_bounds + 4 * k;
            real rr = (real)b[3] * (real)1.5f + th;
            if (seg_sphere_dist_sq(x_old, p, mk(b[0], b[1], b[2])) > rr * rr) continue;
            const mtri *T = &s->tris[k];
            v3 t0 = mk(T->p0[0], T->p0[1], T->p0[2]), t1 = mk(T->p1[0], T->p1[1], T->p1[2]), t2 = mk(T->p2[0], T->p2[1], T->p2[2]);
            real target = th + (real)T->thickness;
            v3 cr = cross(sub(t1, t0), sub(t2, t0));
            real cl = len3(cr);
            if (cl < (real)1e-7f) continue;
            v3 fn = dvs(cr, cl);
            int single = (T->flags & 1u) != 0u;
            real th_hit, hu, hv;
            if (seg_tri(x_old, p, t0, t1, t2, &th_hit, &hu, &hv)) {
                v3 hit = add(x_old, mul(sub(p, x_old), th_hit));
                v3 nrm = fn;
                if (dot(sub(p, x_old), fn) > 0 && !single) nrm = neg(fn);
                contact_response(add(hit, mul(nrm, target)), nrm, T->friction, T->restitution, dt, pinit, &xn, &p);
                front = 1;
            }
            int is_face;
            v3 q = closest_face(p, t0, t1, t2, &is_face);
            v3 d = sub(p, q);
            real dist = len3(d);
            real sd = dot(d, fn);
            if (single) {
                if (sd >= 0) {
                    if (dist < target) {
                        v3 nrm = fn;
                        if (dist > (real)1e-7f) nrm = dvs(d, dist);
                        contact_response(add(p, mul(nrm, target - dist)), nrm, T->friction, T->restitution, dt, pinit, &xn, &p);
                        front = 1;
                    }
                } else if (s->single_sided_recovery && (T->flags & 2u) == 0u) {
                    if (is_face && dist < target * (real)3.0f && dist <= (real)fabs((double)sd) * (real)1.1f + (real)1e-7f &&
                        dist < best_depth) {
                        best_depth = dist; best_n = fn; best_pos = add(p, mul(fn, target - sd));
                        best_fr = T->friction; best_rest = T->restitution; has_rec = 1;
                    }
                }
            } else if (dist < target && sd >= -target) {
                v3 nrm = fn;
                if (dist > (real)1e-7f) nrm = dvs(d, dist);
                contact_response(add(p, mul(nrm, target - dist)), nrm, T->friction, T->restitution, dt, pinit, &xn, &p);
                front = 1;
            }
        }
        if (!front && has_rec) contact_response(best_pos, best_n, best_fr, best_rest, dt, pinit, &xn, &p);
        /* SDF colliders last, velocity reference = the position at this point (collision.wgsl:645-647) */
        if (s->sdf_nb) {
            real pa[3] = {p.x, p.y, p.z}, tg[3], nn[3], fr, rest;
            if (orc_sdf_contact(s->sdf_tex, s->sdf_w, s->sdf_h, s->sdf_d, s->sdf_info, s->sdf_xf, s->sdf_nb, s->ntris,
                                pa, th, tg, nn, &fr, &rest))
                contact_response(mk(tg[0], tg[1], tg[2]), mk(nn[0], nn[1], nn[2]), fr, rest, dt, p, &xn, &p);
        }
        s->p[i] = p;
        s->x[i] = xn;
    }
}

/* update_vel.wgsl:33-53 */
static void k_update_vel(orc_sim *s, real dt_sub) {
    real dt = rmax(dt_sub, (real)1e-8f);
    ORC_PARALLEL_FOR
    for (uint32_t i = 0; i < s->nv; i++) {
        if (s->w[i] > 0) s->v[i] = dvs(sub(s->p[i], s->x[i]), dt);
        else s->v[i] = mk(0, 0, 0);
        s->x[i] = s->p[i];
    }
}

/* ---------------- self collision (self_collision.wgsl) ---------------- */
#define FIXED_SCALE ((real)1000000.0f)
#define MAX_FIXED ((real)50000000.0f)
static inline int32_t to_fixed(real x) { return f2i_sat(rclamp(x * FIXED_SCALE, -MAX_FIXED, MAX_FIXED)); }
static inline void acc_add(int32_t *dst, int32_t v) { *dst = (int32_t)((uint32_t)*dst + (uint32_t)v); }
static void add_disp(accum_t *a, v3 d) {
    acc_add(&a->dx, to_fixed(d.x)); acc_add(&a->dy, to_fixed(d.y)); acc_add(&a->dz, to_fixed(d.z)); a->count += 1;
}
static void add_pin_disp(accum_t *a, v3 d) {
    acc_add(&a->pdx, to_fixed(d.x)); acc_add(&a->pdy, to_fixed(d.y)); acc_add(&a->pdz, to_fixed(d.z)); a->pcount += 1;
}
static void ccd_axis(int32_t ix, int32_t *pos, int32_t *negv) {
    if (ix > 0) { if (ix > *pos) *pos = ix; }
    else if (ix < 0) { if (ix < *negv) *negv = ix; }
}
static void add_ccd(accum_t *a, v3 d) {
    ccd_axis(to_fixed(d.x), &a->cpx, &a->cnx);
    ccd_axis(to_fixed(d.y), &a->cpy, &a->cny);
    ccd_axis(to_fixed(d.z), &a->cpz, &a->cnz);
}

typedef struct { v3 pt, bary; } cbary;
static cbary closest_bary(v3 p, v3 a, v3 b, v3 c) { /* self_collision.wgsl:134-179 */
    cbary r;
    v3 ab = sub(b, a), ac = sub(c, a), ap = sub(p, a);
    real d1 = dot(ab, ap), d2 = dot(ac, ap);
    if (d1 <= 0 && d2 <= 0) { r.pt = a; r.bary = mk(1, 0, 0); return r; }
    v3 bp = sub(p, b);
    real d3 = dot(ab, bp), d4 = dot(ac, bp);
    if (d3 >= 0 && d4 <= d3) { r.pt = b; r.bary = mk(0, 1, 0); return r; }
    real vc = d1 * d4 - d3 * d2;
    if (vc <= 0 && d1 >= 0 && d3 <= 0) { real v = d1 / (d1 - d3); r.pt = add(a, mul(ab, v)); r.bary = mk(1 - v, v, 0); return r; }
    v3 cp = sub(p, c);
    real d5 = dot(ab, cp), d6 = dot(ac, cp);
    if (d6 >= 0 && d5 <= d6) { r.pt = c; r.bary = mk(0, 0, 1); return r; }
    real vb = d5 * d2 - d1 * d6;
    if (vb <= 0 && d2 >= 0 && d6 <= 0) { real w = d2 / (d2 - d6); r.pt = add(a, mul(ac, w)); r.bary = mk(1 - w, 0, w); return r; }
    real va = d3 * d6 - d5 * d4;
    if (va <= 0 && (d4 - d3) >= 0 && (d5 - d6) >= 0) {
        real w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        r.pt = add(b, mul(sub(c, b), w)); r.bary = mk(0, 1 - w, w); return r;
    }
    real denom = (real)1 / (va + vb + vc);
    real v = vb * denom, w = vc * denom;
    r.pt = add(add(a, mul(ab, v)), mul(ac, w)); r.bary = mk(1 - v - w, v, w);
    return r;
}

/* Möller–Trumbore segment/triangle, self_collision.wgsl:182-218 (identical in collision.wgsl:322-361).
 * returns hit and fills t,u,v */
static int seg_tri(v3 po, v3 pn, v3 a, v3 b, v3 c, real *t_out, real *u_out, real *v_out) {
    v3 d = sub(pn, po);
    real ld = dot(d, d);
    if (ld < (real)1e-12f) return 0;
    v3 e1 = sub(b, a), e2 = sub(c, a);
    v3 pv = cross(d, e2);
    real det = dot(e1, pv);
    real tn = dot(e1, e1) * dot(e2, e2);
    if (det * det <= tn * ld * (real)1e-10f || tn <= (real)1e-20f) return 0;
    real inv = (real)1 / det;
    v3 tv = sub(po, a);
    real u = dot(tv, pv) * inv;
    if (u < (real)-0.05f || u > (real)1.05f) return 0;
    v3 qv = cross(tv, e1);
    real v = dot(d, qv) * inv;
    if (v < (real)-0.05f || (u + v) > (real)1.05f) return 0;
    real t = dot(e2, qv) * inv;
    if (t >= 0 && t <= 1) { *t_out = t; *u_out = u; *v_out = v; return 1; }
    return 0;
}

static int within_two_hops(const orc_mesh *m, uint32_t a, uint32_t b) { /* self_collision.wgsl:262-283 */
    if (a == b) return 1;
    uint32_t lo = m->th_off[a], hi = m->th_off[a + 1];
    while (lo < hi) {
        uint32_t mid = lo + (hi - lo) / 2, val = m->th_idx[mid];
        if (val == b) return 1;
        if (val < b) lo = mid + 1; else hi = mid;
    }
    return 0;
}
static int near_folded(const orc_mesh *m, uint32_t a, uint32_t b, real cur) { /* self_collision.wgsl:287-314 */
    if (a == b) return 1;
    uint32_t lo = m->th_off[a], hi = m->th_off[a + 1];
    while (lo < hi) {
        uint32_t mid = lo + (hi - lo) / 2, val = m->th_idx[mid];
        if (val == b) { real l0 = m->th_rest[mid]; return cur < l0 * (real)0.5f ? 0 : 1; }
        if (val < b) lo = mid + 1; else hi = mid;
    }
    return 0;
}
static int near_folded_sq(const orc_mesh *m, uint32_t a, uint32_t b, real cur_sq) { /* ee_direct.wgsl:259-283 */
    if (a == b) return 1;
    uint32_t lo = m->th_off[a], hi = m->th_off[a + 1];
    while (lo < hi) {
        uint32_t mid = lo + (hi - lo) / 2, val = m->th_idx[mid];
        if (val == b) { real l0 = m->th_rest[mid]; return cur_sq < (l0 * l0) * (real)0.25f ? 0 : 1; }
        if (val < b) lo = mid + 1; else hi = mid;
    }
    return 0;
}

/* sc_solve: self_collision.wgsl:317-681.  On a multi-layer cloth with untangling on (the branches
 * guarded by enable_normal_untangling): each free outer vertex records its nearest inner vertex
 * outside its two-hop ring (lines 340-386, 678-681); V-V and V-T responses are skipped where the outer
 * vertex sits behind the inner one (408-427, 546-556, 613-627). */
static void k_sc_solve(orc_sim *s) {
    const orc_mesh *m = s->m;
    const real EPS = (real)1e-7f;
    real cell = s->cell_size;
    uint32_t T = s->table_size;
    for (uint32_t index = 0; index < s->nv; index++) {
        int pinned_i = s->w[index] <= 0;
        v3 xo_i = s->x[index], p_i = s->p[index];
        real move_i0 = len3(sub(p_i, xo_i));
        if (pinned_i && move_i0 < EPS) continue;
        real thick_i = (real)m->thick[index];
        real llen_i = (real)m->local_len[index];
        int32_t cc[3]; cell_of(p_i, cell, cc);
        const int ut = untangle_on(s);
        const int track = ut && !pinned_i && m->layer[index] > 0;
        real best_inner = (real)1e20f;
        uint32_t best_inner_j = 0xFFFFFFFFu;
        for (int32_t cx = cc[0] - 1; cx <= cc[0] + 1; cx++)
        for (int32_t cy = cc[1] - 1; cy <= cc[1] + 1; cy++)
        for (int32_t cz = cc[2] - 1; cz <= cc[2] + 1; cz++) {
            int32_t nc[3] = {cx, cy, cz};
            uint32_t h = hash_coords(nc, T);
            for (uint32_t k = s->hstarts[h]; k < s->hstarts[h + 1]; k++) {
                uint32_t j = s->hsorted[k];
                if (j >= s->nv || j == index) continue;
                v3 p_j = s->p[j];
                int32_t jc[3]; cell_of(p_j, cell, jc);
                if (!(jc[0] == cx && jc[1] == cy && jc[2] == cz)) continue;
                v3 dvv = sub(p_i, p_j);
                real dsq = dot(dvv, dvv);
                real thick_j = (real)m->thick[j], llen_j = (real)m->local_len[j];
                real min_dist = thick_i + thick_j;
                real edge_margin = rmin(llen_i, llen_j) * (real)0.06f;
                real eff = rmax(min_dist, edge_margin);
                real max_geom = eff + (llen_i + llen_j) * (real)1.3f;
                if (track && m->layer[j] < m->layer[index] && dsq < best_inner) {
                    real r = eff * (real)2.5f + llen_j * (real)1.3f;
                    if (dsq <= r * r && !within_two_hops(m, index, j)) { best_inner = dsq; best_inner_j = j; }
                }
                if (!(dsq <= max_geom * max_geom)) continue;
                real dvv_len = RSQRT(rmax(dsq, 0));
                if (near_folded(m, index, j, dvv_len)) continue;
                v3 pj_old = s->x[j];
                int both_dynamic = s->w[j] > 0;
                real eff_sq = eff * eff;
                /* 1. V-V */
                if (!pinned_i && (!both_dynamic || index < j)) {
                    int untangling = 0;
                    if (ut && dsq < eff_sq && dsq > EPS * EPS && m->layer[index] != m->layer[j]) {
                        v3 ni = s->normals[index], nj = s->normals[j];
                        real nil = len3(ni), njl = len3(nj);
                        if (m->layer[index] > m->layer[j]) {
                            if (njl > (real)0.5f) {
                                v3 nju = dvs(nj, njl);
                                int allow = s->w[j] <= 0 || nil <= (real)0.5f || dot(dvs(ni, rmax(nil, EPS)), nju) > 0;
                                untangling = allow && dot(dvv, nju) < 0;
                            }
                        } else if (both_dynamic && nil > (real)0.5f) {
                            v3 niu = dvs(ni, nil);
                            int allow = njl <= (real)0.5f || dot(dvs(nj, rmax(njl, EPS)), niu) > 0;
                            untangling = allow && dot(neg(dvv), niu) < 0;
                        }
                    }
                    if (!untangling && dsq < eff_sq && dsq > EPS * EPS) {
                        s->stats[0]++;
                        v3 n = dvs(dvv, dvv_len);
                        real pen = eff - dvv_len;
                        real wi = s->w[index], wj = s->w[j], ws = wi + wj;
                        if (ws > EPS) {
                            if (wj <= 0) {
                                v3 jm = sub(p_j, pj_old);
                                add_pin_disp(&s->acc[index], add(mul(n, pen), mul(jm, (real)0.5f)));
                            } else {
                                v3 disp = mul(n, pen * (real)0.5f);
                                add_disp(&s->acc[index], mul(disp, wi / ws));
                                add_disp(&s->acc[j], mul(neg(disp), wj / ws));
                            }
                        }
                    }
                }
                /* 2. V-T */
                real mv_i = len3(sub(p_i, xo_i));
                real mv_j = len3(sub(p_j, pj_old));
                real vt_bound = eff + llen_j * (real)1.3f + mv_i + mv_j;
                if (!(dsq <= vt_bound * vt_bound) || within_two_hops(m, index, j)) continue;
                for (uint32_t st = m->star_off[j]; st < m->star_off[j + 1]; st++) {
                    uint32_t v0 = m->star_v0[st], v1 = m->star_v1[st];
                    if (v0 == index || v1 == index) continue;
                    v3 p_v0 = s->p[v0], p_v1 = s->p[v1], p_v0o = s->x[v0], p_v1o = s->x[v1];
                    v3 d0 = sub(p_i, p_v0);
                    real ll0 = (real)m->local_len[v0];
                    real et0 = rmax(thick_i + (real)m->thick[v0], rmin(llen_i, ll0) * (real)0.06f);
                    real b0 = et0 + ll0 * (real)1.3f + mv_i + len3(sub(p_v0, p_v0o));
                    int32_t c0[3]; cell_of(p_v0, cell, c0);
                    int in0 = abs(c0[0] - cc[0]) <= 1 && abs(c0[1] - cc[1]) <= 1 && abs(c0[2] - cc[2]) <= 1 &&
                              dot(d0, d0) <= b0 * b0;
                    v3 d1 = sub(p_i, p_v1);
                    real ll1 = (real)m->local_len[v1];
                    real et1 = rmax(thick_i + (real)m->thick[v1], rmin(llen_i, ll1) * (real)0.06f);
                    real b1 = et1 + ll1 * (real)1.3f + mv_i + len3(sub(p_v1, p_v1o));
                    int32_t c1[3]; cell_of(p_v1, cell, c1);
                    int in1 = abs(c1[0] - cc[0]) <= 1 && abs(c1[1] - cc[1]) <= 1 && abs(c1[2] - cc[2]) <= 1 &&
                              dot(d1, d1) <= b1 * b1;
                    if ((in0 && j > v0) || (in1 && j > v1)) continue;
                    v3 tri_move = dvs(add(add(sub(p_j, pj_old), sub(p_v0, p_v0o)), sub(p_v1, p_v1o)), (real)3.0f);
                    int tri_pin = s->w[j] <= 0 || s->w[v0] <= 0 || s->w[v1] <= 0;
                    v3 rel_old = add(xo_i, tri_move);
                    v3 effv = mk(eff, eff, eff);
                    v3 smin = sub(vmin(rel_old, p_i), effv), smax = add(vmax(rel_old, p_i), effv);
                    v3 tmin = vmin(vmin(p_j, p_v0), p_v1), tmax = vmax(vmax(p_j, p_v0), p_v1);
                    if (smin.x > tmax.x || smin.y > tmax.y || smin.z > tmax.z ||
                        smax.x < tmin.x || smax.y < tmin.y || smax.z < tmin.z) continue;
                    if (within_two_hops(m, index, v0) || within_two_hops(m, index, v1)) continue;
                    v3 move = sub(p_i, rel_old);
                    real move_sq = dot(move, move);
                    int hit = 0; real ht = 1, hu = 0, hv = 0;
                    if (move_sq > (real)1e-12f) hit = seg_tri(rel_old, p_i, p_j, p_v0, p_v1, &ht, &hu, &hv);
                    if (hit) {
                        s->stats[1]++;
                        v3 hit_pt = add(rel_old, mul(move, ht));
                        v3 tcr = cross(sub(p_v0, p_j), sub(p_v1, p_j));
                        real tl = len3(tcr);
                        v3 tn = tl > (real)1e-12f ? dvs(tcr, tl) : mk(0, 1, 0);
                        real h_old_raw = dot(sub(rel_old, p_j), tn);
                        if (ut && !pinned_i && m->layer[index] > m->layer[j] && tl > (real)1e-12f) {
                            v3 ni = s->normals[index];
                            real nil = len3(ni);
                            int allow = tri_pin || nil <= (real)0.5f || dot(dvs(ni, rmax(nil, EPS)), tn) > 0;
                            if (allow && h_old_raw < 0 && dot(move, tn) > 0) continue;
                        }
                        real side = fabs((double)h_old_raw) > 1e-6 ? h_old_raw : dot(neg(move), tn);
                        v3 n_side = side >= 0 ? tn : neg(tn);
                        real h_old = rmax(dot(sub(rel_old, hit_pt), n_side), 0);
                        real h_safe = rmax(h_old, eff * (real)0.25f);
                        v3 p_safe = add(hit_pt, mul(n_side, h_safe));
                        v3 total = sub(p_safe, p_i);
                        real pen_soft = rmax(0, eff - h_safe);
                        if (tri_pin) {
                            if (!pinned_i) add_ccd(&s->acc[index], add(total, mul(n_side, pen_soft)));
                        } else if (pinned_i) {
                            v3 bc = mk(rclamp(1 - hu - hv, 0, 1), rclamp(hu, 0, 1), rclamp(hv, 0, 1));
                            real wj = s->w[j], w0 = s->w[v0], w1 = s->w[v1];
                            real wt = bc.x * bc.x * wj + bc.y * bc.y * w0 + bc.z * bc.z * w1;
                            if (wt > EPS) {
                                v3 full = add(total, mul(n_side, pen_soft));
                                if (wj > 0) add_ccd(&s->acc[j], mul(neg(full), bc.x * wj / wt));
                                if (w0 > 0) add_ccd(&s->acc[v0], mul(neg(full), bc.y * w0 / wt));
                                if (w1 > 0) add_ccd(&s->acc[v1], mul(neg(full), bc.z * w1 / wt));
                            }
                        } else {
                            v3 bc = mk(rclamp(1 - hu - hv, 0, 1), rclamp(hu, 0, 1), rclamp(hv, 0, 1));
                            real wi = s->w[index], wj = s->w[j], w0 = s->w[v0], w1 = s->w[v1];
                            real wt = bc.x * bc.x * wj + bc.y * bc.y * w0 + bc.z * bc.z * w1;
                            real wtot = wi + wt;
                            if (wtot > EPS) {
                                add_ccd(&s->acc[index], mul(total, wi / wtot));
                                if (wj > 0) add_ccd(&s->acc[j], mul(neg(total), bc.x * wj / wtot));
                                if (w0 > 0) add_ccd(&s->acc[v0], mul(neg(total), bc.y * w0 / wtot));
                                if (w1 > 0) add_ccd(&s->acc[v1], mul(neg(total), bc.z * w1 / wtot));
                                if (pen_soft > EPS) {
                                    v3 tp = mul(n_side, pen_soft * (real)0.6f);
                                    add_disp(&s->acc[index], mul(tp, wi / wtot));
                                    if (wj > 0) add_disp(&s->acc[j], mul(neg(tp), bc.x * wj / wtot));
                                    if (w0 > 0) add_disp(&s->acc[v0], mul(neg(tp), bc.y * w0 / wtot));
                                    if (w1 > 0) add_disp(&s->acc[v1], mul(neg(tp), bc.z * w1 / wtot));
                                }
                            }
                        }
                    } else {
                        cbary r = closest_bary(p_i, p_j, p_v0, p_v1);
                        v3 dvt = sub(p_i, r.pt);
                        real dvsq = dot(dvt, dvt);
                        if (!pinned_i && dvsq < eff_sq && dvsq > EPS * EPS) {
                            real dvt_len = RSQRT(dvsq);
                            if (ut && m->layer[index] > m->layer[j]) {
                                v3 tcr = cross(sub(p_v0, p_j), sub(p_v1, p_j));
                                real tl = len3(tcr);
                                if (tl > EPS) {
                                    v3 tn = dvs(tcr, tl);
                                    if (dot(dvt, tn) < 0) {
                                        v3 ni = s->normals[index];
                                        real nil = len3(ni);
                                        if (tri_pin || nil <= (real)0.5f || dot(dvs(ni, rmax(nil, EPS)), tn) > 0) continue;
                                    }
                                }
                            }
                            s->stats[2]++;
                            v3 nvt = dvs(dvt, dvt_len);
                            real pen = eff - dvt_len;
                            if (tri_pin) {
                                add_pin_disp(&s->acc[index], add(mul(nvt, pen), mul(tri_move, (real)0.5f)));
                            } else {
                                real wi = s->w[index], wj = s->w[j], w0 = s->w[v0], w1 = s->w[v1];
                                v3 bc = r.bary;
                                real wt = bc.x * bc.x * wj + bc.y * bc.y * w0 + bc.z * bc.z * w1;
                                real wtot = wi + wt;
                                if (wtot > EPS) {
                                    v3 tp = mul(nvt, pen * (real)0.6f);
                                    add_disp(&s->acc[index], mul(tp, wi / wtot));
                                    if (wj > 0) add_disp(&s->acc[j], mul(neg(tp), bc.x * wj / wtot));
                                    if (w0 > 0) add_disp(&s->acc[v0], mul(neg(tp), bc.y * w0 / wtot));
                                    if (w1 > 0) add_disp(&s->acc[v1], mul(neg(tp), bc.z * w1 / wtot));
                                }
                            }
                        }
                    }
                }
            }
        }
        if (track && best_inner_j < 0x3FFFFEu) s->acc[index].untangle = best_inner_j + 1u;
    }
}

/* ---------------- E-E (self_collision_ee_direct.wgsl) ---------------- */
static void add_disp_ee(accum_t *a, v3 d) { /* ee_direct.wgsl:70-79: the count is incremented twice */
    acc_add(&a->dx, to_fixed(d.x)); acc_add(&a->dy, to_fixed(d.y)); acc_add(&a->dz, to_fixed(d.z)); a->count += 2;
}
static void add_ccd_ee(accum_t *a, v3 d, int count_ee) { /* ee_direct.wgsl:92-116 */
    add_ccd(a, d);
    if (count_ee) {
        real l = len3(d) * FIXED_SCALE;
        real lc = rclamp(l, 0, (real)2000000000.0f);
        uint32_t lf = lc >= (real)4294967296.0 ? 0xFFFFFFFFu : (uint32_t)lc;
        if (lf > 0) a->ee_count += 1;
    }
}
typedef struct { real s, t; } st2;
static st2 closest_segs(v3 p1, v3 q1, v3 p2, v3 q2) { /* ee_direct.wgsl:155-193 */
    st2 r = {0, 0};
    v3 d1 = sub(q1, p1), d2 = sub(q2, p2), rr = sub(p1, p2);
    real a = dot(d1, d1), e = dot(d2, d2), f = dot(d2, rr);
    if (a <= (real)1e-12f && e <= (real)1e-12f) return r;
    if (a <= (real)1e-12f) { r.t = rclamp(f / e, 0, 1); return r; }
    real c = dot(d1, rr);
    if (e <= (real)1e-12f) { r.s = rclamp(-c / a, 0, 1); return r; }
    real b = dot(d1, d2);
    real denom = a * e - b * b;
    real s = 0;
    if (denom > a * e * (real)1e-7f) s = rclamp((b * f - c * e) / denom, 0, 1);
    real t = (b * s + f) / e;
    if (t < 0) { t = 0; s = rclamp(-c / a, 0, 1); }
    else if (t > 1) { t = 1; s = rclamp((b - c) / a, 0, 1); }
    r.s = s; r.t = t;
    return r;
}
static real seg_dist_sq(v3 a0o, v3 a0, v3 a1o, v3 a1, v3 b0o, v3 b0, v3 b1o, v3 b1, real tau, st2 *st) {
    v3 A0 = vmix(a0o, a0, tau), A1 = vmix(a1o, a1, tau), B0 = vmix(b0o, b0, tau), B1 = vmix(b1o, b1, tau);
    *st = closest_segs(A0, A1, B0, B1);
    v3 pa = vmix(A0, A1, st->s), pb = vmix(B0, B1, st->t);
    v3 d = sub(pa, pb);
    return dot(d, d);
}
static real seg_volume(v3 a0o, v3 a0, v3 a1o, v3 a1, v3 b0o, v3 b0, v3 b1o, v3 b1, real tau) {
    v3 A0 = vmix(a0o, a0, tau), A1 = vmix(a1o, a1, tau), B0 = vmix(b0o, b0, tau), B1 = vmix(b1o, b1, tau);
    return dot(cross(sub(A1, A0), sub(B1, B0)), sub(B0, A0));
}
static int bbit(const orc_mesh *m, uint32_t e) { return (m->e_bbits[e >> 5] >> (e & 31)) & 1; }

/* is_outer_edge_behind_inner (self_collision_ee_direct.wgsl:123-148); delta_ab from edge b to edge a */
static int outer_edge_behind_inner(const orc_sim *s, uint32_t la, uint32_t lb, uint32_t a0, uint32_t a1, uint32_t b0,
                                   uint32_t b1, int a_fixed, int b_fixed, v3 delta_ab) {
    const real EPS = (real)1e-7f;
    if (la == lb) return 0;
    v3 na = add(s->normals[a0], s->normals[a1]), nb = add(s->normals[b0], s->normals[b1]);
    real nal = len3(na), nbl = len3(nb);
    if (la > lb) {
        if (nbl <= (real)0.5f) return 0;
        v3 ni = dvs(nb, nbl);
        int allow = b_fixed || nal <= (real)0.5f || dot(dvs(na, rmax(nal, EPS)), ni) > 0;
        return allow && dot(delta_ab, ni) < 0;
    }
    if (nal <= (real)0.5f) return 0;
    v3 ni = dvs(na, nal);
    int allow = a_fixed || nbl <= (real)0.5f || dot(dvs(nb, rmax(nbl, EPS)), ni) > 0;
    return allow && dot(neg(delta_ab), ni) < 0;
}
/* the E-E CCD's untangling exemption (ee_direct.wgsl:506-524, solve_ee.wgsl:386-408) */
static int ee_ccd_untangling(const orc_sim *s, uint32_t la, uint32_t lb, uint32_t u0, uint32_t j, int a_fixed,
                             int b_fixed, v3 a0o, v3 a1o, v3 b0o, v3 b1o, v3 a0, v3 a1, v3 b0, v3 b1) {
    if (la == lb) return 0;
    v3 d0 = mul(sub(add(a0o, a1o), add(b0o, b1o)), (real)0.5f);
    v3 d1 = mul(sub(add(a0, a1), add(b0, b1)), (real)0.5f);
    v3 mv = sub(d1, d0);
    if (la > lb && !a_fixed) {
        v3 ni = s->normals[j], no = s->normals[u0];
        int allow = b_fixed || dot(no, no) <= (real)0.25f;
        return allow && dot(ni, ni) > (real)0.25f && dot(d0, ni) < 0 && dot(mv, ni) > 0;
    }
    if (lb > la && !b_fixed) {
        v3 ni = s->normals[u0], no = s->normals[j];
        int allow = a_fixed || dot(no, no) <= (real)0.25f;
        return allow && dot(ni, ni) > (real)0.25f && dot(neg(d0), ni) < 0 && dot(neg(mv), ni) > 0;
    }
    return 0;
}

static void k_sc_solve_ee(orc_sim *s) {
    const orc_mesh *m = s->m;
    const real EPS = (real)1e-7f;
    real cell = s->cell_size;
    uint32_t T = s->table_size;
    for (uint32_t ea = 0; ea < m->ne; ea++) {
        uint32_t u0 = m->e_v0[ea], u1 = m->e_v1[ea];
        int pin_u0 = s->w[u0] <= 0, pin_u1 = s->w[u1] <= 0;
        int a_has_pin = pin_u0 || pin_u1;
        v3 p_u0 = s->p[u0], p_u1 = s->p[u1], p_u0o = s->x[u0], p_u1o = s->x[u1];
        real th_u0 = (real)m->thick[u0], th_u1 = (real)m->thick[u1];
        real max_th_a = rmax(th_u0, th_u1);
        int bnd_a = bbit(m, ea);
        v3 mta = mk(max_th_a * 2, max_th_a * 2, max_th_a * 2);
        v3 amin = sub(vmin(vmin(p_u0, p_u1), vmin(p_u0o, p_u1o)), mta);
        v3 amax = add(vmax(vmax(p_u0, p_u1), vmax(p_u0o, p_u1o)), mta);
        v3 pmid = mul(add(p_u0, p_u1), (real)0.5f);
        int32_t bc[3]; cell_of(pmid, cell, bc);
        uint32_t search_count = 0;
        int done = 0;
        for (int dz = -1; dz <= 1 && !done; dz++)
        for (int dy = -1; dy <= 1 && !done; dy++)
        for (int dx = -1; dx <= 1 && !done; dx++) {
            int32_t nc[3] = {bc[0] + dx, bc[1] + dy, bc[2] + dz};
            uint32_t h = hash_coords(nc, T);
            uint32_t start = s->estarts[h], end = s->estarts[h + 1];
            if (start >= end) continue;
            for (uint32_t q = start; q < end && !done; q++) {
                uint32_t eb = s->esorted[q];
                if (ea >= eb) continue;
                uint32_t j = m->e_v0[eb], vj = m->e_v1[eb];
                if (j == u0 || j == u1 || vj == u0 || vj == u1) continue;
                v3 p_j = s->p[j], p_vj = s->p[vj];
                v3 bmid = mul(add(p_j, p_vj), (real)0.5f);
                int32_t bcell[3]; cell_of(bmid, cell, bcell);
                if (!(bcell[0] == nc[0] && bcell[1] == nc[1] && bcell[2] == nc[2])) continue;
                v3 p_jo = s->x[j], p_vjo = s->x[vj];
                v3 bmin_ = vmin(vmin(p_j, p_vj), vmin(p_jo, p_vjo)), bmax_ = vmax(vmax(p_j, p_vj), vmax(p_jo, p_vjo));
                if (amin.x > bmax_.x || amin.y > bmax_.y || amin.z > bmax_.z ||
                    amax.x < bmin_.x || amax.y < bmin_.y || amax.z < bmin_.z) continue;
                search_count++;
                if (search_count > s->max_search_iterations) { done = 1; break; }
                int bnd_b = bbit(m, eb);
                int bnd_pair = bnd_a || bnd_b;
                int run_ccd = bnd_pair;
                int has_cached = 0;
                st2 st1c = {0, 0};
                v3 pt1ac = mk(0, 0, 0), pt1bc = mk(0, 0, 0), d1c = mk(0, 0, 0);
                real dsqc = 0, effc = 0;
                real th_j = (real)m->thick[j], th_vj = (real)m->thick[vj];
                if (!run_ccd) {
                    st1c = closest_segs(p_u0, p_u1, p_j, p_vj);
                    pt1ac = vmix(p_u0, p_u1, st1c.s);
                    pt1bc = vmix(p_j, p_vj, st1c.t);
                    d1c = sub(pt1ac, pt1bc);
                    dsqc = dot(d1c, d1c);
                    effc = rmix(th_u0, th_u1, st1c.s) + rmix(th_j, th_vj, st1c.t);
                    real max_eff = max_th_a + rmax(th_j, th_vj);
                    v3 du0 = sub(p_u0, p_u0o), du1 = sub(p_u1, p_u1o), dj = sub(p_j, p_jo), dvj = sub(p_vj, p_vjo);
                    v3 va = mul(add(du0, du1), (real)0.5f), vb = mul(add(dj, dvj), (real)0.5f);
                    real rel = len3(sub(va, vb)) + (real)0.5f * (len3(sub(du1, du0)) + len3(sub(dvj, dj)));
                    real reach = max_eff + rel;
                    if (dsqc > reach * reach) continue;
                    has_cached = 1;
                    v3 d0g = sub(vmix(p_u0o, p_u1o, st1c.s), vmix(p_jo, p_vjo, st1c.t));
                    int g1 = rel > effc * (real)0.5f;
                    int g2 = dot(d0g, d1c) <= 0;
                    run_ccd = g1 || g2;
                }
                if (run_ccd) {
                    real min_sq = (real)1e20f, best_tau = 0, half_span = (real)0.125f;
                    st2 min_st = {0, 0};
                    real prev_vol = seg_volume(p_u0o, p_u0, p_u1o, p_u1, p_jo, p_j, p_vjo, p_vj, 0);
                    real zc_tau = -1;
                    for (uint32_t k = 0; k <= 8; k++) {
                        real tk = (real)k * (real)0.125f;
                        st2 stk;
                        real dk = seg_dist_sq(p_u0o, p_u0, p_u1o, p_u1, p_jo, p_j, p_vjo, p_vj, tk, &stk);
                        if (dk < min_sq) { min_sq = dk; best_tau = tk; min_st = stk; }
                        if (k > 0) {
                            real vol = seg_volume(p_u0o, p_u0, p_u1o, p_u1, p_jo, p_j, p_vjo, p_vj, tk);
                            if (prev_vol * vol <= 0 && zc_tau < 0) zc_tau = tk - (real)0.0625f;
                            prev_vol = vol;
                        }
                    }
                    if (zc_tau >= 0) {
                        st2 stz;
                        real dz_ = seg_dist_sq(p_u0o, p_u0, p_u1o, p_u1, p_jo, p_j, p_vjo, p_vj, zc_tau, &stz);
                        if (dz_ < min_sq) { min_sq = dz_; best_tau = zc_tau; half_span = (real)0.0625f; min_st = stz; }
                    }
                    real tlo = rmax(0, best_tau - half_span), thi = rmin(1, best_tau + half_span);
                    real tau_star = best_tau, dsq_star = min_sq;
                    st2 st_star = min_st;
                    const real third = (real)(1.0f / 3.0f);
                    for (int it = 0; it < 6; it++) {
                        real m1 = tlo + (thi - tlo) * third;
                        real m2 = thi - (thi - tlo) * third;
                        st2 s1, s2;
                        real e1 = seg_dist_sq(p_u0o, p_u0, p_u1o, p_u1, p_jo, p_j, p_vjo, p_vj, m1, &s1);
                        real e2 = seg_dist_sq(p_u0o, p_u0, p_u1o, p_u1, p_jo, p_j, p_vjo, p_vj, m2, &s2);
                        if (e1 < dsq_star) { dsq_star = e1; tau_star = m1; st_star = s1; }
                        if (e2 < dsq_star) { dsq_star = e2; tau_star = m2; st_star = s2; }
                        if (e1 < e2) thi = m2; else tlo = m1;
                    }
                    real ss = st_star.s, ts = st_star.t;
                    real eff_star = rmix(th_u0, th_u1, ss) + rmix(th_j, th_vj, ts);
                    if (dsq_star >= eff_star * eff_star) continue;
                    if (s->exclude_neighbors) {
                        v3 a0s = vmix(p_u0o, p_u0, tau_star), a1s = vmix(p_u1o, p_u1, tau_star);
                        v3 b0s = vmix(p_jo, p_j, tau_star), b1s = vmix(p_vjo, p_vj, tau_star);
                        v3 t0 = sub(a0s, b0s), t1 = sub(a0s, b1s), t2 = sub(a1s, b0s), t3 = sub(a1s, b1s);
                        if (near_folded_sq(m, u0, j, dot(t0, t0)) || near_folded_sq(m, u0, vj, dot(t1, t1)) ||
                            near_folded_sq(m, u1, j, dot(t2, t2)) || near_folded_sq(m, u1, vj, dot(t3, t3))) continue;
                    }
                    if (untangle_on(s) &&
                        ee_ccd_untangling(s, m->layer[u0], m->layer[j], u0, j, s->w[u0] == 0 && s->w[u1] == 0,
                                          s->w[j] == 0 && s->w[vj] == 0, p_u0o, p_u1o, p_jo, p_vjo, p_u0, p_u1, p_j, p_vj))
                        continue;
                    st2 st0 = closest_segs(p_u0o, p_u1o, p_jo, p_vjo);
                    v3 delta0 = sub(vmix(p_u0o, p_u1o, st0.s), vmix(p_jo, p_vjo, st0.t));
                    real dist0 = len3(delta0);
                    if (dist0 > EPS) {
                        v3 n0 = dvs(delta0, dist0);
                        st2 st1 = closest_segs(p_u0, p_u1, p_j, p_vj);
                        v3 delta1 = sub(vmix(p_u0, p_u1, st1.s), vmix(p_j, p_vj, st1.t));
                        real dist1 = len3(delta1);
                        v3 dend = sub(vmix(p_u0, p_u1, ss), vmix(p_j, p_vj, ts));
                        real h_end = dot(dend, n0);
                        int crossover = (h_end <= 0) || (dot(delta1, n0) <= 0) ||
                                        (bnd_pair && dist1 >= eff_star && tau_star < 1);
                        if (h_end < eff_star && crossover) {
                            s->stats[3]++;
                            real pen = eff_star - h_end;
                            int b_has_pin = s->w[j] <= 0 || s->w[vj] <= 0;
                            if (a_has_pin && b_has_pin) continue;
                            int a_all = s->w[u0] == 0 && s->w[u1] == 0;
                            int b_all = s->w[j] == 0 && s->w[vj] == 0;
                            int cnt_pin = !(a_all || b_all);
                            if (b_has_pin) {
                                v3 em = mul(add(sub(p_j, p_jo), sub(p_vj, p_vjo)), (real)0.5f);
                                real a = 1 - ss, b = ss, w0 = s->w[u0], w1 = s->w[u1];
                                real den = a * a * w0 + b * b * w1;
                                if (den > EPS) {
                                    v3 pb = add(mul(n0, pen), mul(em, (real)0.5f));
                                    if (w0 > 0) add_ccd_ee(&s->acc[u0], mul(pb, rmin((a * w0) / den, 2)), cnt_pin);
                                    if (w1 > 0) add_ccd_ee(&s->acc[u1], mul(pb, rmin((b * w1) / den, 2)), cnt_pin);
                                }
                            } else if (a_has_pin) {
                                v3 em = mul(add(sub(p_u0, p_u0o), sub(p_u1, p_u1o)), (real)0.5f);
                                real a = 1 - ts, b = ts, w0 = s->w[j], w1 = s->w[vj];
                                real den = a * a * w0 + b * b * w1;
                                if (den > EPS) {
                                    v3 pb = add(mul(neg(n0), pen), mul(em, (real)0.5f));
                                    if (w0 > 0) add_ccd_ee(&s->acc[j], mul(pb, rmin((a * w0) / den, 2)), cnt_pin);
                                    if (w1 > 0) add_ccd_ee(&s->acc[vj], mul(pb, rmin((b * w1) / den, 2)), cnt_pin);
                                }
                            } else {
                                real wu0 = s->w[u0], wu1 = s->w[u1], wj = s->w[j], wvj = s->w[vj];
                                real we1 = (1 - ss) * (1 - ss) * wu0 + ss * ss * wu1;
                                real we2 = (1 - ts) * (1 - ts) * wj + ts * ts * wvj;
                                real wt = we1 + we2;
                                if (wt > EPS) {
                                    v3 tot = mul(n0, pen);
                                    if (wu0 > 0) add_ccd_ee(&s->acc[u0], mul(tot, (1 - ss) * wu0 / wt), 1);
                                    if (wu1 > 0) add_ccd_ee(&s->acc[u1], mul(tot, ss * wu1 / wt), 1);
                                    if (wj > 0) add_ccd_ee(&s->acc[j], mul(neg(tot), (1 - ts) * wj / wt), 1);
                                    if (wvj > 0) add_ccd_ee(&s->acc[vj], mul(neg(tot), ts * wvj / wt), 1);
                                }
                            }
                            continue;
                        }
                    }
                }
                /* discrete proximity */
                st2 st; v3 delta_ee; real dsq, eff;
                if (has_cached) { st = st1c; delta_ee = d1c; dsq = dsqc; eff = effc; }
                else {
                    st = closest_segs(p_u0, p_u1, p_j, p_vj);
                    v3 pt1 = add(p_u0, mul(sub(p_u1, p_u0), st.s));
                    v3 pt2 = add(p_j, mul(sub(p_vj, p_j), st.t));
                    delta_ee = sub(pt1, pt2);
                    dsq = dot(delta_ee, delta_ee);
                    eff = rmix(th_u0, th_u1, st.s) + rmix(th_j, th_vj, st.t);
                }
                if (dsq >= eff * eff || dsq <= EPS * EPS) continue;
                if (s->exclude_neighbors) {
                    v3 d0 = sub(p_u0, p_j), d1 = sub(p_u0, p_vj), d2 = sub(p_u1, p_j), d3 = sub(p_u1, p_vj);
                    if (near_folded_sq(m, u0, j, dot(d0, d0)) || near_folded_sq(m, u0, vj, dot(d1, d1)) ||
                        near_folded_sq(m, u1, j, dot(d2, d2)) || near_folded_sq(m, u1, vj, dot(d3, d3))) continue;
                }
                real dist = RSQRT(dsq);
                if (untangle_on(s) && outer_edge_behind_inner(s, m->layer[u0], m->layer[j], u0, u1, j, vj,
                                                              s->w[u0] <= 0 && s->w[u1] <= 0,
                                                              s->w[j] <= 0 && s->w[vj] <= 0, delta_ee))
                    continue;
                s->stats[4]++;
                v3 n = dvs(delta_ee, dist);
                real pen = eff - dist;
                real sp = st.s, tp = st.t;
                int b_has_pin = s->w[j] <= 0 || s->w[vj] <= 0;
                if (a_has_pin && b_has_pin) continue;
                if (b_has_pin) {
                    v3 em = mul(add(sub(p_j, p_jo), sub(p_vj, p_vjo)), (real)0.5f);
                    real a = 1 - sp, b = sp, w0 = s->w[u0], w1 = s->w[u1];
                    real den = a * a * w0 + b * b * w1;
                    if (den > EPS) {
                        v3 pb = add(mul(n, pen), mul(em, (real)0.5f));
                        if (w0 > 0) add_pin_disp(&s->acc[u0], mul(pb, rmin((a * w0) / den, 2)));
                        if (w1 > 0) add_pin_disp(&s->acc[u1], mul(pb, rmin((b * w1) / den, 2)));
                    }
                } else if (a_has_pin) {
                    v3 em = mul(add(sub(p_u0, p_u0o), sub(p_u1, p_u1o)), (real)0.5f);
                    real a = 1 - tp, b = tp, w0 = s->w[j], w1 = s->w[vj];
                    real den = a * a * w0 + b * b * w1;
                    if (den > EPS) {
                        v3 pb = add(mul(neg(n), pen), mul(em, (real)0.5f));
                        if (w0 > 0) add_pin_disp(&s->acc[j], mul(pb, rmin((a * w0) / den, 2)));
                        if (w1 > 0) add_pin_disp(&s->acc[vj], mul(pb, rmin((b * w1) / den, 2)));
                    }
                } else {
                    real wu0 = s->w[u0], wu1 = s->w[u1], wj = s->w[j], wvj = s->w[vj];
                    real we1 = (1 - sp) * (1 - sp) * wu0 + sp * sp * wu1;
                    real we2 = (1 - tp) * (1 - tp) * wj + tp * tp * wvj;
                    real wt = we1 + we2;
                    if (wt > EPS) {
                        v3 td = mul(n, pen * (real)0.6f);
                        if (wu0 > 0) add_disp_ee(&s->acc[u0], mul(td, (1 - sp) * wu0 / wt));
                        if (wu1 > 0) add_disp_ee(&s->acc[u1], mul(td, sp * wu1 / wt));
                        if (wj > 0) add_disp_ee(&s->acc[j], mul(neg(td), (1 - tp) * wj / wt));
                        if (wvj > 0) add_disp_ee(&s->acc[vj], mul(neg(td), tp * wvj / wt));
                    }
                }
            }
        }
    }
}

/* self_collision_apply.wgsl:57-177 */
static void k_sc_apply(orc_sim *s) {
    const real EPS = (real)1e-7f;
    const real INV = (real)(1.0f / 1000000.0f);
    for (uint32_t i = 0; i < s->nv; i++) {
        accum_t *a = &s->acc[i];
        int pinned = s->w[i] <= 0;
        v3 fin = mk(0, 0, 0);
        if (a->count > 0) {
            if (!pinned) {
                v3 tot = mul(mk((real)a->dx, (real)a->dy, (real)a->dz), INV);
                v3 eff = dvs(tot, rmax((real)1, RSQRT((real)a->count)));
                if (s->enable_relief) eff = mul(eff, s->relief_factor);
                real pl = len3(eff);
                real ratio = s->max_disp_ratio > 0 ? s->max_disp_ratio : (real)0.5f;
                real ms = rmin((real)s->m->local_len[i] * ratio, (real)0.10f);
                if (pl > ms && pl > EPS) eff = mul(dvs(eff, pl), ms);
                fin = add(fin, eff);
            }
        }
        v3 avg_pin = mk(0, 0, 0);
        if (a->pcount > 0 && !pinned)
            avg_pin = dvs(mul(mk((real)a->pdx, (real)a->pdy, (real)a->pdz), INV), (real)a->pcount);
        if (!pinned) {
            v3 cp = mul(mk((real)a->cpx, (real)a->cpy, (real)a->cpz), INV);
            v3 cn = mul(mk((real)a->cnx, (real)a->cny, (real)a->cnz), INV);
            v3 env = add(cp, cn);
            if (a->ee_count >= 4) {
                real sc = rmax((real)0.5f, RSQRT((real)3.0f / (real)a->ee_count));
                env = mul(env, sc);
            }
            v3 z = mk(0, 0, 0);
            v3 hp = vmax(vmax(env, z), vmax(avg_pin, z));
            v3 hn = vmin(vmin(env, z), vmin(avg_pin, z));
            fin = add(fin, add(hp, hn));
        }
        if (len3(fin) > EPS) s->p[i] = add(s->p[i], fin);
        memset(a, 0, sizeof(*a));
    }
}

/* ---------------- pair cache (self_collision_algorithm 1) ---------------- */
void orc_sim_set_pair_cache(orc_sim *s, int on, int final_fallback, uint32_t quota_vt, uint32_t quota_ee,
                            uint32_t margin_mode, float safety, float hscale, float hmax, int brute) {
    s->pair_cache = on; s->pair_final_fallback = final_fallback;
    s->pair_quota_vt = quota_vt < 1 ? 1 : quota_vt > 8 ? 8 : quota_vt;
    s->pair_quota_ee = quota_ee < 1 ? 1 : quota_ee > 8 ? 8 : quota_ee;
    s->pair_margin_mode = margin_mode ? 1 : 0;
    s->pair_safety = safety; s->pair_hscale = hscale; s->pair_hmax = hmax;
    s->pair_brute = brute;
}
/* The held pairs in slot order and the counters of the last collect: returns the number of V-T
 * (which 0) or E-E (which 1) pairs, written to out (4 indices each) when out is not NULL. */
uint32_t orc_sim_active_pairs(const orc_sim *s, int which, uint32_t *out, uint32_t *counters) {
    const uint32_t *src = which ? s->pair_ee : s->pair_vt;
    uint32_t q = which ? s->pair_quota_ee : s->pair_quota_vt, n = 0;
    for (size_t k = 0; k < (size_t)s->nv * q; k++) {
        if (src[4 * k] == 0xFFFFFFFFu) continue;
        if (out) memcpy(out + 4 * (size_t)n, src + 4 * k, 16);
        n++;
    }
    if (counters) memcpy(counters, s->pair_counters, 16);
    return n;
}
/* self_collision_collect_pairs.wgsl:122-156 (no a == b shortcut) */
static int near_folded_pairs(const orc_mesh *m, uint32_t a, uint32_t b, real cur) {
    uint32_t lo = m->th_off[a], hi = m->th_off[a + 1];
    while (lo < hi) {
        uint32_t mid = lo + (hi - lo) / 2, val = m->th_idx[mid];
        if (val == b) { real l0 = m->th_rest[mid]; return cur < l0 * (real)0.5f ? 0 : 1; }
        if (val < b) lo = mid + 1; else hi = mid;
    }
    return 0;
}
static real pair_horizon(const orc_sim *s, uint32_t v) {
    return rmin(len3(s->v[v]) * s->pair_dt_frame * s->pair_hscale, s->pair_hmax);
}
static int32_t iabs_wrap(int32_t d) { return d < 0 ? (int32_t)(0u - (uint32_t)d) : d; }
/* can_ee_anchor_reach, collect_pairs.wgsl:250-279 */
static int ee_anchor_reach(const orc_sim *s, uint32_t u, uint32_t w) {
    const orc_mesh *m = s->m;
    if (s->w[u] <= 0) return 0;
    int32_t cu[3], cw[3];
    cell_of(s->p[u], s->cell_size, cu); cell_of(s->p[w], s->cell_size, cw);
    for (int a = 0; a < 3; a++) if (iabs_wrap((int32_t)((uint32_t)cu[a] - (uint32_t)cw[a])) > 1) return 0;
    v3 d = sub(s->p[u], s->p[w]);
    real dsq = dot(d, d);
    real mf = (real)(float)s->pair_margin_mode;
    real hu = pair_horizon(s, u) * mf, hw = pair_horizon(s, w) * mf;
    real bound = (real)m->thick[u] + (real)m->thick[w] + ((real)m->local_len[u] + (real)m->local_len[w]) * (real)1.3f +
                 s->pair_safety + hu + hw;
    if (dsq > bound * bound) return 0;
    if (m->island[u] == m->island[w] && s->exclude_neighbors)
        if (near_folded_pairs(m, u, w, RSQRT(rmax(dsq, 0)))) return 0;
    return 1;
}
static int smaller_anchor(const orc_sim *s, uint32_t u, uint32_t w, uint32_t cu, uint32_t cw) {
    if (u > cu || (u == cu && w >= cw)) return 0;
    return ee_anchor_reach(s, u, w);
}
typedef struct { real d[8]; uint32_t e[8][4]; uint32_t n; } topk_t;
/* collect_pairs.wgsl:427-476 */
static void topk_offer(topk_t *L, uint32_t quota, real dist, const uint32_t ent[4], uint32_t *eval, uint32_t *drop) {
    for (uint32_t q = 0; q < L->n; q++)
        if (L->e[q][1] == ent[1] && L->e[q][2] == ent[2] && L->e[q][3] == ent[3]) return;
    (*eval)++;
    uint32_t pos = L->n;
    for (uint32_t q = 0; q < L->n; q++) {
        const uint32_t *o = L->e[q];
        int tie = dist == L->d[q] && (ent[1] < o[1] || (ent[1] == o[1] && (ent[2] < o[2] || (ent[2] == o[2] && ent[3] < o[3]))));
        if (dist < L->d[q] || tie) { pos = q; break; }
    }
    if (pos >= quota) { (*drop)++; return; }
    uint32_t up = L->n < quota - 1 ? L->n : quota - 1;
    for (uint32_t q = up; q > pos; q--) { L->d[q] = L->d[q - 1]; memcpy(L->e[q], L->e[q - 1], 16); }
    L->d[pos] = dist; memcpy(L->e[pos], ent, 16);
    if (L->n < quota) L->n++;
}
/* self_collision_collect_pairs.wgsl:294-661.  The candidates j are the vertices of the 27 cells
 * around vertex i, cell by cell (x outer, z inner), in ascending index within a cell: from the
 * hash tables, or (pair_brute) from a scan of every vertex, which must give the same lists. */
static void k_pair_collect(orc_sim *s) {
    const orc_mesh *m = s->m;
    const real EPS = (real)1e-7f;
    const real cell = s->cell_size;
    const real mf = (real)(float)s->pair_margin_mode;
    uint32_t nv = s->nv;
    memset(s->pair_counters, 0, sizeof(s->pair_counters));
    uint32_t *cand = (uint32_t *)malloc((nv ? nv : 1) * 4);
    for (uint32_t index = 0; index < nv; index++) {
        int pinned_i = s->w[index] <= 0;
        v3 p_i = s->p[index];
        real thick_i = (real)m->thick[index], lli = (real)m->local_len[index];
        int32_t base[3]; cell_of(p_i, cell, base);
        topk_t VT, EE; VT.n = EE.n = 0;
        uint32_t vt_eval = 0, vt_drop = 0, ee_eval = 0, ee_drop = 0;
        real hi_eff = pair_horizon(s, index) * mf;
        for (int gx = -1; gx <= 1; gx++)
        for (int gy = -1; gy <= 1; gy++)
        for (int gz = -1; gz <= 1; gz++) {
            int32_t cc[3] = {(int32_t)((uint32_t)base[0] + (uint32_t)gx), (int32_t)((uint32_t)base[1] + (uint32_t)gy),
                             (int32_t)((uint32_t)base[2] + (uint32_t)gz)};
            uint32_t nc = 0;
            if (s->pair_brute) {
                for (uint32_t j = 0; j < nv; j++) cand[nc++] = j;
            } else {
                uint32_t h = hash_coords(cc, s->table_size);
                for (uint32_t k = s->hstarts[h]; k < s->hstarts[h + 1]; k++) cand[nc++] = s->hsorted[k];
            }
            for (uint32_t kk = 0; kk < nc; kk++) {
                uint32_t j = cand[kk];
                if (j >= nv || j == index) continue;
                v3 p_j = s->p[j];
                int32_t jc[3]; cell_of(p_j, cell, jc);
                if (jc[0] != cc[0] || jc[1] != cc[1] || jc[2] != cc[2]) continue;
                int pinned_j = s->w[j] <= 0;
                if (pinned_i && pinned_j) continue;
                real hj = pair_horizon(s, j) * mf;
                v3 dvv = sub(p_i, p_j);
                real dsq = dot(dvv, dvv);
                real eff = thick_i + (real)m->thick[j];
                real llj = (real)m->local_len[j];
                real broad = eff + (lli + llj) * (real)1.3f + s->pair_safety + hi_eff + hj;
                if (!(dsq <= broad * broad)) continue;
                int same_island = m->island[index] == m->island[j];
                real dvv_len = RSQRT(rmax(dsq, 0));
                if (same_island && s->exclude_neighbors && near_folded_pairs(m, index, j, dvv_len)) continue;
                real check = eff + s->pair_safety + hi_eff + hj;
                /* V-T */
                real vt_bound = eff + llj * (real)1.3f + s->pair_safety + hi_eff + hj;
                int near2 = s->exclude_neighbors && within_two_hops(m, index, j);
                if (dsq <= vt_bound * vt_bound && !near2) {
                    uint32_t ss = m->star_off[j], se = m->star_off[j + 1];
                    uint32_t ns = se > ss ? (se - ss < 8 ? se - ss : 8) : 0;
                    for (uint32_t st = 0; st < ns; st++) {
                        uint32_t v0 = m->star_v0[ss + st], v1 = m->star_v1[ss + st];
                        if (index == v0 || index == v1 || v0 >= nv || v1 >= nv) continue;
                        uint32_t c0 = j, c1 = v0, c2 = v1;
                        if (v0 < j && v0 < v1) { c0 = v0; c1 = v1; c2 = j; }
                        else if (v1 < j && v1 < v0) { c0 = v1; c1 = j; c2 = v0; }
                        cbary cb = closest_bary(p_i, s->p[c0], s->p[c1], s->p[c2]);
                        v3 dvt = sub(p_i, cb.pt);
                        real dsq_vt = dot(dvt, dvt);
                        if (dsq_vt < check * check) {
                            if (s->exclude_neighbors && (within_two_hops(m, index, v0) || within_two_hops(m, index, v1)))
                                continue;
                            uint32_t ent[4] = {index, c0, c1, c2};
                            topk_offer(&VT, s->pair_quota_vt, dsq_vt, ent, &vt_eval, &vt_drop);
                        }
                    }
                }
                /* E-E (its bound equals the broad bound) */
                uint32_t ai0 = m->adj_off[index], ai1 = m->adj_off[index + 1];
                uint32_t aj0 = m->adj_off[j], aj1 = m->adj_off[j + 1];
                if (!(ai0 < ai1 && aj0 < aj1)) continue;
                uint32_t nei = ai1 - ai0 < 8 ? ai1 - ai0 : 8, nej = aj1 - aj0 < 8 ? aj1 - aj0 : 8;
                for (uint32_t ei = 0; ei < nei; ei++) {
                    uint32_t ui = m->adj_idx[ai0 + ei];
                    if (ui >= nv) continue;
                    v3 p_ui = s->p[ui];
                    real llui = (real)m->local_len[ui];
                    v3 dui = sub(p_ui, p_i);
                    real len_sq_i = dot(dui, dui);
                    real max_len_i = (lli + llui) * (real)1.5f;
                    if (len_sq_i > max_len_i * max_len_i || len_sq_i <= (real)1e-12f) continue;
                    for (uint32_t ej = 0; ej < nej; ej++) {
                        uint32_t vj = m->adj_idx[aj0 + ej];
                        if (vj >= nv) continue;
                        if (ui == j || ui == vj || index == vj) continue;
                        uint32_t c_j = j < vj ? j : vj, c_vj = j < vj ? vj : j;
                        int smaller = smaller_anchor(s, index, vj, index, j) || smaller_anchor(s, ui, j, index, j) ||
                                      smaller_anchor(s, ui, vj, index, j) || smaller_anchor(s, j, index, index, j) ||
                                      smaller_anchor(s, j, ui, index, j) || smaller_anchor(s, vj, index, index, j) ||
                                      smaller_anchor(s, vj, ui, index, j);
                        if (pinned_i || smaller) continue;
                        v3 p_vj = s->p[vj];
                        real llvj = (real)m->local_len[vj];
                        v3 djv = sub(p_vj, p_j);
                        real len_sq_j = dot(djv, djv);
                        real max_len_j = (llj + llvj) * (real)1.5f;
                        if (len_sq_j > max_len_j * max_len_j || len_sq_j <= (real)1e-12f) continue;
                        v3 p_cj = j < vj ? p_j : p_vj, p_cvj = j < vj ? p_vj : p_j;
                        st2 st = closest_segs(p_i, p_ui, p_cj, p_cvj);
                        v3 pt1 = add(p_i, mul(sub(p_ui, p_i), st.s));
                        v3 pt2 = add(p_cj, mul(sub(p_cvj, p_cj), st.t));
                        v3 dee = sub(pt1, pt2);
                        real dsq_ee = dot(dee, dee);
                        if (!(dsq_ee < check * check && dsq_ee > EPS * EPS)) continue;
                        if (s->exclude_neighbors) {
                            if (near_folded_pairs(m, index, vj, len3(sub(p_i, p_vj))) ||
                                near_folded_pairs(m, ui, j, len3(sub(p_ui, p_j))) ||
                                near_folded_pairs(m, ui, vj, len3(sub(p_ui, p_vj))))
                                continue;
                        }
                        uint32_t ent[4] = {index, ui, c_j, c_vj};
                        topk_offer(&EE, s->pair_quota_ee, dsq_ee, ent, &ee_eval, &ee_drop);
                    }
                }
            }
        }
        for (uint32_t k = 0; k < s->pair_quota_vt; k++) {
            uint32_t *o = s->pair_vt + 4 * ((size_t)index * s->pair_quota_vt + k);
            if (k < VT.n) memcpy(o, VT.e[k], 16); else { o[0] = 0xFFFFFFFFu; o[1] = o[2] = o[3] = 0; }
        }
        for (uint32_t k = 0; k < s->pair_quota_ee; k++) {
            uint32_t *o = s->pair_ee + 4 * ((size_t)index * s->pair_quota_ee + k);
            if (k < EE.n) memcpy(o, EE.e[k], 16); else { o[0] = 0xFFFFFFFFu; o[1] = o[2] = o[3] = 0; }
        }
        s->pair_counters[0] += vt_eval; s->pair_counters[1] += ee_eval;
        s->pair_counters[2] += vt_drop; s->pair_counters[3] += ee_drop;
    }
    free(cand);
}
/* self_collision_solve_vt.wgsl:227-414 */
/* record_untangle_cand (self_collision_solve_vt.wgsl:97-107): atomic max of
 * (1023 - distance rank) << 22 | (j + 1) */
static void record_untangle(orc_sim *s, uint32_t vi, uint32_t j, real dist, real cap) {
    if (vi >= s->nv || j >= 0x3FFFFEu || dist >= cap) return;
    real ratio = rclamp(dist / rmax(cap, (real)1e-7f), 0, 1);
    uint32_t rank = 1023u - (uint32_t)(float)(ratio * (real)1023.0f);
    uint32_t packed = (rank << 22) | ((j + 1u) & 0x3FFFFFu);
    if (packed > s->acc[vi].untangle) s->acc[vi].untangle = packed;
}
static void k_pair_solve_vt(orc_sim *s) {
    const orc_mesh *m = s->m;
    const real EPS = (real)1e-7f;
    uint32_t nv = s->nv;
    for (size_t k = 0; k < (size_t)nv * s->pair_quota_vt; k++) {
        const uint32_t *pr = s->pair_vt + 4 * k;
        if (pr[0] == 0xFFFFFFFFu) continue;
        uint32_t index = pr[0], j = pr[1], v0 = pr[2], v1 = pr[3];
        if (index >= nv || j >= nv || v0 >= nv || v1 >= nv) continue;
        if (index == j || index == v0 || index == v1 || v0 == v1 || j == v0 || j == v1) continue;
        int pinned_i = s->w[index] <= 0;
        v3 p_i = s->p[index], xo_i = s->x[index];
        v3 p_j = s->p[j], p_v0 = s->p[v0], p_v1 = s->p[v1];
        v3 pj_old = s->x[j], p_v0o = s->x[v0], p_v1o = s->x[v1];
        real eff = (real)m->thick[index] + (real)m->thick[j];
        v3 tri_move = dvs(add(add(sub(p_j, pj_old), sub(p_v0, p_v0o)), sub(p_v1, p_v1o)), (real)3.0f);
        real wi = s->w[index], wj = s->w[j], w0 = s->w[v0], w1 = s->w[v1];
        int tri_pin = wj <= 0 || w0 <= 0 || w1 <= 0;
        const int ut = untangle_on(s) && !pinned_i && m->layer[index] > m->layer[j];
        v3 ni_unit = mk(0, 0, 0);
        real nil = 0;
        if (ut) {
            nil = len3(s->normals[index]);
            if (nil > (real)0.5f) ni_unit = dvs(s->normals[index], rmax(nil, EPS));
        }
        v3 rel_old = add(xo_i, tri_move);
        real ht = 0, hu = 0, hv = 0;
        if (seg_tri(rel_old, p_i, p_j, p_v0, p_v1, &ht, &hu, &hv)) {
            v3 move = sub(p_i, rel_old);
            v3 hit_pt = add(rel_old, mul(move, ht));
            v3 tcr = cross(sub(p_v0, p_j), sub(p_v1, p_j));
            real tl = len3(tcr);
            v3 tn = tl > (real)1e-12f ? dvs(tcr, tl) : mk(0, 1, 0);
            real h_old_raw = dot(sub(rel_old, p_j), tn);
            if (ut && tl > (real)1e-12f && (tri_pin || nil <= (real)0.5f || dot(ni_unit, tn) > 0) && h_old_raw < 0 &&
                dot(move, tn) > 0) {
                record_untangle(s, index, j, len3(sub(p_i, hit_pt)), eff * (real)2.5f);
                continue;
            }
            real side = (real)fabs((double)h_old_raw) > (real)1e-6f ? h_old_raw : dot(neg(move), tn);
            v3 n_side = side >= 0 ? tn : neg(tn);
            real h_old = rmax(dot(sub(rel_old, hit_pt), n_side), 0);
            real h_safe = rmax(h_old, eff * (real)0.25f);
            v3 total = sub(add(hit_pt, mul(n_side, h_safe)), p_i);
            real pen_soft = rmax(0, eff - h_safe);
            v3 bc = mk(rclamp(1 - hu - hv, 0, 1), rclamp(hu, 0, 1), rclamp(hv, 0, 1));
            if (tri_pin) {
                if (!pinned_i) add_ccd(&s->acc[index], add(total, mul(n_side, pen_soft)));
            } else if (pinned_i) {
                real wt = bc.x * bc.x * wj + bc.y * bc.y * w0 + bc.z * bc.z * w1;
                if (wt > EPS) {
                    v3 full = add(total, mul(n_side, pen_soft));
                    if (wj > 0) add_ccd(&s->acc[j], mul(neg(full), bc.x * wj / wt));
                    if (w0 > 0) add_ccd(&s->acc[v0], mul(neg(full), bc.y * w0 / wt));
                    if (w1 > 0) add_ccd(&s->acc[v1], mul(neg(full), bc.z * w1 / wt));
                }
            } else {
                real wt = bc.x * bc.x * wj + bc.y * bc.y * w0 + bc.z * bc.z * w1;
                real wtot = wi + wt;
                if (wtot > EPS) {
                    add_ccd(&s->acc[index], mul(total, wi / wtot));
                    if (wj > 0) add_ccd(&s->acc[j], mul(neg(total), bc.x * wj / wtot));
                    if (w0 > 0) add_ccd(&s->acc[v0], mul(neg(total), bc.y * w0 / wtot));
                    if (w1 > 0) add_ccd(&s->acc[v1], mul(neg(total), bc.z * w1 / wtot));
                    if (pen_soft > EPS) {
                        v3 tp = mul(n_side, pen_soft * (real)0.6f);
                        add_disp(&s->acc[index], mul(tp, wi / wtot));
                        if (wj > 0) add_disp(&s->acc[j], mul(neg(tp), bc.x * wj / wtot));
                        if (w0 > 0) add_disp(&s->acc[v0], mul(neg(tp), bc.y * w0 / wtot));
                        if (w1 > 0) add_disp(&s->acc[v1], mul(neg(tp), bc.z * w1 / wtot));
                    }
                }
            }
            continue;
        }
        cbary cb = closest_bary(p_i, p_j, p_v0, p_v1);
        v3 dvt = sub(p_i, cb.pt);
        real dist = len3(dvt);
        if (ut) {
            v3 tcr = cross(sub(p_v0, p_j), sub(p_v1, p_j));
            real tl = len3(tcr);
            v3 tn = tl > EPS ? dvs(tcr, rmax(tl, EPS)) : mk(0, 0, 0);
            real h = tl > EPS ? dot(dvt, tn) : 0;
            if (tl > EPS && (tri_pin || nil <= (real)0.5f || dot(ni_unit, tn) > 0) && h < 0) {
                record_untangle(s, index, j, dist, eff * (real)2.5f);
                continue;
            }
        }
        if (pinned_i || !(dist < eff) || !(dist > EPS)) continue;
        v3 nvt = dvs(dvt, dist);
        real pen = eff - dist;
        if (tri_pin) { add_pin_disp(&s->acc[index], add(mul(nvt, pen), mul(tri_move, (real)0.5f))); continue; }
        v3 bc = cb.bary;
        real wt = bc.x * bc.x * wj + bc.y * bc.y * w0 + bc.z * bc.z * w1;
        real wtot = wi + wt;
        if (wtot > EPS) {
            v3 tp = mul(nvt, pen * (real)0.6f);
            add_disp(&s->acc[index], mul(tp, wi / wtot));
            if (wj > 0) add_disp(&s->acc[j], mul(neg(tp), bc.x * wj / wtot));
            if (w0 > 0) add_disp(&s->acc[v0], mul(neg(tp), bc.y * w0 / wtot));
            if (w1 > 0) add_disp(&s->acc[v1], mul(neg(tp), bc.z * w1 / wtot));
        }
    }
}
/* self_collision_solve_ee.wgsl:231-589 */
static void k_pair_solve_ee(orc_sim *s) {
    const orc_mesh *m = s->m;
    const real EPS = (real)1e-7f;
    uint32_t nv = s->nv;
    for (size_t k = 0; k < (size_t)nv * s->pair_quota_ee; k++) {
        const uint32_t *pr = s->pair_ee + 4 * k;
        if (pr[0] == 0xFFFFFFFFu) continue;
        uint32_t u0 = pr[0], u1 = pr[1], j = pr[2], vj = pr[3];
        if (u0 >= nv || u1 >= nv || j >= nv || vj >= nv) continue;
        if (u0 == u1 || j == vj || u0 == j || u0 == vj || u1 == j || u1 == vj) continue;
        real wu0 = s->w[u0], wu1 = s->w[u1], wj = s->w[j], wvj = s->w[vj];
        int a_has_pin = wu0 <= 0 || wu1 <= 0, b_has_pin = wj <= 0 || wvj <= 0;
        if (a_has_pin && b_has_pin) continue;
        v3 p_u0 = s->p[u0], p_u1 = s->p[u1], p_j = s->p[j], p_vj = s->p[vj];
        v3 p_u0o = s->x[u0], p_u1o = s->x[u1], p_jo = s->x[j], p_vjo = s->x[vj];
        real th_u0 = m->thick[u0], th_u1 = m->thick[u1], th_j = m->thick[j], th_vj = m->thick[vj];
        st2 st1c = closest_segs(p_u0, p_u1, p_j, p_vj);
        v3 d1c = sub(vmix(p_u0, p_u1, st1c.s), vmix(p_j, p_vj, st1c.t));
        real dsqc = dot(d1c, d1c);
        real effc = rmix(th_u0, th_u1, st1c.s) + rmix(th_j, th_vj, st1c.t);
        real max_eff = rmax(th_u0, th_u1) + rmax(th_j, th_vj);
        v3 du0 = sub(p_u0, p_u0o), du1 = sub(p_u1, p_u1o), dj = sub(p_j, p_jo), dvj = sub(p_vj, p_vjo);
        v3 va = mul(add(du0, du1), (real)0.5f), vb = mul(add(dj, dvj), (real)0.5f);
        real rel = len3(sub(va, vb)) + (real)0.5f * (len3(sub(du1, du0)) + len3(sub(dvj, dj)));
        real reach = max_eff + rel;
        if (dsqc > reach * reach) continue;
        v3 d0g = sub(vmix(p_u0o, p_u1o, st1c.s), vmix(p_jo, p_vjo, st1c.t));
        int done = 0;
        if ((rel > effc * (real)0.5f) || (dot(d0g, d1c) <= 0)) {
            real min_sq = (real)1e20f, best_tau = 0, half_span = (real)0.125f;
            st2 min_st = {0, 0};
            real prev_vol = seg_volume(p_u0o, p_u0, p_u1o, p_u1, p_jo, p_j, p_vjo, p_vj, 0);
            real zc = -1;
            for (uint32_t kk = 0; kk <= 8; kk++) {
                real tk = (real)(float)kk * (real)0.125f;
                st2 stk;
                real dk = seg_dist_sq(p_u0o, p_u0, p_u1o, p_u1, p_jo, p_j, p_vjo, p_vj, tk, &stk);
                if (dk < min_sq) { min_sq = dk; best_tau = tk; min_st = stk; }
                if (kk > 0) {
                    real vol = seg_volume(p_u0o, p_u0, p_u1o, p_u1, p_jo, p_j, p_vjo, p_vj, tk);
                    if (prev_vol * vol <= 0 && zc < 0) zc = tk - (real)0.0625f;
                    prev_vol = vol;
                }
            }
            if (zc >= 0) {
                st2 stz;
                real dzs = seg_dist_sq(p_u0o, p_u0, p_u1o, p_u1, p_jo, p_j, p_vjo, p_vj, zc, &stz);
                if (dzs < min_sq) { min_sq = dzs; best_tau = zc; half_span = (real)0.0625f; min_st = stz; }
            }
            real tlo = rmax(0, best_tau - half_span), thi = rmin(1, best_tau + half_span);
            real dsq_star = min_sq;
            st2 st_star = min_st;
            for (int it = 0; it < 6; it++) {
                real m1 = tlo + (thi - tlo) * (real)(1.0f / 3.0f);
                real m2 = thi - (thi - tlo) * (real)(1.0f / 3.0f);
                st2 s1, s2;
                real e1 = seg_dist_sq(p_u0o, p_u0, p_u1o, p_u1, p_jo, p_j, p_vjo, p_vj, m1, &s1);
                real e2 = seg_dist_sq(p_u0o, p_u0, p_u1o, p_u1, p_jo, p_j, p_vjo, p_vj, m2, &s2);
                if (e1 < dsq_star) { dsq_star = e1; st_star = s1; }
                if (e2 < dsq_star) { dsq_star = e2; st_star = s2; }
                if (e1 < e2) thi = m2; else tlo = m1;
            }
            real ss = st_star.s, ts = st_star.t;
            real eff_star = rmix(th_u0, th_u1, ss) + rmix(th_j, th_vj, ts);
            if (dsq_star >= eff_star * eff_star) continue;
            if (untangle_on(s) &&
                ee_ccd_untangling(s, m->layer[u0], m->layer[j], u0, j, s->w[u0] == 0 && s->w[u1] == 0,
                                  s->w[j] == 0 && s->w[vj] == 0, p_u0o, p_u1o, p_jo, p_vjo, p_u0, p_u1, p_j, p_vj))
                continue;
            st2 st0 = closest_segs(p_u0o, p_u1o, p_jo, p_vjo);
            v3 delta0 = sub(vmix(p_u0o, p_u1o, st0.s), vmix(p_jo, p_vjo, st0.t));
            real dist0 = len3(delta0);
            if (dist0 > EPS) {
                v3 n0 = dvs(delta0, dist0);
                real h_end = dot(sub(vmix(p_u0, p_u1, ss), vmix(p_j, p_vj, ts)), n0);
                int crossover = (h_end <= 0) || (dot(d1c, n0) <= 0);
                if (h_end < eff_star && crossover) {
                    real pen = eff_star - h_end;
                    int a_all = wu0 == 0 && wu1 == 0, b_all = wj == 0 && wvj == 0;
                    int cnt_pin = !(a_all || b_all);
                    if (b_has_pin) {
                        v3 em = mul(add(sub(p_j, p_jo), sub(p_vj, p_vjo)), (real)0.5f);
                        real a = 1 - ss, b = ss, den = a * a * wu0 + b * b * wu1;
                        if (den > EPS) {
                            v3 pb = add(mul(n0, pen), mul(em, (real)0.5f));
                            if (wu0 > 0) add_ccd_ee(&s->acc[u0], mul(pb, rmin((a * wu0) / den, 2)), cnt_pin);
                            if (wu1 > 0) add_ccd_ee(&s->acc[u1], mul(pb, rmin((b * wu1) / den, 2)), cnt_pin);
                        }
                    } else if (a_has_pin) {
                        v3 em = mul(add(sub(p_u0, p_u0o), sub(p_u1, p_u1o)), (real)0.5f);
                        real a = 1 - ts, b = ts, den = a * a * wj + b * b * wvj;
                        if (den > EPS) {
                            v3 pb = add(mul(neg(n0), pen), mul(em, (real)0.5f));
                            if (wj > 0) add_ccd_ee(&s->acc[j], mul(pb, rmin((a * wj) / den, 2)), cnt_pin);
                            if (wvj > 0) add_ccd_ee(&s->acc[vj], mul(pb, rmin((b * wvj) / den, 2)), cnt_pin);
                        }
                    } else {
                        real we1 = (1 - ss) * (1 - ss) * wu0 + ss * ss * wu1;
                        real we2 = (1 - ts) * (1 - ts) * wj + ts * ts * wvj;
                        real wt = we1 + we2;
                        if (wt > EPS) {
                            v3 tot = mul(n0, pen);
                            if (wu0 > 0) add_ccd_ee(&s->acc[u0], mul(tot, (1 - ss) * wu0 / wt), 1);
                            if (wu1 > 0) add_ccd_ee(&s->acc[u1], mul(tot, ss * wu1 / wt), 1);
                            if (wj > 0) add_ccd_ee(&s->acc[j], mul(neg(tot), (1 - ts) * wj / wt), 1);
                            if (wvj > 0) add_ccd_ee(&s->acc[vj], mul(neg(tot), ts * wvj / wt), 1);
                        }
                    }
                    done = 1;
                }
            }
        }
        if (done) continue;
        if (dsqc >= effc * effc || dsqc <= EPS * EPS) continue;
        real dist = RSQRT(dsqc);
        if (untangle_on(s) && outer_edge_behind_inner(s, m->layer[u0], m->layer[j], u0, u1, j, vj,
                                                      s->w[u0] <= 0 && s->w[u1] <= 0, s->w[j] <= 0 && s->w[vj] <= 0, d1c))
            continue;
        v3 n = dvs(d1c, dist);
        real pen = effc - dist;
        real sp = st1c.s, tp = st1c.t;
        if (b_has_pin) {
            v3 em = mul(add(sub(p_j, p_jo), sub(p_vj, p_vjo)), (real)0.5f);
            real a = 1 - sp, b = sp, den = a * a * wu0 + b * b * wu1;
            if (den > EPS) {
                v3 pb = add(mul(n, pen), mul(em, (real)0.5f));
                if (wu0 > 0) add_pin_disp(&s->acc[u0], mul(pb, rmin((a * wu0) / den, 2)));
                if (wu1 > 0) add_pin_disp(&s->acc[u1], mul(pb, rmin((b * wu1) / den, 2)));
            }
        } else if (a_has_pin) {
            v3 em = mul(add(sub(p_u0, p_u0o), sub(p_u1, p_u1o)), (real)0.5f);
            real a = 1 - tp, b = tp, den = a * a * wj + b * b * wvj;
            if (den > EPS) {
                v3 pb = add(mul(neg(n), pen), mul(em, (real)0.5f));
                if (wj > 0) add_pin_disp(&s->acc[j], mul(pb, rmin((a * wj) / den, 2)));
                if (wvj > 0) add_pin_disp(&s->acc[vj], mul(pb, rmin((b * wvj) / den, 2)));
            }
        } else {
            real we1 = (1 - sp) * (1 - sp) * wu0 + sp * sp * wu1;
            real we2 = (1 - tp) * (1 - tp) * wj + tp * tp * wvj;
            real wt = we1 + we2;
            if (wt > EPS) {
                v3 td = mul(n, pen * (real)0.6f);
                if (wu0 > 0) add_disp(&s->acc[u0], mul(td, (1 - sp) * wu0 / wt));
                if (wu1 > 0) add_disp(&s->acc[u1], mul(td, sp * wu1 / wt));
                if (wj > 0) add_disp(&s->acc[j], mul(neg(td), (1 - tp) * wj / wt));
                if (wvj > 0) add_disp(&s->acc[vj], mul(neg(td), tp * wvj / wt));
            }
        }
    }
}

/* compute_normals.wgsl:25-58: normalised sum of the star triangles' cross products at x */
static void k_compute_normals(orc_sim *s) {
    const orc_mesh *m = s->m;
    for (uint32_t i = 0; i < s->nv; i++) {
        v3 p_i = s->x[i], sum = mk(0, 0, 0);
        for (uint32_t k = m->star_off[i]; k < m->star_off[i + 1]; k++)
            sum = add(sum, cross(sub(s->x[m->star_v0[k]], p_i), sub(s->x[m->star_v1[k]], p_i)));
        real l = len3(sum);
        s->normals[i] = l > (real)1e-7f ? dvs(sum, l) : mk(0, 0, 0);
    }
}

/* closest point on a triangle, position only (self_collision_untangle.wgsl:85-129) */
static v3 closest_pos(v3 p, v3 a, v3 b, v3 c) {
    v3 ab = sub(b, a), ac = sub(c, a), ap = sub(p, a);
    real d1 = dot(ab, ap), d2 = dot(ac, ap);
    if (d1 <= 0 && d2 <= 0) return a;
    v3 bp = sub(p, b);
    real d3 = dot(ab, bp), d4 = dot(ac, bp);
    if (d3 >= 0 && d4 <= d3) return b;
    real vc = d1 * d4 - d3 * d2;
    if (vc <= 0 && d1 >= 0 && d3 <= 0) return add(a, mul(ab, d1 / (d1 - d3)));
    v3 cp = sub(p, c);
    real d5 = dot(ab, cp), d6 = dot(ac, cp);
    if (d6 >= 0 && d5 <= d6) return c;
    real vb = d5 * d2 - d1 * d6;
    if (vb <= 0 && d2 >= 0 && d6 <= 0) return add(a, mul(ac, d2 / (d2 - d6)));
    real va = d3 * d6 - d5 * d4;
    if (va <= 0 && (d4 - d3) >= 0 && (d5 - d6) >= 0) return add(b, mul(sub(c, b), (d4 - d3) / ((d4 - d3) + (d5 - d6))));
    real den = 1 / (va + vb + vc);
    real v = vb * den, w = vc * den;
    return add(add(a, mul(ab, v)), mul(ac, w));
}

/* self_collision_untangle.wgsl:163-394: a free outer-layer vertex with a recorded inner candidate j1
 * takes j1's nearest 1-ring inner neighbour j2 outside its two-hop ring, the nearest triangle of the
 * two stars (ties within 1e-10: the front side, then the larger |height|), and if it lies behind it
 * inside the 45-degree cone and within half a thickness of its normal line, with a compatible vertex
 * normal and behind the interpolated pseudo-normal, is pushed out through the triangle's normal. */
static void k_sc_untangle(orc_sim *s) {
    const orc_mesh *m = s->m;
    const real EPS = (real)1e-7f;
    if (!untangle_on(s)) return;
    for (uint32_t index = 0; index < s->nv; index++) {
        uint32_t cand = s->acc[index].untangle;
        if (!cand) continue;
        s->acc[index].untangle = 0;
        if (s->w[index] <= 0 || m->layer[index] == 0) continue;
        uint32_t j1e = cand & 0x3FFFFFu;
        if (!j1e) continue;
        uint32_t j1 = j1e - 1u;
        if (j1 >= s->nv || j1 == index) continue;
        v3 p_i = s->p[index];
        real thick_i = (real)m->thick[index], llen_i = (real)m->local_len[index];
        uint32_t j2 = 0xFFFFFFFFu;
        real best_nb = (real)1e20f;
        for (uint32_t a = m->adj_off[j1]; a < m->adj_off[j1 + 1]; a++) {
            uint32_t nb = m->adj_idx[a];
            if (nb < s->nv && nb != index && m->layer[nb] < m->layer[index]) {
                v3 d = sub(p_i, s->p[nb]);
                real dsq = dot(d, d);
                if (dsq < best_nb && !within_two_hops(m, index, nb)) { best_nb = dsq; j2 = nb; }
            }
        }
        real best_dsq = (real)1e20f, best_h = 0;
        uint32_t bj = 0xFFFFFFFFu, bv0 = 0, bv1 = 0;
        for (int c = 0; c < 2; c++) {
            uint32_t j = c == 0 ? j1 : j2;
            if (j == 0xFFFFFFFFu) continue;
            v3 p_j = s->p[j];
            real eff = rmax(thick_i + (real)m->thick[j], rmin(llen_i, (real)m->local_len[j]) * (real)0.06f);
            real cap = eff * (real)2.5f, cap_sq = cap * cap;
            for (uint32_t st = m->star_off[j]; st < m->star_off[j + 1]; st++) {
                uint32_t v0 = m->star_v0[st], v1 = m->star_v1[st];
                if (v0 == index || v1 == index) continue;
                if (c == 1 && (v0 == j1 || v1 == j1)) continue;
                v3 p_v0 = s->p[v0], p_v1 = s->p[v1];
                v3 tmin = vmin(vmin(p_j, p_v0), p_v1), tmax = vmax(vmax(p_j, p_v0), p_v1), z = mk(0, 0, 0);
                v3 ad = add(vmax(sub(tmin, p_i), z), vmax(sub(p_i, tmax), z));
                real lim = rmin(cap_sq, best_dsq + (real)1e-10f);
                if (dot(ad, ad) >= lim) continue;
                v3 d = sub(p_i, closest_pos(p_i, p_j, p_v0, p_v1));
                real dsq = dot(d, d);
                if (dsq >= cap_sq) continue;
                if (bj != 0xFFFFFFFFu && dsq > best_dsq + (real)1e-10f) continue;
                v3 tcr = cross(sub(p_v0, p_j), sub(p_v1, p_j));
                real tl = len3(tcr);
                if (tl <= EPS) continue;
                real h = dot(d, dvs(tcr, tl));
                int same_d = (real)fabs((double)(dsq - best_dsq)) <= (real)1e-10f;
                int better_tie = (h >= 0 && best_h < 0) ||
                                 (((h >= 0) == (best_h >= 0)) && (real)fabs((double)h) > (real)fabs((double)best_h));
                if (bj == 0xFFFFFFFFu || dsq < best_dsq - (real)1e-10f || (same_d && better_tie)) {
                    if (within_two_hops(m, index, j) || within_two_hops(m, index, v0) || within_two_hops(m, index, v1))
                        continue;
                    best_dsq = dsq; best_h = h; bj = j; bv0 = v0; bv1 = v1;
                }
            }
        }
        if (bj == 0xFFFFFFFFu || !(best_h < 0)) continue;
        v3 p_j = s->p[bj], p_v0 = s->p[bv0], p_v1 = s->p[bv1];
        cbary r = closest_bary(p_i, p_j, p_v0, p_v1);
        v3 d = sub(p_i, r.pt);
        real dist = len3(d);
        v3 tcr = cross(sub(p_v0, p_j), sub(p_v1, p_j));
        real tl = len3(tcr);
        if (!(tl > EPS && dist > EPS)) continue;
        v3 tn = dvs(tcr, tl);
        real h = dot(d, tn);
        if (!(h < 0)) continue;
        v3 ni = s->normals[index];
        real nil = len3(ni);
        int tri_pin = s->w[bj] <= 0 || s->w[bv0] <= 0 || s->w[bv1] <= 0;
        int allow = tri_pin || nil <= (real)0.5f || dot(dvs(ni, rmax(nil, EPS)), tn) > 0;
        real eff = rmax(thick_i + (real)m->thick[bj], rmin(llen_i, (real)m->local_len[bj]) * (real)0.06f);
        real cos_behind = -h / dist;
        real r_tan = RSQRT(rmax(dist * dist - h * h, 0));
        real r_lim = rmax(eff * (real)0.5f, EPS);
        if (!(allow && cos_behind > (real)0.707f && r_tan < r_lim)) continue;
        v3 bc = r.bary;
        v3 nint = add(add(mul(s->normals[bj], bc.x), mul(s->normals[bv0], bc.y)), mul(s->normals[bv1], bc.z));
        real nl = len3(nint);
        if (!(nl <= (real)0.5f || dot(d, nint) < 0)) continue;
        real cone_w = rclamp((cos_behind - (real)0.707f) / (real)0.2f, 0, 1);
        real r_w = rclamp(1 - r_tan / r_lim, 0, 1);
        real pen = rmin(eff - h, eff * (real)2.0f) * cone_w * r_w;
        if (!(pen > EPS)) continue;
        if (tri_pin) { add_disp(&s->acc[index], mul(tn, pen)); continue; }
        real wi = s->w[index], wj = s->w[bj], w0 = s->w[bv0], w1 = s->w[bv1];
        real wt = bc.x * bc.x * wj + bc.y * bc.y * w0 + bc.z * bc.z * w1;
        real wtot = wi + wt;
        if (!(wtot > EPS)) continue;
        v3 disp = mul(tn, pen * (real)0.5f);
        add_disp(&s->acc[index], mul(disp, wi / wtot));
        if (wj > 0) add_disp(&s->acc[bj], mul(neg(disp), bc.x * wj / wtot));
        if (w0 > 0) add_disp(&s->acc[bv0], mul(neg(disp), bc.y * w0 / wtot));
        if (w1 > 0) add_disp(&s->acc[bv1], mul(neg(disp), bc.z * w1 / wtot));
    }
}

/* the vertex normals compute_normals.wgsl gives at the current positions x (nv x 3) */
void orc_sim_normals(orc_sim *s, float *out) {
    k_compute_normals(s);
    for (uint32_t i = 0; i < s->nv; i++) {
        out[3 * i] = (float)s->normals[i].x; out[3 * i + 1] = (float)s->normals[i].y; out[3 * i + 2] = (float)s->normals[i].z;
    }
}

static void self_collision_pass(orc_sim *s) {
    const orc_mesh *m = s->m;
    if (untangle_on(s)) k_compute_normals(s);
    hash_build(s->p, s->nv, s->cell_size, s->table_size, s->hkeys, s->hstarts, s->hsorted);
    if (m->ne > 0) {
        v3 *mid = (v3 *)malloc(m->ne * sizeof(v3));
        for (uint32_t e = 0; e < m->ne; e++) mid[e] = mul(add(s->p[m->e_v0[e]], s->p[m->e_v1[e]]), (real)0.5f);
        hash_build(mid, m->ne, s->cell_size, s->table_size, s->ekeys, s->estarts, s->esorted);
        free(mid);
    }
    k_sc_solve(s);
    k_sc_untangle(s);
    if (m->ne > 0) k_sc_solve_ee(s);
    k_sc_apply(s);
}

/* one frame: dispatch.rs:7-354 with adaptive substeps off, decoupled colliders,
 * self-collision mode 0 / algorithm 0, no sewing; the pins run at the start of every
 * iteration and once after the collider (dispatch.rs:176-181, 238-250) */
/* edge_collision.wgsl:98-150 */
typedef struct { real s, t, dist; } eseg;
static eseg edge_closest(v3 p0, v3 p1, v3 q0, v3 q1) {
    v3 d1 = sub(p1, p0), d2 = sub(q1, q0), r = sub(p0, q0);
    real a = dot(d1, d1), e = dot(d2, d2), f = dot(d2, r), s_ = 0, t = 0;
    eseg res;
    if (a <= (real)1e-12f && e <= (real)1e-12f) { res.s = 0; res.t = 0; res.dist = len3(sub(p0, q0)); return res; }
    if (a <= (real)1e-12f) {
        t = rclamp(f / e, 0, 1);
    } else {
        real c = dot(d1, r);
        if (e <= (real)1e-12f) {
            s_ = rclamp(-c / a, 0, 1);
        } else {
            real b = dot(d1, d2), denom = a * e - b * b;
            s_ = denom > a * e * (real)1e-7f ? rclamp((b * f - c * e) / denom, 0, 1) : 0;
            t = (b * s_ + f) / e;
            if (t < 0) { t = 0; s_ = rclamp(-c / a, 0, 1); }
            else if (t > 1) { t = 1; s_ = rclamp((b - c) / a, 0, 1); }
        }
    }
    res.s = s_; res.t = t; res.dist = len3(sub(add(p0, mul(d1, s_)), add(q0, mul(d2, t))));
    return res;
}
/* edge_collision.wgsl:153-276: every cloth edge over the 27 cells (z outer) of its midpoint */
static void k_edge_collision(orc_sim *s) {
    const orc_mesh *m = s->m;
    const real EPS = (real)1e-7f;
    const real cell = (real)m->cell_size;
    if (s->nce == 0) return;
    for (uint32_t k = 0; k < m->ne; k++) {
        uint32_t u0 = m->e_v0[k], u1 = m->e_v1[k];
        real w0 = s->w[u0], w1 = s->w[u1];
        if (w0 + w1 <= EPS) continue;
        v3 p0 = s->p[u0], p1 = s->p[u1];
        v3 mid = mul(add(p0, p1), (real)0.5f);
        real cth = rmax((real)m->thick[u0], (real)m->thick[u1]);
        int32_t cc[3]; cell_of(mid, cell, cc);
        v3 amin = vmin(p0, p1), amax = vmax(p0, p1);
        for (int dz = -1; dz <= 1; dz++)
        for (int dy = -1; dy <= 1; dy++)
        for (int dx = -1; dx <= 1; dx++) {
            int32_t nc[3] = {(int32_t)((uint32_t)cc[0] + (uint32_t)dx), (int32_t)((uint32_t)cc[1] + (uint32_t)dy),
                             (int32_t)((uint32_t)cc[2] + (uint32_t)dz)};
            uint32_t h = hash_coords(nc, 32768);
            for (uint32_t i = s->cstarts[h]; i < s->cstarts[h + 1]; i++) {
                uint32_t ci = s->csorted[i];
                const float *E = s->cedges + 11 * (size_t)ci;
                v3 q0 = mk(E[0], E[1], E[2]), q1 = mk(E[4], E[5], E[6]);
                real target = (cth + (real)E[3]) * s->edge_scale + s->edge_offset;
                v3 cmin = vmin(q0, q1), cmax = vmax(q0, q1);
                if (amax.x + target < cmin.x || amin.x - target > cmax.x || amax.y + target < cmin.y ||
                    amin.y - target > cmax.y || amax.z + target < cmin.z || amin.z - target > cmax.z)
                    continue;
                eseg r = edge_closest(p0, p1, q0, q1);
                if (!(r.dist < target)) continue;
                v3 diff = sub(add(p0, mul(sub(p1, p0), r.s)), add(q0, mul(sub(q1, q0), r.t)));
                v3 en = mk(E[8], E[9], E[10]);
                int single = (s->cedge_flags[ci] & 1u) != 0, has_n = dot(en, en) > (real)0.25f;
                v3 n = mk(0, 0, 1);
                real pen = target - r.dist;
                if (single && has_n) {
                    real sp = dot(diff, en);
                    if (sp >= 0) { n = r.dist > EPS ? dvs(diff, r.dist) : en; pen = target - r.dist; }
                    else { n = en; pen = rmin(target - sp, target * (real)2.0f); }
                } else if (r.dist > EPS) {
                    n = dvs(diff, r.dist);
                } else {
                    v3 cr = cross(sub(p1, p0), sub(q1, q0));
                    real cl = len3(cr);
                    if (cl > EPS) n = dvs(cr, cl);
                }
                v3 push = mul(n, pen);
                real a = 1 - r.s, b = r.s, denom = a * a * w0 + b * b * w1;
                if (denom > EPS) {
                    real f0 = rmin((a * w0) / denom, 2), f1 = rmin((b * w1) / denom, 2);
                    if (w0 > 0) add_pin_disp(&s->acc[u0], mul(push, f0));
                    if (w1 > 0) add_pin_disp(&s->acc[u1], mul(push, f1));
                }
            }
        }
    }
}

/* The outer self-collision pass of a substep (dispatch.rs:1554-1645): the pair cache collects on
 * the frame's first substep and solves the held pairs on the others; with the final fallback the
 * last substep takes the direct path. */
static void self_collision_outer(orc_sim *s, int first, int last) {
    if (!s->pair_cache || (s->pair_final_fallback && last)) { self_collision_pass(s); return; }
    if (untangle_on(s)) k_compute_normals(s);
    if (first) {
        hash_build(s->p, s->nv, s->cell_size, s->table_size, s->hkeys, s->hstarts, s->hsorted);
        k_pair_collect(s);
    }
    k_pair_solve_vt(s);
    k_pair_solve_ee(s);
    k_sc_untangle(s);
    k_sc_apply(s);
}

static void encode_steps(orc_sim *s, real dt, uint32_t substeps) {
    if (s->nv == 0) return;
    real dt_sub = dt / (real)substeps;
    const orc_mesh *m = s->m;
    s->pair_dt_frame = dt;
    for (uint32_t sub = 0; sub < substeps; sub++) {
        memset(s->lam_d, 0, (m->nd ? m->nd : 1) * sizeof(real));
        memset(s->lam_b, 0, (m->nb ? m->nb : 1) * sizeof(real));
        k_sew_shrink(s, dt_sub);
        k_predict(s, dt_sub);
        k_wrinkle(s, dt_sub);
        for (uint32_t it = 0; it < s->iterations; it++) {
            k_pins(s, dt_sub);
            distance_pass(s, dt_sub);
            sewing_pass(s, dt_sub);
            for (uint32_t c = 0; c < m->nb_col; c++) {
                ORC_PARALLEL_FOR
                for (uint32_t k = m->b_col_off[c]; k < m->b_col_off[c] + m->b_col_cnt[c]; k++) k_bending_one(s, k, dt_sub);
            }
        }
        if (s->ncols > 0 || s->ntris > 0 || s->sdf_nb > 0) k_collision(s, dt_sub);
        k_pins(s, dt_sub);
        if (s->edge_collision && m->ne > 0 && (s->ncols > 0 || s->ntris > 0 || s->sdf_nb > 0)) {
            k_edge_collision(s);
            k_sc_apply(s);
        }
        if (s->self_collision) self_collision_outer(s, sub == 0, sub + 1 == substeps);
        sewing_pass(s, dt_sub);
        k_update_vel(s, dt_sub);
    }
}

/* ClothSimulator.step (lib.rs:1770-1786 -> dispatch.rs:610-671); returns effective substeps */
uint32_t orc_sim_step(orc_sim *s, float dt, uint32_t substeps) {
    uint32_t S = substeps > 0 ? substeps : 20;
    dynamic_sdf_update(s);
    encode_steps(s, (real)dt, S);
    return S;
}
void orc_sim_step_single_substep(orc_sim *s, float dt_sub) { encode_steps(s, (real)dt_sub, 1); }

/* Pieces for tests: run only the parts of one substep up to (and excluding) the
 * velocity update, so that the constraint residual can be evaluated. */
void orc_sim_substep_parts(orc_sim *s, float dt_sub, int do_collide, int do_self, int do_update) {
    const orc_mesh *m = s->m;
    real dt = (real)dt_sub;
    memset(s->lam_d, 0, (m->nd ? m->nd : 1) * sizeof(real));
    memset(s->lam_b, 0, (m->nb ? m->nb : 1) * sizeof(real));
    k_sew_shrink(s, dt);
    k_predict(s, dt);
    k_wrinkle(s, dt);
    for (uint32_t it = 0; it < s->iterations; it++) {
        k_pins(s, dt);
        distance_pass(s, dt);
        sewing_pass(s, dt);
        for (uint32_t c = 0; c < m->nb_col; c++) {
            ORC_PARALLEL_FOR
            for (uint32_t k = m->b_col_off[c]; k < m->b_col_off[c] + m->b_col_cnt[c]; k++) k_bending_one(s, k, dt);
        }
    }
    if (do_collide && (s->ncols > 0 || s->ntris > 0 || s->sdf_nb > 0)) k_collision(s, dt);
    k_pins(s, dt);
    if (do_collide && s->edge_collision && m->ne > 0 && (s->ncols > 0 || s->ntris > 0 || s->sdf_nb > 0)) {
        k_edge_collision(s);
        k_sc_apply(s);
    }
    s->pair_dt_frame = dt;
    if (do_self && s->self_collision) self_collision_outer(s, 1, 1);
    sewing_pass(s, dt);
    if (do_update) k_update_vel(s, dt);
}

/* Relative XPBD residual r = ||h|| / ||L0|| over distance, shear and bending constraints,
 * h_j = C_j(p) + alpha~_j lambda_j + gamma_j grad C_j . (p - x): the quantity the damped
 * projection of distance.wgsl:60-124 / bending.wgsl:59-126 drives to zero (its numerator),
 * with alpha~ = alpha / dt^2 and gamma = min(damping * 1e-5 / dt, 1e4) (zero without
 * constraint damping).  Constraints the projection skips (w0 + w1 <= 1e-7, |p0 - p1| <= 1e-7,
 * compliance >= 1e9) are excluded.  Also returns max |h_j| / L0_j.  No reference counterpart. */
void orc_sim_residual(const orc_sim *s, float dt_sub, double *out_rel, double *out_maxrel) {
    const orc_mesh *m = s->m;
    double dt = dt_sub > 1e-8f ? dt_sub : 1e-8;
    double hh = 0, ll = 0, mx = 0;
    for (uint32_t k = 0; k < m->nd; k++) {
        uint32_t a = m->d_v0[k], b = m->d_v1[k];
        double ws = (double)s->w[a] + (double)s->w[b];
        double L = m->d_rest[k];
        if (ws <= 1e-7) continue;
        double dx = (double)s->p[a].x - s->p[b].x, dy = (double)s->p[a].y - s->p[b].y, dz = (double)s->p[a].z - s->p[b].z;
        double d = sqrt(dx * dx + dy * dy + dz * dz);
        if (d <= 1e-7) continue;
        double c = d - L;
        double comp = c >= 0 ? (double)s->d_tc[k] : (double)s->d_cc[k];
        if (comp >= 1e9) continue;
        double h = c + comp / (dt * dt) * (double)s->lam_d[k];
        double damp = m->d_type[k] == 1 ? (double)s->shear_damping
                                        : (c >= 0 ? (double)s->tension_damping : (double)s->compression_damping);
        if (damp > 0) {
            double gamma = fmin(damp * (double)1e-5f / dt, 1e4);
            double rx = ((double)s->p[a].x - s->x[a].x) - ((double)s->p[b].x - s->x[b].x);
            double ry = ((double)s->p[a].y - s->x[a].y) - ((double)s->p[b].y - s->x[b].y);
            double rz = ((double)s->p[a].z - s->x[a].z) - ((double)s->p[b].z - s->x[b].z);
            h += gamma * (dx * rx + dy * ry + dz * rz) / d;
        }
        hh += h * h; ll += L * L;
        if (L > 0 && fabs(h) / L > mx) mx = fabs(h) / L;
    }
    for (uint32_t k = 0; k < m->nb; k++) {
        double comp = s->b_comp[k];
        if (comp >= 1e9) continue;
        uint32_t a = m->b_v2[k], b = m->b_v3[k];
        double ws = (double)s->w[a] + (double)s->w[b];
        if (ws <= 1e-7) continue;
        double L = m->b_rest[k];
        double dx = (double)s->p[a].x - s->p[b].x, dy = (double)s->p[a].y - s->p[b].y, dz = (double)s->p[a].z - s->p[b].z;
        double d = sqrt(dx * dx + dy * dy + dz * dz);
        if (d <= 1e-7) continue;
        double h = (d - L) + comp / (dt * dt) * (double)s->lam_b[k];
        if (s->bending_damping > 0) {
            double gamma = fmin((double)s->bending_damping * (double)1e-5f / dt, 1e4);
            double rx = ((double)s->p[a].x - s->x[a].x) - ((double)s->p[b].x - s->x[b].x);
            double ry = ((double)s->p[a].y - s->x[a].y) - ((double)s->p[b].y - s->x[b].y);
            double rz = ((double)s->p[a].z - s->x[a].z) - ((double)s->p[b].z - s->x[b].z);
            h += gamma * (dx * rx + dy * ry + dz * rz) / d;
        }
        hh += h * h; ll += L * L;
        if (L > 0 && fabs(h) / L > mx) mx = fabs(h) / L;
    }
    *out_rel = ll > 0 ? sqrt(hh) / sqrt(ll) : 0;
    *out_maxrel = mx;
}

int orc_real_bytes(void) { return (int)sizeof(real); }
#ifdef _OPENMP
#include <omp.h>
int orc_num_threads(void) { return omp_get_max_threads(); }
#else
int orc_num_threads(void) { return 1; }
#endif
void orc_sim_stats(orc_sim *s, uint64_t *out, int clear) {
    for (int i = 0; i < 6; i++) { out[i] = s->stats[i]; if (clear) s->stats[i] = 0; }
}
