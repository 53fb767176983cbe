/*
 * CPU ORACLE for the mesh SDF bake and the SDF collider -- TEST INFRASTRUCTURE ONLY.
 *
 * Plain scalar C restatement (compiled with xpbd_oracle.c into liborc_f32.so / liborc_f64.so;
 * ORC_REAL sets the type of the per-voxel distance arithmetic).  Only tests/ load it, through
 * oracle/oracle.py.  Reference functions restated (paths relative to the reference root):
 *   close_open_boundary_loops ...... crates/cloth_core/src/sdf_baker.rs:244-462
 *   mesh_sdf_domain / bone info .... crates/cloth_core/src/sdf_baker.rs:939-1013
 *   collect_mesh_bake_triangles .... crates/cloth_core/src/sdf_baker.rs:1016-1065
 *   dense bake ..................... crates/cloth_core/src/shaders/bake_mesh_sdf.wgsl:80-121
 *   coarse top-8 pass .............. crates/cloth_core/src/shaders/bake_mesh_sdf_hier_coarse_main.wgsl
 *   fine pass ...................... crates/cloth_core/src/shaders/bake_mesh_sdf_hier_fine_main.wgsl
 *   hierarchical host .............. crates/cloth_core/src/sdf_baker.rs:1625-1882
 *   SDF collider ................... crates/cloth_core/src/shaders/collision.wgsl:116-265
 * The dense bake is the brute-force scan in triangle order (first strict minimum wins), with no
 * culling.  pack2x16float is taken as round-to-nearest-even to binary16, the rounding of the
 * hardware conversion; the reference's driver rounding of that pack is not observable here.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifndef ORC_REAL
#define ORC_REAL float
#endif
typedef ORC_REAL real;

typedef struct { real x, y, z; } s3;
/* With -fopenmp (liborc_f32_omp.so) the voxel slabs of a bake are split over threads; every voxel
 * is computed alone, so the texels are the serial ones bit for bit. */
#ifdef _OPENMP
#define ORC_PARALLEL_FOR _Pragma("omp parallel for schedule(dynamic, 1)")
#else
#define ORC_PARALLEL_FOR
#endif

static inline s3 sk(real x, real y, real z) { s3 r = {x, y, z}; return r; }
static inline s3 sadd(s3 a, s3 b) { return sk(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline s3 ssub(s3 a, s3 b) { return sk(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline s3 smul(s3 a, real s) { return sk(a.x * s, a.y * s, a.z * s); }
static inline real sdot(s3 a, s3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline real ssqrt(real x) { return (real)sqrt((double)x); }

/* ---- binary16 ---- */
uint16_t orc_f32_to_f16(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    const uint32_t sign = (u >> 16) & 0x8000u;
    const uint32_t ex = (u >> 23) & 0xFFu;
    uint32_t man = u & 0x7FFFFFu;
    if (ex == 0xFFu) return (uint16_t)(sign | 0x7C00u | (man ? 0x200u : 0u));
    int e = (int)ex - 127 + 15;
    if (e >= 31) return (uint16_t)(sign | 0x7C00u);
    if (e <= 0) {
        if (e < -10) return (uint16_t)sign;
        man |= 0x800000u;
        const int shift = 14 - e;
        uint32_t h = man >> shift;
        const uint32_t rem = man & ((1u << shift) - 1u), half = 1u << (shift - 1);
        if (rem > half || (rem == half && (h & 1u))) h++;
        return (uint16_t)(sign | h);
    }
    uint32_t h = ((uint32_t)e << 10) | (man >> 13);
    const uint32_t rem = man & 0x1FFFu;
    if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) h++;  /* a carry into the exponent is the right result */
    return (uint16_t)(sign | h);
}

float orc_f16_to_f32(uint16_t h) {
    const uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
    uint32_t ex = (h >> 10) & 0x1Fu, man = h & 0x3FFu, u;
    if (ex == 0) {
        if (man == 0) {
            u = sign;
        } else {
            ex = 127 - 15 + 1;
            while (!(man & 0x400u)) { man <<= 1; ex--; }
            man &= 0x3FFu;
            u = sign | (ex << 23) | (man << 13);
        }
    } else if (ex == 31) {
        u = sign | 0x7F800000u | (man << 13);
    } else {
        u = sign | ((ex - 15 + 127) << 23) | (man << 13);
    }
    float f;
    memcpy(&f, &u, 4);
    return f;
}

static inline uint32_t pack_dist(real d) { return (uint32_t)orc_f32_to_f16((float)d) | 0x3C000000u; }

/* ---- host preparation (always f32, as the reference's host code) ---- */
static void cross3(const float *a, const float *b, float *o) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}
static void normalize3(const float *v, float *o) {
    const float ls = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
    if (ls > 1e-12f) {
        const float inv = 1.0f / (float)sqrt((double)ls);
        o[0] = v[0] * inv; o[1] = v[1] * inv; o[2] = v[2] * inv;
    } else {
        o[0] = 0.0f; o[1] = 1.0f; o[2] = 0.0f;
    }
}
static int tri_ok(const int32_t *t, uint32_t n) {
    return t[0] >= 0 && t[1] >= 0 && t[2] >= 0 && (uint32_t)t[0] < n && (uint32_t)t[1] < n && (uint32_t)t[2] < n &&
           t[0] != t[1] && t[1] != t[2] && t[2] != t[0];
}

typedef struct { int32_t a, b; uint32_t count; } ecount;
static int cmp_ecount(const void *x, const void *y) {
    const ecount *p = (const ecount *)x, *q = (const ecount *)y;
    if (p->a != q->a) return p->a < q->a ? -1 : 1;
    if (p->b != q->b) return p->b < q->b ? -1 : 1;
    return 0;
}
typedef struct { int32_t v[3]; } tri3;
static int cmp_tri3(const void *x, const void *y) {
    const tri3 *p = (const tri3 *)x, *q = (const tri3 *)y;
    for (int k = 0; k < 3; k++)
        if (p->v[k] != q->v[k]) return p->v[k] < q->v[k] ? -1 : 1;
    return 0;
}
static void sort3(int32_t *s) {
    int32_t t;
    if (s[0] > s[1]) { t = s[0]; s[0] = s[1]; s[1] = t; }
    if (s[1] > s[2]) { t = s[1]; s[1] = s[2]; s[2] = t; }
    if (s[0] > s[1]) { t = s[0]; s[0] = s[1]; s[1] = t; }
}
static uint32_t edge_count(const ecount *ec, size_t n, int32_t u, int32_t v) {
    ecount key = {u < v ? u : v, u < v ? v : u, 0};
    const ecount *f = (const ecount *)bsearch(&key, ec, n, sizeof(ecount), cmp_ecount);
    return f ? f->count : 0;
}

/* Caps closed boundary loops.  verts (3 x nv floats) and tris (3 x nt) are grown in place through
 * the returned arrays; *nv_io / *nt_io are updated.  With bone weights (wb[b] holds nv floats for
 * each of nbones bones, grown in place) each cap centre takes the mean weight of its loop
 * (sdf_baker.rs:427-439).  Returns the loops closed. */
static uint32_t close_loops_weighted(float **verts_io, uint32_t *nv_io, int32_t **tris_io, uint32_t *nt_io,
                                     float **wb, uint32_t nbones) {
    const uint32_t n = *nv_io, m = *nt_io;
    if (m < 4 || n < 4) return 0;
    float *V = *verts_io;
    int32_t *T = *tris_io;
    /* edge multiplicities and the set of existing triangles */
    ecount *ec = (ecount *)malloc(sizeof(ecount) * 3 * m);
    tri3 *ex = (tri3 *)malloc(sizeof(tri3) * (m + 1));
    size_t nec = 0, nex = 0;
    for (uint32_t t = 0; t < m; t++) {
        const int32_t *tr = &T[3 * t];
        if (!tri_ok(tr, n)) continue;
        tri3 s = {{tr[0], tr[1], tr[2]}};
        sort3(s.v);
        ex[nex++] = s;
        for (int e = 0; e < 3; e++) {
            const int32_t u = tr[e], v = tr[(e + 1) % 3];
            ecount c = {u < v ? u : v, u < v ? v : u, 1};
            ec[nec++] = c;
        }
    }
    qsort(ec, nec, sizeof(ecount), cmp_ecount);
    size_t w = 0;
    for (size_t i = 0; i < nec; i++) {
        if (w && ec[w - 1].a == ec[i].a && ec[w - 1].b == ec[i].b) ec[w - 1].count++;
        else ec[w++] = ec[i];
    }
    nec = w;
    qsort(ex, nex, sizeof(tri3), cmp_tri3);
    /* directed boundary edges in triangle order with the face normal of their triangle */
    int32_t *bu = (int32_t *)malloc(sizeof(int32_t) * 3 * m), *bv = (int32_t *)malloc(sizeof(int32_t) * 3 * m);
    float *bn = (float *)malloc(sizeof(float) * 9 * m);
    size_t nb = 0;
    for (uint32_t t = 0; t < m; t++) {
        const int32_t *tr = &T[3 * t];
        if (!tri_ok(tr, n)) continue;
        const float *p0 = &V[3 * (size_t)tr[0]], *p1 = &V[3 * (size_t)tr[1]], *p2 = &V[3 * (size_t)tr[2]];
        const float e01[3] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
        const float e02[3] = {p2[0] - p0[0], p2[1] - p0[1], p2[2] - p0[2]};
        float c[3], nrm[3];
        cross3(e01, e02, c);
        normalize3(c, nrm);
        for (int e = 0; e < 3; e++) {
            const int32_t u = tr[e], v = tr[(e + 1) % 3];
            if (edge_count(ec, nec, u, v) == 1u) {
                bu[nb] = u; bv[nb] = v;
                memcpy(&bn[3 * nb], nrm, sizeof(nrm));
                nb++;
            }
        }
    }
    uint32_t closed = 0;
    if (nb) {
        char *used = (char *)calloc(nb, 1);
        int32_t *loop = (int32_t *)malloc(sizeof(int32_t) * (n + 2));
        size_t *loop_e = (size_t *)malloc(sizeof(size_t) * (nb + 1));
        for (size_t start = 0; start < nb; start++) {
            if (used[start]) continue;
            used[start] = 1;
            const int32_t su = bu[start];
            size_t k = 0, ke = 0;
            loop[k++] = su;
            loop[k++] = bv[start];
            loop_e[ke++] = start;
            int32_t curr = bv[start];
            int is_closed = 0;
            while (k <= n) {
                if (curr == su) { k--; is_closed = 1; break; }
                int found = 0;
                for (size_t e = 0; e < nb; e++) /* candidates of `curr` in insertion order */
                    if (bu[e] == curr && !used[e]) {
                        used[e] = 1;
                        loop_e[ke++] = e;
                        curr = bv[e];
                        found = 1;
                        break;
                    }
                if (!found) break;
                loop[k++] = curr;
            }
            if (!is_closed || k < 3 || k >= n) continue;
            const float inv_k = 1.0f / (float)k;
            float center[3] = {0.0f, 0.0f, 0.0f};
            for (size_t i = 0; i < k; i++)
                for (int a = 0; a < 3; a++) center[a] += V[3 * (size_t)loop[i] + a];
            for (int a = 0; a < 3; a++) center[a] *= inv_k;
            float cap[3] = {0.0f, 0.0f, 0.0f};
            for (size_t i = 0; i < k; i++) {
                const float *pu = &V[3 * (size_t)loop[i]], *pv = &V[3 * (size_t)loop[(i + 1) % k]];
                const float d1[3] = {pu[0] - pv[0], pu[1] - pv[1], pu[2] - pv[2]};
                const float d2[3] = {center[0] - pv[0], center[1] - pv[1], center[2] - pv[2]};
                float cn[3];
                cross3(d1, d2, cn);
                for (int a = 0; a < 3; a++) cap[a] += cn[a];
            }
            float capn[3];
            normalize3(cap, capn);
            float dsum = 0.0f;
            for (size_t i = 0; i < ke; i++) {
                const float *an = &bn[3 * loop_e[i]];
                dsum += an[0] * capn[0] + an[1] * capn[1] + an[2] * capn[2];
            }
            const float mean = dsum / (float)(ke ? ke : 1);
            if (mean <= -0.85f) continue;
            if (k == 3) {
                tri3 s = {{loop[0], loop[1], loop[2]}};
                sort3(s.v);
                /* linear search: the set also grows with the caps added so far */
                int have = 0;
                for (size_t i = 0; i < nex; i++)
                    if (!cmp_tri3(&ex[i], &s)) { have = 1; break; }
                if (have) continue;
                ex = (tri3 *)realloc(ex, sizeof(tri3) * (nex + 1));
                ex[nex++] = s;
                T = (int32_t *)realloc(T, sizeof(int32_t) * 3 * (*nt_io + 1));
                T[3 * *nt_io] = loop[0]; T[3 * *nt_io + 1] = loop[2]; T[3 * *nt_io + 2] = loop[1];
                (*nt_io)++;
                closed++;
                continue;
            }
            const int32_t ci = (int32_t)*nv_io;
            V = (float *)realloc(V, sizeof(float) * 3 * (*nv_io + 1));
            memcpy(&V[3 * (size_t)ci], center, sizeof(center));
            for (uint32_t b = 0; b < nbones; b++) {
                float ws = 0.0f;
                for (size_t i = 0; i < k; i++) ws += wb[b][loop[i]];
                wb[b] = (float *)realloc(wb[b], sizeof(float) * (*nv_io + 1));
                wb[b][ci] = ws * inv_k;
            }
            (*nv_io)++;
            T = (int32_t *)realloc(T, sizeof(int32_t) * 3 * (*nt_io + k));
            for (size_t i = 0; i < k; i++) {
                T[3 * *nt_io] = loop[(i + 1) % k]; T[3 * *nt_io + 1] = loop[i]; T[3 * *nt_io + 2] = ci;
                (*nt_io)++;
            }
            closed++;
        }
        free(used); free(loop); free(loop_e);
    }
    free(ec); free(ex); free(bu); free(bv); free(bn);
    *verts_io = V;
    *tris_io = T;
    return closed;
}

uint32_t orc_close_open_boundary_loops(float **verts_io, uint32_t *nv_io, int32_t **tris_io, uint32_t *nt_io) {
    return close_loops_weighted(verts_io, nv_io, tris_io, nt_io, NULL, 0);
}

typedef struct { float min[3], max[3], size[3]; uint32_t w, h, d; } odomain;

static odomain sdf_domain(const float *V, uint32_t nv, float voxel, float margin, float thickness) {
    odomain dm;
    for (int a = 0; a < 3; a++) { dm.min[a] = INFINITY; dm.max[a] = -INFINITY; }
    for (uint32_t i = 0; i < nv; i++)
        for (int a = 0; a < 3; a++) {
            dm.min[a] = fminf(dm.min[a], V[3 * i + a]);
            dm.max[a] = fmaxf(dm.max[a], V[3 * i + a]);
        }
    for (int a = 0; a < 3; a++) dm.size[a] = fmaxf(dm.max[a] - dm.min[a], 0.05f);
    for (int a = 0; a < 3; a++) {
        const float pad = fmaxf(dm.size[a] * margin, 0.04f) + fmaxf(thickness, 0.0f) + 0.02f;
        dm.min[a] -= pad;
        dm.max[a] += pad;
        dm.size[a] = dm.max[a] - dm.min[a];
    }
    const float v = fmaxf(voxel, 0.001f);
    uint32_t *dims[3] = {&dm.w, &dm.h, &dm.d};
    for (int a = 0; a < 3; a++) {
        double r = ceil((double)(dm.size[a] / v));
        if (r < 8.0) r = 8.0;
        if (r > 2048.0) r = 2048.0;
        *dims[a] = (uint32_t)r;
    }
    return dm;
}

typedef struct { real p0[3], p1[3], p2[3], n[3]; } otri;

static otri *bake_tris(const float *V, uint32_t nv, const int32_t *T, uint32_t nt, uint32_t *count) {
    otri *out = (otri *)malloc(sizeof(otri) * (nt ? nt : 1));
    uint32_t c = 0;
    for (uint32_t t = 0; t < nt; t++) {
        const int32_t *tr = &T[3 * t];
        if (tr[0] < 0 || tr[1] < 0 || tr[2] < 0 || (uint32_t)tr[0] >= nv || (uint32_t)tr[1] >= nv || (uint32_t)tr[2] >= nv)
            continue;
        const float *p0 = &V[3 * (size_t)tr[0]], *p1 = &V[3 * (size_t)tr[1]], *p2 = &V[3 * (size_t)tr[2]];
        const float e01[3] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
        const float e02[3] = {p2[0] - p0[0], p2[1] - p0[1], p2[2] - p0[2]};
        float cr[3], nrm[3];
        cross3(e01, e02, cr);
        normalize3(cr, nrm);
        for (int a = 0; a < 3; a++) {
            out[c].p0[a] = p0[a]; out[c].p1[a] = p1[a]; out[c].p2[a] = p2[a]; out[c].n[a] = nrm[a];
        }
        c++;
    }
    if (!c) {
        memset(&out[0], 0, sizeof(otri));
        out[0].n[1] = 1;
        c = 1;
    }
    *count = c;
    return out;
}

/* Ericson, Real-Time Collision Detection, 5.1.5 */
static s3 closest_tri(s3 p, s3 a, s3 b, s3 c) {
    const s3 ab = ssub(b, a), ac = ssub(c, a), ap = ssub(p, a);
    const real d1 = sdot(ab, ap), d2 = sdot(ac, ap);
    if (d1 <= 0 && d2 <= 0) return a;
    const s3 bp = ssub(p, b);
    const real d3 = sdot(ab, bp), d4 = sdot(ac, bp);
    if (d3 >= 0 && d4 <= d3) return b;
    const real vc = d1 * d4 - d3 * d2;
    if (vc <= 0 && d1 >= 0 && d3 <= 0) return sadd(a, smul(ab, d1 / (d1 - d3)));
    const s3 cp = ssub(p, c);
    const real d5 = sdot(ab, cp), d6 = sdot(ac, cp);
    if (d6 >= 0 && d5 <= d6) return c;
    const real vb = d5 * d2 - d1 * d6;
    if (vb <= 0 && d2 >= 0 && d6 <= 0) return sadd(a, smul(ac, d2 / (d2 - d6)));
    const real va = d3 * d6 - d5 * d4;
    if (va <= 0 && (d4 - d3) >= 0 && (d5 - d6) >= 0) {
        const real w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        return sadd(b, smul(ssub(c, b), w));
    }
    const real denom = (real)1 / (va + vb + vc);
    const real v = vb * denom, w = vc * denom;
    return sadd(sadd(a, smul(ab, v)), smul(ac, w));
}

static inline s3 tp(const real *x) { return sk(x[0], x[1], x[2]); }

static s3 voxel_center(const odomain *dm, uint32_t x, uint32_t y, uint32_t z) {
    const real ux = ((real)x + (real)0.5) * ((real)1 / (real)dm->w);
    const real uy = ((real)y + (real)0.5) * ((real)1 / (real)dm->h);
    const real uz = ((real)z + (real)0.5) * ((real)1 / (real)dm->d);
    return sk((real)dm->min[0] + ux * ((real)dm->max[0] - (real)dm->min[0]),
              (real)dm->min[1] + uy * ((real)dm->max[1] - (real)dm->min[1]),
              (real)dm->min[2] + uz * ((real)dm->max[2] - (real)dm->min[2]));
}

/* one triangle: squared distance and signed distance */
static void tri_eval(s3 p, const otri *t, real *dsq, real *sg) {
    const s3 q = closest_tri(p, tp(t->p0), tp(t->p1), tp(t->p2));
    const s3 diff = ssub(p, q);
    *dsq = sdot(diff, diff);
    *sg = ssqrt(*dsq) * (sdot(diff, tp(t->n)) >= 0 ? (real)1 : (real)-1);
}

typedef struct {
    uint32_t w, h, d;
    float bone_info[20];
    float coarse_voxel;
    uint32_t *texels, *nearest;
} orc_sdf;

static void bone_info(const odomain *dm, float friction, float thickness, float restitution, float *o) {
    memset(o, 0, 20 * sizeof(float));
    for (int a = 0; a < 3; a++) {
        o[a] = dm->min[a];
        o[4 + a] = dm->max[a];
        o[8 + a] = 1.0f / dm->size[a];
        o[12 + a] = -dm->min[a] / dm->size[a];
    }
    o[7] = (float)dm->w;
    o[16] = friction; o[17] = thickness; o[18] = restitution;
}

/* bake_mesh_sdf.wgsl (K = 1) / the coarse pass (K = 8) over every voxel of dm */
static void scan(const odomain *dm, const otri *tr, uint32_t nt, int K, uint32_t *out, uint32_t *nearest) {
    ORC_PARALLEL_FOR
    for (uint32_t z = 0; z < dm->d; z++)
        for (uint32_t y = 0; y < dm->h; y++)
            for (uint32_t x = 0; x < dm->w; x++) {
                const s3 p = voxel_center(dm, x, y, z);
                real bsq[8], bsg[8];
                uint32_t bid[8];
                for (int k = 0; k < 8; k++) { bsq[k] = (real)1e10; bsg[k] = (real)10; bid[k] = 0; }
                for (uint32_t i = 0; i < nt; i++) {
                    const s3 q = closest_tri(p, tp(tr[i].p0), tp(tr[i].p1), tp(tr[i].p2));
                    const s3 diff = ssub(p, q);
                    const real dsq = sdot(diff, diff);
                    if (!(dsq < bsq[K - 1])) continue;
                    const real sg = ssqrt(dsq) * (sdot(diff, tp(tr[i].n)) >= 0 ? (real)1 : (real)-1);
                    real is = dsq, ig = sg;
                    uint32_t ii = i;
                    for (int k = 0; k < K; k++)
                        if (is < bsq[k]) {
                            const real ts = bsq[k], tg = bsg[k];
                            const uint32_t ti = bid[k];
                            bsq[k] = is; bsg[k] = ig; bid[k] = ii;
                            is = ts; ig = tg; ii = ti;
                        }
                }
                const size_t flat = ((size_t)z * dm->h + y) * dm->w + x;
                out[flat] = pack_dist(bsg[0]);
                if (nearest)
                    for (int k = 0; k < 8; k++) nearest[flat * 8 + k] = bid[k];
            }
}

/* bake_mesh_sdf_hier_fine_main.wgsl, voxel by voxel */
static void fine(const odomain *dm, const odomain *cd, const otri *tr, uint32_t nt, const uint32_t *cdist,
                 const uint32_t *ctri, real band, uint32_t *out) {
    const s3 lmin = sk(dm->min[0], dm->min[1], dm->min[2]), lmax = sk(dm->max[0], dm->max[1], dm->max[2]);
    const s3 size = ssub(lmax, lmin);
    const s3 spacing = sk(size.x / (real)cd->w, size.y / (real)cd->h, size.z / (real)cd->d);
    const real halfdiag = (real)0.5 * ssqrt(sdot(spacing, spacing));
    ORC_PARALLEL_FOR
    for (uint32_t z = 0; z < dm->d; z++)
        for (uint32_t y = 0; y < dm->h; y++)
            for (uint32_t x = 0; x < dm->w; x++) {
                const s3 uvw = sk(((real)x + (real)0.5) * ((real)1 / (real)dm->w),
                                  ((real)y + (real)0.5) * ((real)1 / (real)dm->h),
                                  ((real)z + (real)0.5) * ((real)1 / (real)dm->d));
                const s3 p = sk(lmin.x + uvw.x * size.x, lmin.y + uvw.y * size.y, lmin.z + uvw.z * size.z);
                const s3 rel = ssub(p, lmin);
                int32_t par[3];
                const real rr[3] = {rel.x / spacing.x, rel.y / spacing.y, rel.z / spacing.z};
                const uint32_t cdims[3] = {cd->w, cd->h, cd->d};
                for (int a = 0; a < 3; a++) {
                    const double f = floor((double)rr[a]);
                    int32_t v = f >= 2147483520.0 ? 2147483647 : f <= -2147483648.0 ? INT32_MIN : (int32_t)f;
                    if (v < 0) v = 0;
                    if (v > (int32_t)cdims[a] - 1) v = (int32_t)cdims[a] - 1;
                    par[a] = v;
                }
                const size_t pidx = ((size_t)par[2] * cd->h + (uint32_t)par[1]) * cd->w + (uint32_t)par[0];
                const uint32_t dcp = cdist[pidx];
                const real dc = (real)orc_f16_to_f32((uint16_t)(dcp & 0xFFFFu));
                const size_t flat = ((size_t)z * dm->h + y) * dm->w + x;
                if ((real)fabs((double)dc) - halfdiag >= band) {
                    out[flat] = dcp;
                    continue;
                }
                uint32_t best_i = 0xFFFFFFFFu;
                real best_sq = (real)1e10, best_sg = (real)10, second_sq = (real)1e10, second_sg = (real)10;
                for (int dz = -2; dz <= 2; dz++)
                    for (int dy = -2; dy <= 2; dy++)
                        for (int dx = -2; dx <= 2; dx++) {
                            int32_t c[3] = {par[0] + dx, par[1] + dy, par[2] + dz};
                            for (int a = 0; a < 3; a++) {
                                if (c[a] < 0) c[a] = 0;
                                if (c[a] > (int32_t)cdims[a] - 1) c[a] = (int32_t)cdims[a] - 1;
                            }
                            const size_t base = (((size_t)c[2] * cd->h + (uint32_t)c[1]) * cd->w + (uint32_t)c[0]) * 8;
                            for (int k = 0; k < 8; k++) {
                                const uint32_t ti = ctri[base + k];
                                if (ti >= nt || ti == best_i) continue;
                                const otri *t = &tr[ti];
                                s3 tmin = tp(t->p0), tmax = tp(t->p0);
                                const s3 pts[2] = {tp(t->p1), tp(t->p2)};
                                for (int j = 0; j < 2; j++) {
                                    tmin = sk(tmin.x < pts[j].x ? tmin.x : pts[j].x, tmin.y < pts[j].y ? tmin.y : pts[j].y,
                                              tmin.z < pts[j].z ? tmin.z : pts[j].z);
                                    tmax = sk(tmax.x > pts[j].x ? tmax.x : pts[j].x, tmax.y > pts[j].y ? tmax.y : pts[j].y,
                                              tmax.z > pts[j].z ? tmax.z : pts[j].z);
                                }
                                const s3 lo = ssub(tmin, p), hi = ssub(p, tmax);
                                const s3 dv = sk((lo.x > 0 ? lo.x : 0) + (hi.x > 0 ? hi.x : 0),
                                                 (lo.y > 0 ? lo.y : 0) + (hi.y > 0 ? hi.y : 0),
                                                 (lo.z > 0 ? lo.z : 0) + (hi.z > 0 ? hi.z : 0));
                                if (sdot(dv, dv) >= second_sq) continue;
                                real dsq, sg;
                                tri_eval(p, t, &dsq, &sg);
                                if (dsq < best_sq) {
                                    second_sq = best_sq; second_sg = best_sg;
                                    best_sq = dsq; best_sg = sg; best_i = ti;
                                } else if (dsq < second_sq) {
                                    second_sq = dsq; second_sg = sg;
                                }
                            }
                        }
                int amb = ssqrt(second_sq) - ssqrt(best_sq) <= (real)0.0002f;
                if (amb) amb = (best_sg >= 0) != (second_sg >= 0);
                if (amb) {
                    best_sq = (real)1e10;
                    best_sg = (real)10;
                    for (uint32_t i = 0; i < nt; i++) {
                        real dsq, sg;
                        tri_eval(p, &tr[i], &dsq, &sg);
                        if (dsq < best_sq) { best_sq = dsq; best_sg = sg; }
                    }
                }
                out[flat] = pack_dist(best_sg);
            }
}

/* method 0 dense, 1 hierarchical, 2 coarse pass alone; NULL on failure (empty mesh, band too wide) */
orc_sdf *orc_sdf_bake(const float *verts, uint32_t nv, const int32_t *tris, uint32_t nt, float voxel, float margin,
                      float friction, float thickness, float restitution, int close_open, int method) {
    if (!nv) return NULL;
    float *V = (float *)malloc(sizeof(float) * 3 * nv);
    int32_t *T = (int32_t *)malloc(sizeof(int32_t) * 3 * (nt ? nt : 1));
    memcpy(V, verts, sizeof(float) * 3 * nv);
    if (nt) memcpy(T, tris, sizeof(int32_t) * 3 * nt);
    uint32_t n = nv, m = nt;
    if (close_open) orc_close_open_boundary_loops(&V, &n, &T, &m);
    uint32_t ntr = 0;
    otri *tr = bake_tris(V, n, T, m, &ntr);
    orc_sdf *r = (orc_sdf *)calloc(1, sizeof(orc_sdf));
    const float cv = fmaxf(voxel * 4.0f, 0.001f);
    const odomain dm = sdf_domain(V, n, method == 2 ? cv : voxel, margin, thickness);
    bone_info(&dm, friction, thickness, restitution, r->bone_info);
    r->w = dm.w; r->h = dm.h; r->d = dm.d;
    const size_t nvox = (size_t)dm.w * dm.h * dm.d;
    r->texels = (uint32_t *)malloc(4 * nvox);
    if (method == 0) {
        scan(&dm, tr, ntr, 1, r->texels, NULL);
    } else if (method == 2) {
        r->coarse_voxel = cv;
        r->nearest = (uint32_t *)malloc(4 * nvox * 8);
        scan(&dm, tr, ntr, 8, r->texels, r->nearest);
    } else {
        const odomain cd = sdf_domain(V, n, cv, margin, thickness);
        const size_t nc = (size_t)cd.w * cd.h * cd.d;
        uint32_t *cdist = (uint32_t *)malloc(4 * nc), *ctri = (uint32_t *)malloc(4 * nc * 8);
        scan(&cd, tr, ntr, 8, cdist, ctri);
        const float h_fine = fmaxf(fmaxf(dm.size[0] / (float)dm.w, dm.size[1] / (float)dm.h), dm.size[2] / (float)dm.d);
        const float band = 4.0f * h_fine;
        const float hcm = fminf(fminf(cd.size[0] / (float)cd.w, cd.size[1] / (float)cd.h), cd.size[2] / (float)cd.d);
        if (band > 2.0f * hcm) {
            free(cdist); free(ctri); free(r->texels); free(r); r = NULL;
        } else {
            fine(&dm, &cd, tr, ntr, cdist, ctri, (real)band, r->texels);
            free(cdist); free(ctri);
        }
    }
    free(V); free(T); free(tr);
    return r;
}

void orc_sdf_info(const orc_sdf *r, uint32_t *dims, float *bone_info20, float *coarse_voxel) {
    dims[0] = r->w; dims[1] = r->h; dims[2] = r->d;
    memcpy(bone_info20, r->bone_info, sizeof(r->bone_info));
    *coarse_voxel = r->coarse_voxel;
}
void orc_sdf_texels(const orc_sdf *r, uint32_t *out) { memcpy(out, r->texels, 4 * (size_t)r->w * r->h * r->d); }
void orc_sdf_nearest(const orc_sdf *r, uint32_t *out) {
    if (r->nearest) memcpy(out, r->nearest, 4 * (size_t)r->w * r->h * r->d * 8);
}
void orc_sdf_free(orc_sdf *r) {
    if (!r) return;
    free(r->texels); free(r->nearest); free(r);
}

/* ---- SDF collider (collision.wgsl:116-265), trilinear filtering in the working type ---- */
static real texel_c(const uint32_t *tex, uint32_t w, uint32_t h, int32_t i, int32_t j, int32_t k, int ch) {
    const uint32_t t = tex[((size_t)k * h + (uint32_t)j) * w + (uint32_t)i];
    return (real)orc_f16_to_f32((uint16_t)(ch ? t >> 16 : t & 0xFFFFu));
}
static int32_t clampi(int32_t v, int32_t lo, int32_t hi) { return v < lo ? lo : v > hi ? hi : v; }
static void sample(const uint32_t *tex, uint32_t w, uint32_t h, uint32_t d, real u, real v, real ww, real *r, real *g) {
    const real x = u * (real)w - (real)0.5, y = v * (real)h - (real)0.5, z = ww * (real)d - (real)0.5;
    const real fx = (real)floor((double)x), fy = (real)floor((double)y), fz = (real)floor((double)z);
    const real tx = x - fx, ty = y - fy, tz = z - fz;
    const int32_t xi = (int32_t)fx, yi = (int32_t)fy, zi = (int32_t)fz;
    const int32_t x0 = clampi(xi, 0, (int32_t)w - 1), x1 = clampi(xi + 1, 0, (int32_t)w - 1);
    const int32_t y0 = clampi(yi, 0, (int32_t)h - 1), y1 = clampi(yi + 1, 0, (int32_t)h - 1);
    const int32_t z0 = clampi(zi, 0, (int32_t)d - 1), z1 = clampi(zi + 1, 0, (int32_t)d - 1);
    for (int ch = 0; ch < 2; ch++) {
#define TX(i, j, k) texel_c(tex, w, h, i, j, k, ch)
#define LERP(a, b, t) ((a) + ((b) - (a)) * (t))
        const real c00 = LERP(TX(x0, y0, z0), TX(x1, y0, z0), tx), c10 = LERP(TX(x0, y1, z0), TX(x1, y1, z0), tx);
        const real c01 = LERP(TX(x0, y0, z1), TX(x1, y0, z1), tx), c11 = LERP(TX(x0, y1, z1), TX(x1, y1, z1), tx);
        const real val = LERP(LERP(c00, c10, ty), LERP(c01, c11, ty), tz);
#undef TX
#undef LERP
        if (ch) *g = val; else *r = val;
    }
}
static s3 xform(const float *m, s3 v, real wc) {
    return sk((real)m[0] * v.x + (real)m[4] * v.y + (real)m[8] * v.z + (real)m[12] * wc,
              (real)m[1] * v.x + (real)m[5] * v.y + (real)m[9] * v.z + (real)m[13] * wc,
              (real)m[2] * v.x + (real)m[6] * v.y + (real)m[10] * v.z + (real)m[14] * wc);
}
static real clampr(real x, real lo, real hi) { return x < lo ? lo : x > hi ? hi : x; }

/* p3 / xn3: the vertex's position and committed-position estimate (updated in place by the
 * contact response); p_init3: the velocity reference.  Returns 1 when a contact was applied and
 * writes the response's target, normal, friction and restitution. */
int orc_sdf_contact(const uint32_t *tex, uint32_t w, uint32_t h, uint32_t d, const float *info, const float *xf,
                    uint32_t nb, uint32_t ntris_mesh, const real *p3, real vth, real *target3, real *normal3,
                    real *fr, real *rest) {
    if (!nb) return 0;
    const s3 p = tp(p3);
    real best_dist = (real)1e6, best_fr = 0, best_rest = 0, best_target = 0;
    s3 best_n = sk(0, 0, 0);
    uint32_t hits = 0;
    for (uint32_t b = 0; b < nb; b++) {
        const float *I = info + 20 * (size_t)b;
        const real radius = I[15];
        if (radius > 0) {
            const s3 dw = ssub(p, sk(I[3], I[7], I[11]));
            const real rt = radius + vth;
            if (sdot(dw, dw) > rt * rt) continue;
        }
        const float *wm = xf + 32 * (size_t)b, *im = wm + 16;
        const s3 pl = xform(im, p, 1);
        const real target = vth + (real)I[17];
        const real margin = target + (real)0.05f;
        if (pl.x < (real)I[0] - margin || pl.x > (real)I[4] + margin || pl.y < (real)I[1] - margin ||
            pl.y > (real)I[5] + margin || pl.z < (real)I[2] - margin || pl.z > (real)I[6] + margin)
            continue;
        const s3 sc = sk(I[8], I[9], I[10]), of = sk(I[12], I[13], I[14]);
        const s3 uvw = sk(pl.x * sc.x + of.x, pl.y * sc.y + of.y, pl.z * sc.z + of.z);
        const s3 umin = sk((real)I[0] * sc.x + of.x, (real)I[1] * sc.y + of.y, (real)I[2] * sc.z + of.z);
        const s3 umax = sk((real)I[4] * sc.x + of.x, (real)I[5] * sc.y + of.y, (real)I[6] * sc.z + of.z);
        if (uvw.x < umin.x || uvw.x > umax.x || uvw.y < umin.y || uvw.y > umax.y || uvw.z < umin.z || uvw.z > umax.z)
            continue;
        const s3 es = smul(ssub(umax, umin), nb == 1 ? (real)0.002f : (real)0.015f);
        const s3 lo = sadd(umin, es), hi = ssub(umax, es);
        const s3 us = sk(clampr(uvw.x, lo.x, hi.x), clampr(uvw.y, lo.y, hi.y), clampr(uvw.z, lo.z, hi.z));
        real raw, alpha, tmp;
        sample(tex, w, h, d, us.x, us.y, us.z, &raw, &alpha);
        const real threshold = (real)I[19] > (real)0.01f ? (real)I[19] : (real)0.01f;
        real retract = 0;
        if (ntris_mesh > 0) {
            const real t = clampr((alpha - threshold) / ((real)0.7f - threshold), 0, 1);
            retract = ((real)1 - t * t * ((real)3 - (real)2 * t)) * (real)0.06f;
        }
        const real eff = alpha > threshold ? raw + retract : (real)10;
        if (!(eff < target)) continue;
        const s3 eu = smul(sc, (real)0.005f);
        real dxp, dxn, dyp, dyn, dzp, dzn;
        sample(tex, w, h, d, clampr(uvw.x + eu.x, lo.x, hi.x), us.y, us.z, &dxp, &tmp);
        sample(tex, w, h, d, clampr(uvw.x - eu.x, lo.x, hi.x), us.y, us.z, &dxn, &tmp);
        sample(tex, w, h, d, us.x, clampr(uvw.y + eu.y, lo.y, hi.y), us.z, &dyp, &tmp);
        sample(tex, w, h, d, us.x, clampr(uvw.y - eu.y, lo.y, hi.y), us.z, &dyn, &tmp);
        sample(tex, w, h, d, us.x, us.y, clampr(uvw.z + eu.z, lo.z, hi.z), &dzp, &tmp);
        sample(tex, w, h, d, us.x, us.y, clampr(uvw.z - eu.z, lo.z, hi.z), &dzn, &tmp);
        real dx = dxp - dxn, dy = dyp - dyn, dz = dzp - dzn;
        if (fabs((double)dx) < 1e-6) dx = uvw.x < (lo.x + hi.x) * (real)0.5 ? dxp - raw : raw - dxn;
        if (fabs((double)dy) < 1e-6) dy = uvw.y < (lo.y + hi.y) * (real)0.5 ? dyp - raw : raw - dyn;
        if (fabs((double)dz) < 1e-6) dz = uvw.z < (lo.z + hi.z) * (real)0.5 ? dzp - raw : raw - dzn;
        const s3 g = sk(dx, dy, dz);
        const real gl = ssqrt(sdot(g, g));
        s3 nl = sk(0, 0, 1);
        if (gl > (real)1e-6f) {
            nl = sk(g.x / gl, g.y / gl, g.z / gl);
        } else {
            const s3 dc = ssub(pl, sk(((real)I[0] + (real)I[4]) * (real)0.5, ((real)I[1] + (real)I[5]) * (real)0.5,
                                      ((real)I[2] + (real)I[6]) * (real)0.5));
            const real lc = ssqrt(sdot(dc, dc));
            if (lc > (real)1e-6f) nl = sk(dc.x / lc, dc.y / lc, dc.z / lc);
        }
        const s3 nw0 = xform(wm, nl, 0);
        const real nwl = ssqrt(sdot(nw0, nw0));
        const s3 nw = sk(nw0.x / nwl, nw0.y / nwl, nw0.z / nwl);
        if (hits == 0 || eff < best_dist) {
            best_dist = eff; best_n = nw; best_fr = I[16]; best_rest = I[18]; best_target = target;
        }
        hits++;
    }
    if (hits > 0 && best_dist < best_target) {
        const real pen = best_target - best_dist;
        const s3 t = sadd(p, smul(best_n, pen));
        target3[0] = t.x; target3[1] = t.y; target3[2] = t.z;
        normal3[0] = best_n.x; normal3[1] = best_n.y; normal3[2] = best_n.z;
        *fr = best_fr;
        *rest = best_rest;
        return 1;
    }
    return 0;
}

/* ---- flat entry points for ctypes ---- */
void orc_f32_to_f16_n(const float *in, uint16_t *out, size_t n) {
    for (size_t i = 0; i < n; i++) out[i] = orc_f32_to_f16(in[i]);
}
void orc_f16_to_f32_n(const uint16_t *in, float *out, size_t n) {
    for (size_t i = 0; i < n; i++) out[i] = orc_f16_to_f32(in[i]);
}
/* Closes loops of (verts, tris); writes the grown mesh when the capacities (in elements) suffice.
 * Returns the loops closed; *nv_out / *nt_out receive the grown counts. */
uint32_t orc_close_loops_copy(const float *verts, uint32_t nv, const int32_t *tris, uint32_t nt, float *out_verts,
                              size_t cap_v, int32_t *out_tris, size_t cap_t, uint32_t *nv_out, uint32_t *nt_out) {
    float *V = (float *)malloc(sizeof(float) * 3 * (nv ? nv : 1));
    int32_t *T = (int32_t *)malloc(sizeof(int32_t) * 3 * (nt ? nt : 1));
    if (nv) memcpy(V, verts, sizeof(float) * 3 * nv);
    if (nt) memcpy(T, tris, sizeof(int32_t) * 3 * nt);
    uint32_t n = nv, m = nt;
    const uint32_t loops = orc_close_open_boundary_loops(&V, &n, &T, &m);
    *nv_out = n;
    *nt_out = m;
    if (out_verts && cap_v >= (size_t)3 * n) memcpy(out_verts, V, sizeof(float) * 3 * n);
    if (out_tris && cap_t >= (size_t)3 * m) memcpy(out_tris, T, sizeof(int32_t) * 3 * m);
    free(V);
    free(T);
    return loops;
}

/* ------------------------------------------------------------------ bone SDF atlas bake */
/* compute_3d_atlas_layout (sdf_baker.rs:137-163): o = cols, rows, layers, res, width, height, depth */
void orc_atlas_layout(uint32_t n, uint32_t res, uint32_t max_dim, uint32_t *o) {
    if (!n) { o[0] = o[1] = o[2] = 0; o[3] = res; o[4] = o[5] = o[6] = 0; return; }
    uint32_t cols = (uint32_t)ceil(pow((double)n, 1.0 / 3.0));
    if (cols < 1) cols = 1;
    uint32_t rows = (uint32_t)ceil(sqrt((double)((n + cols - 1) / cols)));
    if (rows < 1) rows = 1;
    uint32_t layers = (n + cols * rows - 1) / (cols * rows);
    if (layers < 1) layers = 1;
    uint32_t mt = cols > rows ? cols : rows;
    if (layers > mt) mt = layers;
    uint32_t r = res;
    if ((uint64_t)mt * res > max_dim) { r = max_dim / mt; if (r < 16) r = 16; }
    o[0] = cols; o[1] = rows; o[2] = layers; o[3] = r; o[4] = cols * r; o[5] = rows * r; o[6] = layers * r;
}

/* transform_point (sdf_baker.rs:167-178) */
static void bone_xform(const float *m, const float *p, float *o) {
    const float x = m[0] * p[0] + m[1] * p[1] + m[2] * p[2] + m[3];
    const float y = m[4] * p[0] + m[5] * p[1] + m[6] * p[2] + m[7];
    const float z = m[8] * p[0] + m[9] * p[1] + m[10] * p[2] + m[11];
    const float w = m[12] * p[0] + m[13] * p[1] + m[14] * p[2] + m[15];
    if (fabsf(w) > 1e-7f && fabsf(w - 1.0f) > 1e-5f) {
        const float iw = 1.0f / w;
        o[0] = x * iw; o[1] = y * iw; o[2] = z * iw;
    } else {
        o[0] = x; o[1] = y; o[2] = z;
    }
}

typedef struct { float p[3][3], w[3], n[3], bmask; } obtri;
typedef struct { float lmin[3], lmax[3]; uint32_t start, count, col, row, layer; } obparams;
typedef struct orc_bone_sdf {
    uint32_t w, h, d, nact;
    uint32_t *texels, *active;
    float *infos, *binds;
} orc_bone_sdf;

static inline s3 fv(const float *p) { return sk(p[0], p[1], p[2]); }
static inline s3 scross(s3 a, s3 b) { return sk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
static inline s3 snorm(s3 v) { const real l = ssqrt(sdot(v, v)); return sk(v.x / l, v.y / l, v.z / l); }
static inline real smax(real a, real b) { return (a != a) ? b : (b != b) ? a : (a > b ? a : b); }

/* bake_sdf.wgsl:34-196 for one voxel of one bone: the nearest triangle by Ericson's region test
 * (Real-Time Collision Detection 5.1.5) in triangle order, near ties within 1e-7 max(1, d^2) broken
 * toward the larger |normal projection|, the open-rim sign fix, and the weight at the nearest point. */
static uint32_t bone_voxel(const obparams *P, const obtri *tr, uint32_t res, uint32_t vx, uint32_t vy, uint32_t vz) {
    const real fr = (real)res;
    const real u = ((real)vx + (real)0.5) / fr, v = ((real)vy + (real)0.5) / fr, ww = ((real)vz + (real)0.5) / fr;
    const s3 lmin = fv(P->lmin), lmax = fv(P->lmax);
    const s3 lp = ssub(lmax, lmin);
    const s3 p = sk(lmin.x + u * lp.x, lmin.y + v * lp.y, lmin.z + ww * lp.z);
    real min_dsq = (real)1e10f, best_abs = -1, best_sd = 10, best_alpha = 0;
    for (uint32_t i = 0; i < P->count; i++) {
        const obtri *T = &tr[P->start + i];
        const s3 a = fv(T->p[0]), b = fv(T->p[1]), c = fv(T->p[2]);
        const s3 ab = ssub(b, a), ac = ssub(c, a), ap = ssub(p, a);
        const real d1 = sdot(ab, ap), d2 = sdot(ac, ap);
        s3 cp = a, bary = sk(1, 0, 0);
        if (!(d1 <= 0 && d2 <= 0)) {
            const s3 bp = ssub(p, b);
            const real d3 = sdot(ab, bp), d4 = sdot(ac, bp);
            const real vc = d1 * d4 - d3 * d2;
            if (d3 >= 0 && d4 <= d3) { cp = b; bary = sk(0, 1, 0); }
            else if (vc <= 0 && d1 >= 0 && d3 <= 0) {
                const real t = d1 / (d1 - d3);
                cp = sadd(a, smul(ab, t)); bary = sk(1 - t, t, 0);
            } else {
                const s3 cq = ssub(p, c);
                const real d5 = sdot(ab, cq), d6 = sdot(ac, cq);
                const real vb = d5 * d2 - d1 * d6;
                if (d6 >= 0 && d5 <= d6) { cp = c; bary = sk(0, 0, 1); }
                else if (vb <= 0 && d2 >= 0 && d6 <= 0) {
                    const real t = d2 / (d2 - d6);
                    cp = sadd(a, smul(ac, t)); bary = sk(1 - t, 0, t);
                } else {
                    const real va = d3 * d6 - d5 * d4;
                    if (va <= 0 && (d4 - d3) >= 0 && (d5 - d6) >= 0) {
                        const real t = (d4 - d3) / ((d4 - d3) + (d5 - d6));
                        cp = sadd(b, smul(ssub(c, b), t)); bary = sk(0, 1 - t, t);
                    } else {
                        const real den = 1 / (va + vb + vc);
                        const real sv = vb * den, sw = vc * den;
                        cp = sadd(sadd(a, smul(ab, sv)), smul(ac, sw)); bary = sk(1 - sv - sw, sv, sw);
                    }
                }
            }
        }
        const s3 diff = ssub(p, cp);
        const real dsq = sdot(diff, diff);
        const s3 tn = fv(T->n);
        const real proj = sdot(diff, tn), aproj = (real)fabs((double)proj);
        const real eps_sq = (real)1e-7f * (min_dsq > 1 ? min_dsq : (real)1);
        const int closer = dsq + eps_sq < min_dsq;
        const int tie = (real)fabs((double)(dsq - min_dsq)) <= eps_sq && aproj > best_abs;
        if (!(closer || tie)) continue;
        if (dsq < min_dsq) min_dsq = dsq;
        best_abs = aproj;
        const real d = ssqrt(dsq);
        real sign = proj >= 0 ? 1 : -1;
        const uint32_t bm = (uint32_t)(T->bmask + 0.5f);
        if (bm && sign < 0) {
            real op = -1;
            if ((bm & 1u) && bary.z <= (real)1e-5f) op = smax(op, sdot(diff, snorm(scross(ssub(b, a), tn))));
            if ((bm & 2u) && bary.x <= (real)1e-5f) op = smax(op, sdot(diff, snorm(scross(ssub(c, b), tn))));
            if ((bm & 4u) && bary.y <= (real)1e-5f) op = smax(op, sdot(diff, snorm(scross(ssub(a, c), tn))));
            if (op > 0 && op >= aproj * (real)0.25) sign = 1;
        }
        best_sd = d * sign;
        best_alpha = bary.x * T->w[0] + bary.y * T->w[1] + bary.z * T->w[2];
    }
    return (uint32_t)orc_f32_to_f16((float)best_sd) | ((uint32_t)orc_f32_to_f16((float)best_alpha) << 16);
}

/* bake_sdf.wgsl over bones [0, nbake) of a parameter array: params holds 16 words per bone in the
 * GpuBakeParams layout (local min[3], tri_start, local max[3], tri_count, col, row, layer, res, width,
 * height, depth, pitch), tris 16 floats per triangle in the GpuBakeTriangle layout (p0, w0, p1, w1, p2,
 * w2, normal, open-edge mask); each bone's res^3 tile is written into tex (width x height rows, Z/Y/X). */
void orc_bone_bake_tiles(const uint32_t *params, uint32_t nbake, const float *tris, uint32_t *tex) {
    for (uint32_t b = 0; b < nbake; b++) {
        const uint32_t *q = params + 16 * (size_t)b;
        obparams P;
        memcpy(P.lmin, q, 12); memcpy(P.lmax, q + 4, 12);
        P.start = q[3]; P.count = q[7]; P.col = q[8]; P.row = q[9]; P.layer = q[10];
        const uint32_t res = q[11], tw = q[12], th = q[13], pitch = q[15] ? q[15] : tw;
        obtri *tr = (obtri *)malloc(sizeof(obtri) * (P.count ? P.count : 1));
        for (uint32_t i = 0; i < P.count; i++) {
            const float *t = tris + 16 * (size_t)(P.start + i);
            for (int c = 0; c < 3; c++) {
                for (int k = 0; k < 3; k++) tr[i].p[c][k] = t[4 * c + k];
                tr[i].w[c] = t[4 * c + 3];
            }
            for (int k = 0; k < 3; k++) tr[i].n[k] = t[12 + k];
            tr[i].bmask = t[15];
        }
        obparams L = P;
        L.start = 0;
        for (uint32_t vz = 0; vz < res; vz++)
            for (uint32_t vy = 0; vy < res; vy++)
                for (uint32_t vx = 0; vx < res; vx++) {
                    const size_t gx = (size_t)P.col * res + vx, gy = (size_t)P.row * res + vy,
                                 gz = (size_t)P.layer * res + vz;
                    tex[(gz * th + gy) * pitch + gx] = bone_voxel(&L, tr, res, vx, vy, vz);
                }
        free(tr);
    }
}

/* bake_bone_sdf_gpu_ext (sdf_baker.rs:494-930).  weights: nbones x nv (bone-major); binds: nbones
 * row-major 4x4 bind (world -> bone) matrices.  Returns NULL when no bone has a weight above the
 * threshold or resolution is 0. */
orc_bone_sdf *orc_bone_sdf_bake(const float *verts, uint32_t nv, const int32_t *tris, uint32_t nt,
                                const float *weights, const float *binds, uint32_t nbones, uint32_t resolution,
                                float margin, float wth, float blend_k, float friction, float thickness,
                                float restitution, int close_open) {
    if (!resolution) return NULL;
    float *V = (float *)malloc(sizeof(float) * 3 * (nv ? nv : 1));
    int32_t *T = (int32_t *)malloc(sizeof(int32_t) * 3 * (nt ? nt : 1));
    if (nv) memcpy(V, verts, sizeof(float) * 3 * nv);
    if (nt) memcpy(T, tris, sizeof(int32_t) * 3 * nt);
    float **W = (float **)malloc(sizeof(float *) * (nbones ? nbones : 1));
    for (uint32_t b = 0; b < nbones; b++) {
        W[b] = (float *)malloc(sizeof(float) * (nv ? nv : 1));
        if (nv) memcpy(W[b], weights + (size_t)b * nv, sizeof(float) * nv);
    }
    uint32_t n = nv, m = nt;
    if (close_open) close_loops_weighted(&V, &n, &T, &m, W, nbones);
    /* undirected edge multiplicities over the valid triangles (sdf_baker.rs:203-240) */
    ecount *ec = (ecount *)malloc(sizeof(ecount) * 3 * (m ? m : 1));
    size_t nec = 0;
    for (uint32_t t = 0; t < m; t++) {
        const int32_t *tr = &T[3 * t];
        if (!tri_ok(tr, n)) continue;
        for (int e = 0; e < 3; e++) {
            const int32_t u = tr[e], v = tr[(e + 1) % 3];
            ecount c = {u < v ? u : v, u < v ? v : u, 1};
            ec[nec++] = c;
        }
    }
    qsort(ec, nec, sizeof(ecount), cmp_ecount);
    size_t q = 0;
    for (size_t i = 0; i < nec; i++) {
        if (q && ec[q - 1].a == ec[i].a && ec[q - 1].b == ec[i].b) ec[q - 1].count++;
        else ec[q++] = ec[i];
    }
    nec = q;
    uint32_t *act = (uint32_t *)malloc(sizeof(uint32_t) * (nbones ? nbones : 1)), na = 0;
    for (uint32_t b = 0; b < nbones; b++) {
        float mx = 0.0f;
        for (uint32_t v = 0; v < n; v++) mx = fmaxf(mx, W[b][v]);
        if (mx > wth) act[na++] = b;
    }
    orc_bone_sdf *r = NULL;
    if (na) {
        uint32_t L[7];
        orc_atlas_layout(na, resolution, 2048, L);
        const uint32_t cols = L[0], rows = L[1], layers = L[2], res = L[3];
        r = (orc_bone_sdf *)calloc(1, sizeof(orc_bone_sdf));
        r->w = L[4]; r->h = L[5]; r->d = L[6]; r->nact = na;
        r->active = act;
        act = NULL;
        r->infos = (float *)malloc(sizeof(float) * 20 * na);
        r->binds = (float *)malloc(sizeof(float) * 16 * na);
        obparams *P = (obparams *)malloc(sizeof(obparams) * na);
        obtri *tr = (obtri *)malloc(sizeof(obtri) * ((size_t)na * m + 1));
        uint32_t ntr = 0;
        for (uint32_t ob = 0; ob < na; ob++) {
            const uint32_t sb = r->active[ob];
            const float *mat = binds + 16 * (size_t)sb;
            memcpy(&r->binds[16 * (size_t)ob], mat, 16 * sizeof(float));
            float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY}, size[3];
            uint32_t valid = 0;
            for (uint32_t v = 0; v < n; v++)
                if (W[sb][v] > wth) {
                    float pl[3];
                    bone_xform(mat, &V[3 * (size_t)v], pl);
                    for (int k = 0; k < 3; k++) { mn[k] = fminf(mn[k], pl[k]); mx[k] = fmaxf(mx[k], pl[k]); }
                    valid++;
                }
            if (!valid)
                for (int k = 0; k < 3; k++) { mn[k] = -0.1f; mx[k] = 0.1f; }
            for (int k = 0; k < 3; k++) size[k] = fmaxf(mx[k] - mn[k], 0.05f);
            for (int k = 0; k < 3; k++) {
                const float pad = size[k] * margin + fmaxf(thickness, 0.0f);
                mn[k] -= pad; mx[k] += pad;
                size[k] = mx[k] - mn[k];
            }
            const uint32_t bpl = cols * rows, l = ob / bpl, rem = ob % bpl, row = rem / cols, col = rem % cols;
            float *info = &r->infos[20 * (size_t)ob];
            const float nn[3] = {(float)cols, (float)rows, (float)layers}, ix[3] = {(float)col, (float)row, (float)l};
            for (int k = 0; k < 3; k++) {
                info[k] = mn[k];
                info[4 + k] = mx[k];
                info[8 + k] = (1.0f / size[k]) / nn[k];
                info[12 + k] = (-mn[k] / size[k] + ix[k]) / nn[k];
            }
            info[3] = (float)ob; info[7] = (float)res; info[11] = blend_k; info[15] = 0.0f;
            info[16] = friction; info[17] = thickness; info[18] = restitution; info[19] = wth;
            const uint32_t start = ntr;
            for (uint32_t t = 0; t < m; t++) {
                const int32_t *tt = &T[3 * t];
                if ((uint32_t)tt[0] >= n || (uint32_t)tt[1] >= n || (uint32_t)tt[2] >= n) continue;
                const float w0 = W[sb][tt[0]], w1 = W[sb][tt[1]], w2 = W[sb][tt[2]];
                if (!(w0 > 0.0f || w1 > 0.0f || w2 > 0.0f)) continue;
                obtri o;
                for (int c = 0; c < 3; c++) bone_xform(mat, &V[3 * (size_t)tt[c]], o.p[c]);
                int ov = 1;
                for (int k = 0; k < 3; k++) {
                    const float tmin = fminf(fminf(o.p[0][k], o.p[1][k]), o.p[2][k]);
                    const float tmax = fmaxf(fmaxf(o.p[0][k], o.p[1][k]), o.p[2][k]);
                    ov = ov && tmin <= mx[k] && tmax >= mn[k];
                }
                if (!ov) continue;
                const float e01[3] = {o.p[1][0] - o.p[0][0], o.p[1][1] - o.p[0][1], o.p[1][2] - o.p[0][2]};
                const float e02[3] = {o.p[2][0] - o.p[0][0], o.p[2][1] - o.p[0][1], o.p[2][2] - o.p[0][2]};
                float cr[3];
                cross3(e01, e02, cr);
                normalize3(cr, o.n);
                uint32_t bm = 0;
                if (edge_count(ec, nec, tt[0], tt[1]) == 1u) bm |= 1u;
                if (edge_count(ec, nec, tt[1], tt[2]) == 1u) bm |= 2u;
                if (edge_count(ec, nec, tt[2], tt[0]) == 1u) bm |= 4u;
                o.w[0] = w0; o.w[1] = w1; o.w[2] = w2;
                o.bmask = (float)bm;
                tr[ntr++] = o;
            }
            P[ob].start = start; P[ob].count = ntr - start;
            P[ob].col = col; P[ob].row = row; P[ob].layer = l;
            memcpy(P[ob].lmin, mn, sizeof(mn));
            memcpy(P[ob].lmax, mx, sizeof(mx));
        }
        const size_t nvox = (size_t)r->w * r->h * r->d;
        r->texels = (uint32_t *)calloc(nvox ? nvox : 1, sizeof(uint32_t));
        for (uint32_t ob = 0; ob < na; ob++) {
            const obparams *pp = &P[ob];
            ORC_PARALLEL_FOR
            for (int32_t vz = 0; vz < (int32_t)res; vz++)
                for (uint32_t vy = 0; vy < res; vy++)
                    for (uint32_t vx = 0; vx < res; vx++) {
                        const size_t gx = pp->col * res + vx, gy = pp->row * res + vy, gz = pp->layer * res + (uint32_t)vz;
                        r->texels[(gz * r->h + gy) * r->w + gx] = bone_voxel(pp, tr, res, vx, vy, (uint32_t)vz);
                    }
        }
        free(P);
        free(tr);
    }
    free(act);
    free(ec);
    for (uint32_t b = 0; b < nbones; b++) free(W[b]);
    free(W); free(V); free(T);
    return r;
}

void orc_bone_sdf_info(const orc_bone_sdf *r, uint32_t *dims, uint32_t *n_active) {
    dims[0] = r->w; dims[1] = r->h; dims[2] = r->d;
    *n_active = r->nact;
}
void orc_bone_sdf_arrays(const orc_bone_sdf *r, uint32_t *texels, float *infos, uint32_t *active, float *binds) {
    if (texels) memcpy(texels, r->texels, 4 * (size_t)r->w * r->h * r->d);
    if (infos) memcpy(infos, r->infos, sizeof(float) * 20 * r->nact);
    if (active) memcpy(active, r->active, sizeof(uint32_t) * r->nact);
    if (binds) memcpy(binds, r->binds, sizeof(float) * 16 * r->nact);
}
void orc_bone_sdf_free(orc_bone_sdf *r) {
    if (!r) return;
    free(r->texels); free(r->active); free(r->infos); free(r->binds);
    free(r);
}
