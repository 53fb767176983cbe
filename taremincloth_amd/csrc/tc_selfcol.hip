// Spatial hash and DIRECT self-collision kernels for gfx950.
//
// Reference semantics (paths relative to the reference repository root):
//   hash build        crates/cloth_core/src/shaders/spatial_grid_sort.wgsl:25-195
//   edge hash build   crates/cloth_core/src/shaders/spatial_grid_sort_edges.wgsl:51-199
//   V-V / V-T solve   crates/cloth_core/src/shaders/self_collision.wgsl:71-681
//   E-E solve         crates/cloth_core/src/shaders/self_collision_ee_direct.wgsl:66-776
//   apply             crates/cloth_core/src/shaders/self_collision_apply.wgsl:57-177
//
// Hash build, vertices and edge midpoints together: count (atomics) -> exclusive
// scan of each table (one workgroup per row of 1024 entries, one over the row
// totals, one add) -> scatter (atomic cursor) -> per-bucket sort, so every
// bucket lists its elements in ascending index order.  The reference's scatter
// order inside a bucket depends on atomic timing; the sorted order makes the
// hash tables, and with them the E-E search cut-off, deterministic.
//
// The solve kernels accumulate displacements in 1e-6 m fixed point with
// integer atomics, which makes the result independent of thread order.
#include "tc_kernels.h"

namespace tc {

static inline uint32_t grid_for(size_t n, uint32_t block) { return uint32_t((n + block - 1) / block); }

// The vertex table (0) and the edge-midpoint table (1) have the same bucket count and are built
// together: every launch below serves both, the table picked by the thread or workgroup index.
struct HashPair {
    HashBufs t[2];
    uint32_t n[2];      // elements: vertices, edges
    uint32_t table;     // buckets per table
};

__global__ void __launch_bounds__(256) k_hash_clear(HashPair hp) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < 2u * hp.table) hp.t[i / hp.table].counts[i % hp.table] = 0u;
}

__global__ void __launch_bounds__(256) k_hash_count(HashPair hp, const float4 *__restrict__ P,
                                                    const uint2 *__restrict__ e, float cell) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= hp.n[0] + hp.n[1]) return;
    uint32_t h, k;
    if (i < hp.n[0]) {
        k = i;
        h = hash_cell(cell_of(xyz(P[k]), cell), hp.table);
        hp.t[0].keys[k] = h;
        atomicAdd(&hp.t[0].counts[h], 1u);
    } else {
        k = i - hp.n[0];
        const uint2 ed = e[k];
        const f3 mid = (xyz(P[ed.x]) + xyz(P[ed.y])) * 0.5f;
        h = hash_cell(cell_of(mid, cell), hp.table);
        hp.t[1].keys[k] = h;
        atomicAdd(&hp.t[1].counts[h], 1u);
    }
}

// Exclusive scan of counts[0..table) into starts[0..table], starts[table] = total,
// for both tables in three launches over rows of 1024 entries (at most SCAN_ROWS rows per
// table), coalesced throughout.  k_scan_rows (one workgroup per row and table) scans each row
// with wave64 prefix sums and the 16 wavefront totals, writes the in-row exclusive sums, keeps
// the row total in counts[table + row] and resets counts[0..table) to zero for use as scatter
// cursors; k_scan_totals (one workgroup per table) turns the row totals into their exclusive
// prefix and writes starts[table]; k_scan_add adds each row's prefix.
constexpr uint32_t SCAN_ROWS = 4096;
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t x, uint32_t lane) {
    for (uint32_t o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    return x;
}

__global__ void __launch_bounds__(1024) k_scan_rows(HashPair hp, uint32_t rows) {
    __shared__ uint32_t wave_tot[16];
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    const uint32_t tb = blockIdx.x / rows, row = blockIdx.x % rows, table = hp.table;
    uint32_t *__restrict__ counts = hp.t[tb].counts;
    uint32_t *__restrict__ starts = hp.t[tb].starts;
    const uint32_t k = (row << 10) + tid;
    const uint32_t c = k < table ? counts[k] : 0u;
    const uint32_t incl = wave_incl_scan(c, lane);
    if (lane == 63u) wave_tot[wave] = incl;
    __syncthreads();
    uint32_t base = 0;
    for (uint32_t w = 0; w < wave; ++w) base += wave_tot[w];
    if (k < table) {
        starts[k] = base + incl - c;
        counts[k] = 0u;
    }
    if (tid == 1023u) counts[table + row] = base + incl;
}

// rows <= SCAN_ROWS = 4 x 1024: thread t owns rows 4t .. 4t+3
__global__ void __launch_bounds__(1024) k_scan_totals(HashPair hp, uint32_t rows) {
    __shared__ uint32_t wave_tot[16];
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6, table = hp.table;
    uint32_t *tot = hp.t[blockIdx.x].counts + table;
    uint32_t v[4], sum = 0;
#pragma unroll
    for (uint32_t u = 0; u < 4; ++u) {
        const uint32_t r = 4u * tid + u;
        v[u] = r < rows ? tot[r] : 0u;
        sum += v[u];
    }
    const uint32_t incl = wave_incl_scan(sum, lane);
    if (lane == 63u) wave_tot[wave] = incl;
    __syncthreads();
    uint32_t run = 0;
    for (uint32_t w = 0; w < wave; ++w) run += wave_tot[w];
    run += incl - sum;
#pragma unroll
    for (uint32_t u = 0; u < 4; ++u) {
        const uint32_t r = 4u * tid + u;
        if (r < rows) tot[r] = run;
        run += v[u];
    }
    if (tid == 1023u) hp.t[blockIdx.x].starts[table] = run;
}

__global__ void __launch_bounds__(1024) k_scan_add(HashPair hp, uint32_t rows) {
    const uint32_t tb = blockIdx.x / rows, row = blockIdx.x % rows;
    const uint32_t k = (row << 10) + threadIdx.x;
    if (k < hp.table) hp.t[tb].starts[k] += hp.t[tb].counts[hp.table + row];
}

__global__ void __launch_bounds__(256) k_hash_scatter(HashPair hp) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= hp.n[0] + hp.n[1]) return;
    const uint32_t tb = i < hp.n[0] ? 0u : 1u, k = tb ? i - hp.n[0] : i;
    const HashBufs &h = hp.t[tb];
    const uint32_t b = h.keys[k];
    h.sorted[h.starts[b] + atomicAdd(&h.counts[b], 1u)] = k;
}

__global__ void __launch_bounds__(256) k_bucket_sort(HashPair hp) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 2u * hp.table) return;
    const uint32_t tb = i / hp.table, h = i % hp.table;
    const uint32_t *__restrict__ starts = hp.t[tb].starts;
    uint32_t *__restrict__ sorted = hp.t[tb].sorted;
    const uint32_t b = starts[h], e = starts[h + 1];
    for (uint32_t k = b + 1; k < e; ++k) {
        const uint32_t x = sorted[k];
        uint32_t j = k;
        while (j > b && sorted[j - 1] > x) { sorted[j] = sorted[j - 1]; --j; }
        sorted[j] = x;
    }
}

hipError_t launch_hashes(hipStream_t s, const float4 *P, uint32_t nv, const uint2 *edges, uint32_t ne, float cell,
                         uint32_t table, const HashBufs &vh, const HashBufs &eh) {
    if (table == 0u || table > SCAN_ROWS * 1024u) return hipErrorInvalidValue;
    HashPair hp;
    hp.t[0] = vh;
    hp.t[1] = eh;
    hp.n[0] = nv;
    hp.n[1] = ne;
    hp.table = table;
    const uint32_t rows = (table + 1023u) >> 10, nall = nv + ne;
    hipLaunchKernelGGL(k_hash_clear, dim3(grid_for(2 * size_t(table), 256)), dim3(256), 0, s, hp);
    if (nall) hipLaunchKernelGGL(k_hash_count, dim3(grid_for(nall, 256)), dim3(256), 0, s, hp, P, edges, cell);
    hipLaunchKernelGGL(k_scan_rows, dim3(2 * rows), dim3(1024), 0, s, hp, rows);
    hipLaunchKernelGGL(k_scan_totals, dim3(2), dim3(1024), 0, s, hp, rows);
    hipLaunchKernelGGL(k_scan_add, dim3(2 * rows), dim3(1024), 0, s, hp, rows);
    if (nall) hipLaunchKernelGGL(k_hash_scatter, dim3(grid_for(nall, 256)), dim3(256), 0, s, hp);
    hipLaunchKernelGGL(k_bucket_sort, dim3(grid_for(2 * size_t(table), 256)), dim3(256), 0, s, hp);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// fixed-point accumulation (self_collision.wgsl:71-116)
__device__ __forceinline__ int32_t to_fixed(float x) {
    return f2i_sat(clampf(x * 1000000.0f, -50000000.0f, 50000000.0f));
}
__device__ __forceinline__ void add_disp(Accum *a, f3 d) {
    atomicAdd(&a->dx, to_fixed(d.x));
    atomicAdd(&a->dy, to_fixed(d.y));
    atomicAdd(&a->dz, to_fixed(d.z));
    atomicAdd(&a->count, 1u);
}
__device__ __forceinline__ void add_pin_disp(Accum *a, f3 d) {
    atomicAdd(&a->pdx, to_fixed(d.x));
    atomicAdd(&a->pdy, to_fixed(d.y));
    atomicAdd(&a->pdz, to_fixed(d.z));
    atomicAdd(&a->pcount, 1u);
}
__device__ __forceinline__ void ccd_axis(int32_t v, int32_t *pos, int32_t *negv) {
    if (v > 0) atomicMax(pos, v);
    else if (v < 0) atomicMin(negv, v);
}
__device__ __forceinline__ void add_ccd(Accum *a, f3 d) {
    ccd_axis(to_fixed(d.x), &a->cpx, &a->cnx);
    ccd_axis(to_fixed(d.y), &a->cpy, &a->cny);
    ccd_axis(to_fixed(d.z), &a->cpz, &a->cnz);
}
// self_collision_ee_direct.wgsl:70-79: the E-E proximity path counts each contribution twice
__device__ __forceinline__ void add_disp_ee(Accum *a, f3 d) {
    atomicAdd(&a->dx, to_fixed(d.x));
    atomicAdd(&a->dy, to_fixed(d.y));
    atomicAdd(&a->dz, to_fixed(d.z));
    atomicAdd(&a->count, 2u);
}
__device__ __forceinline__ void add_ccd_ee(Accum *a, f3 d, bool count_ee) {
    add_ccd(a, d);
    if (count_ee) {
        const float l = clampf(length(d) * 1000000.0f, 0.0f, 2000000000.0f);
        if ((uint32_t)l > 0u) atomicAdd(&a->ee_count, 1u);
    }
}

// self_collision.wgsl:134-179
__device__ __forceinline__ f3 closest_bary(f3 p, f3 a, f3 b, f3 c, f3 &bary) {
    const f3 ab = b - a, ac = c - a, ap = p - a;
    const float d1 = dot(ab, ap), d2 = dot(ac, ap);
    if (d1 <= 0.0f && d2 <= 0.0f) { bary = mk3(1.0f, 0.0f, 0.0f); return a; }
    const f3 bp = p - b;
    const float d3 = dot(ab, bp), d4 = dot(ac, bp);
    if (d3 >= 0.0f && d4 <= d3) { bary = mk3(0.0f, 1.0f, 0.0f); return b; }
    const float vc = d1 * d4 - d3 * d2;
    if (vc <= 0.0f && d1 >= 0.0f && d3 <= 0.0f) {
        const float v = d1 / (d1 - d3);
        bary = mk3(1.0f - v, v, 0.0f);
        return a + ab * v;
    }
    const f3 cp = p - c;
    const float d5 = dot(ab, cp), d6 = dot(ac, cp);
    if (d6 >= 0.0f && d5 <= d6) { bary = mk3(0.0f, 0.0f, 1.0f); return c; }
    const float vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0f && d2 >= 0.0f && d6 <= 0.0f) {
        const float w = d2 / (d2 - d6);
        bary = mk3(1.0f - w, 0.0f, w);
        return a + ac * w;
    }
    const float va = d3 * d6 - d5 * d4;
    if (va <= 0.0f && (d4 - d3) >= 0.0f && (d5 - d6) >= 0.0f) {
        const float w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        bary = mk3(0.0f, 1.0f - w, w);
        return b + (c - b) * w;
    }
    const float denom = 1.0f / (va + vb + vc);
    const float v = vb * denom, w = vc * denom;
    bary = mk3(1.0f - v - w, v, w);
    return a + ab * v + ac * w;
}

// self_collision.wgsl:182-218; returns hit with (t, u, v)
__device__ __forceinline__ bool seg_tri_uv(f3 po, f3 pn, f3 a, f3 b, f3 c, float &t_o, float &u_o, float &v_o) {
    const f3 d = pn - po;
    const float ld = dot(d, d);
    if (ld < 1e-12f) return false;
    const f3 e1 = b - a, e2 = c - a;
    const f3 pv = cross(d, e2);
    const float det = dot(e1, pv);
    const float tn = dot(e1, e1) * dot(e2, e2);
    if (det * det <= tn * ld * 1e-10f || tn <= 1e-20f) return false;
    const float inv = 1.0f / det;
    const f3 tv = po - a;
    const float u = dot(tv, pv) * inv;
    if (u < -0.05f || u > 1.05f) return false;
    const f3 qv = cross(tv, e1);
    const float v = dot(d, qv) * inv;
    if (v < -0.05f || (u + v) > 1.05f) return false;
    const float t = dot(e2, qv) * inv;
    if (t >= 0.0f && t <= 1.0f) { t_o = t; u_o = u; v_o = v; return true; }
    return false;
}

// two-hop lookups (self_collision.wgsl:262-314, ee_direct.wgsl:259-283):
// returns -1 if b is not in the two-hop list of a, else its slot
__device__ __forceinline__ int32_t two_hop_slot(const SelfColTopo &t, uint32_t a, uint32_t b) {
    uint32_t lo = t.th_off[a], hi = t.th_off[a + 1];
    while (lo < hi) {
        const uint32_t mid = lo + (hi - lo) / 2, val = t.th_idx[mid];
        if (val == b) return int32_t(mid);
        if (val < b) lo = mid + 1; else hi = mid;
    }
    return -1;
}
__device__ __forceinline__ bool within_two_hops(const SelfColTopo &t, uint32_t a, uint32_t b) {
    return a == b || two_hop_slot(t, a, b) >= 0;
}
__device__ __forceinline__ bool near_folded_sq(const SelfColTopo &t, uint32_t a, uint32_t b, float cur_sq) {
    if (a == b) return true;
    const int32_t s = two_hop_slot(t, a, b);
    if (s < 0) return false;
    const float l0 = t.th_rest[s];
    return !(cur_sq < (l0 * l0) * 0.25f);
}

// Work distribution of the two solve kernels.  A wavefront takes SC_GROUPS
// groups of two elements (vertices or edges) one after the other; in each group,
// lanes 0-26 and 32-58 walk the hash bucket of one of the element's 27 neighbour
// cells, numbered in the order of the reference's E-E loop (dz outer, dx inner:
// cell = 9 (dz+1) + 3 (dy+1) + (dx+1)).  The candidates that pass the cheap
// tests go into a per-wavefront LDS queue, which the wavefront works off 64 pairs
// at a time, one pair per lane, so that the expensive pair tests (CCD sweeps,
// two-hop lookups) run with full lanes; the queue carries over from one group to
// the next.  The fixed-point accumulation makes the result independent of this
// order.
constexpr uint32_t SC_BLOCK = 256;
constexpr uint32_t SC_GROUPS = 4;
constexpr uint32_t SC_PER_BLOCK = (SC_BLOCK / 64) * SC_GROUPS * 2;  // elements per workgroup
constexpr uint32_t SC_QCAP = 128;  // queue entries per wavefront: < 64 left over + up to 64 pushed per step
constexpr uint32_t SC_NONE = 0xFFFFFFFFu, SC_STOP = 0xFFFFFFFEu;

__device__ __forceinline__ i3 neighbour_cell(i3 c, uint32_t k) {
    return {c.x + (int32_t(k % 3u) - 1), c.y + (int32_t((k / 3u) % 3u) - 1), c.z + (int32_t(k / 9u) - 1)};
}

// Lockstep walk of every lane's bucket range [q, qe).  accept(k) tests entry k of
// the bucket list and returns the partner element, SC_NONE to skip it, or SC_STOP
// to end this lane's walk; pair(self, partner) processes one queued pair.  qn is
// the number of queued pairs (the same in every lane); a full batch of 64 is
// processed as soon as it is there, the rest only when `flush` is set.  Every
// lane of the wavefront must call this (the ballots and the queue are shared).
template <class Accept, class Pair>
__device__ __forceinline__ void sc_walk(uint2 *queue, uint32_t &qn, bool flush, uint32_t lane, uint32_t self,
                                        uint32_t q, uint32_t qe, Accept accept, Pair pair) {
    for (;;) {
        const bool more = __any(q < qe) != 0;
        bool push = false;
        uint32_t other = 0;
        if (q < qe) {
            other = accept(q);
            ++q;
            if (other == SC_STOP) q = qe;
            else push = other != SC_NONE;
        }
        const uint64_t m = __ballot(push);
        if (push) queue[qn + uint32_t(__popcll(m & ((uint64_t(1) << lane) - 1u)))] = make_uint2(self, other);
        qn += uint32_t(__popcll(m));
        if (qn >= 64u || (flush && !more && qn != 0u)) {
            const uint32_t take = min(qn, 64u), rest = qn - take;
            __builtin_amdgcn_wave_barrier();  // the pushes above come first
            uint2 pr = make_uint2(0u, 0u), mv = make_uint2(0u, 0u);
            if (lane < take) pr = queue[lane];
            if (lane < rest) mv = queue[64u + lane];
            __builtin_amdgcn_wave_barrier();
            if (lane < rest) queue[lane] = mv;
            qn = rest;
            if (lane < take) pair(pr.x, pr.y);
        }
        if (!more && (qn == 0u || !flush)) break;
    }
}

// V-V and V-T response of vertex `index` against vertex j and j's star, for a
// pair that passed the cell and reach tests (self_collision.wgsl:388-681).  UT: the multi-layer
// untangling exemptions (a separate instantiation keeps single-layer cloth at its register count).
template <bool UT>
__device__ __forceinline__ void vv_pair(const float4 *__restrict__ P, const float4 *__restrict__ X,
                                        const SelfColTopo &T, Accum *__restrict__ acc, const SelfColArgs &A,
                                        uint32_t index, uint32_t j) {
    const float EPS = 1e-7f;
    const float4 qi = P[index], xi = X[index];
    const bool pinned_i = qi.w <= 0.0f;
    const f3 xo_i = xyz(xi), p_i = xyz(qi);
    const float mv_i = length(p_i - xo_i);
    const float thick_i = xi.w;
    const float llen_i = T.local_len[index];
    const i3 cc = cell_of(p_i, A.cell);
    const float4 qj = P[j], xj = X[j];
    const f3 p_j = xyz(qj);
    const f3 dvv = p_i - p_j;
    const float dsq = dot(dvv, dvv);
    const float thick_j = xj.w, llen_j = T.local_len[j];
    const float eff = fmaxf(thick_i + thick_j, fminf(llen_i, llen_j) * 0.06f);
    const float dvv_len = sqrtf(fmaxf(dsq, 0.0f));
    // one lookup serves both tests on (index, j): near_folded and within_two_hops
    const int32_t s_ij = two_hop_slot(T, index, j);
    if (s_ij >= 0 && !(dvv_len < T.th_rest[s_ij] * 0.5f)) return;
    const f3 pj_old = xyz(xj);
    const float wi = qi.w, wj = qj.w;
    const bool both_dynamic = wj > 0.0f;
    const float eff_sq = eff * eff;
    constexpr bool ut = UT;
    const uint32_t layer_i = ut ? T.layer[index] : 0u, layer_j = ut ? T.layer[j] : 0u;
    // 1. vertex-vertex
    if (!pinned_i && (!both_dynamic || index < j)) {
        bool untangling = false;  // the outer vertex sits behind the inner one: the untangle pass moves it
        if (ut && dsq < eff_sq && dsq > EPS * EPS && layer_i != layer_j) {
            const float4 ni4 = T.normals[index], nj4 = T.normals[j];
            const f3 ni = xyz(ni4), nj = xyz(nj4);
            const float nil = length(ni), njl = length(nj);
            if (layer_i > layer_j) {
                if (njl > 0.5f) {
                    const f3 nju = nj / njl;
                    const bool allow = wj <= 0.0f || nil <= 0.5f || dot(ni / fmaxf(nil, EPS), nju) > 0.0f;
                    untangling = allow && dot(dvv, nju) < 0.0f;
                }
            } else if (both_dynamic && nil > 0.5f) {
                const f3 niu = ni / nil;
                const bool allow = njl <= 0.5f || dot(nj / fmaxf(njl, EPS), niu) > 0.0f;
                untangling = allow && dot(-dvv, niu) < 0.0f;
            }
        }
        if (!untangling && dsq < eff_sq && dsq > EPS * EPS) {
            const f3 n = dvv / dvv_len;
            const float pen = eff - dvv_len;
            const float ws = wi + wj;
            if (ws > EPS) {
                if (wj <= 0.0f) {
                    add_pin_disp(&acc[index], n * pen + (p_j - pj_old) * 0.5f);
                } else {
                    const f3 disp = n * (pen * 0.5f);
                    add_disp(&acc[index], disp * (wi / ws));
                    add_disp(&acc[j], (-disp) * (wj / ws));
                }
            }
        }
    }
    // 2. vertex-triangle over the star of j
    const float mv_j = length(p_j - pj_old);
    const float vt_bound = eff + llen_j * 1.3f + mv_i + mv_j;
    if (!(dsq <= vt_bound * vt_bound) || s_ij >= 0) return;
    for (uint32_t st = T.star_off[j]; st < T.star_off[j + 1]; ++st) {
        const uint2 pr = T.star[st];
        const uint32_t v0 = pr.x, v1 = pr.y;
        if (v0 == index || v1 == index) continue;
        const float4 q0 = P[v0], q1 = P[v1], x0 = X[v0], x1 = X[v1];
        const f3 p_v0 = xyz(q0), p_v1 = xyz(q1), p_v0o = xyz(x0), p_v1o = xyz(x1);
        const f3 d0 = p_i - p_v0;
        const float ll0 = T.local_len[v0];
        const float et0 = fmaxf(thick_i + x0.w, fminf(llen_i, ll0) * 0.06f);
        const float b0 = et0 + ll0 * 1.3f + mv_i + length(p_v0 - p_v0o);
        const i3 c0 = cell_of(p_v0, A.cell);
        const bool in0 = abs(c0.x - cc.x) <= 1 && abs(c0.y - cc.y) <= 1 && abs(c0.z - cc.z) <= 1 &&
                         dot(d0, d0) <= b0 * b0;
        const f3 d1 = p_i - p_v1;
        const float ll1 = T.local_len[v1];
        const float et1 = fmaxf(thick_i + x1.w, fminf(llen_i, ll1) * 0.06f);
        const float b1 = et1 + ll1 * 1.3f + mv_i + length(p_v1 - p_v1o);
        const i3 c1 = cell_of(p_v1, A.cell);
        const bool in1 = abs(c1.x - cc.x) <= 1 && abs(c1.y - cc.y) <= 1 && abs(c1.z - cc.z) <= 1 &&
                         dot(d1, d1) <= b1 * b1;
        if ((in0 && j > v0) || (in1 && j > v1)) continue;
        const f3 tri_move = (((p_j - pj_old) + (p_v0 - p_v0o)) + (p_v1 - p_v1o)) / 3.0f;
        const bool tri_pin = wj <= 0.0f || q0.w <= 0.0f || q1.w <= 0.0f;
        const f3 rel_old = xo_i + tri_move;
        const f3 ev = mk3(eff, eff, eff);
        const f3 smin = min3(rel_old, p_i) - ev, smax = max3(rel_old, p_i) + ev;
        const f3 tmin = min3(min3(p_j, p_v0), p_v1), tmax = max3(max3(p_j, p_v0), p_v1);
        if (smin.x > tmax.x || smin.y > tmax.y || smin.z > tmax.z || smax.x < tmin.x || smax.y < tmin.y ||
            smax.z < tmin.z)
            continue;
        if (within_two_hops(T, index, v0) || within_two_hops(T, index, v1)) continue;
        const f3 move = p_i - rel_old;
        const float move_sq = dot(move, move);
        float ht = 1.0f, hu = 0.0f, hv = 0.0f;
        bool hit = false;
        if (move_sq > 1e-12f) hit = seg_tri_uv(rel_old, p_i, p_j, p_v0, p_v1, ht, hu, hv);
        const float w0 = q0.w, w1 = q1.w;
        if (hit) {
            const f3 hit_pt = rel_old + move * ht;
            const f3 tcr = cross(p_v0 - p_j, p_v1 - p_j);
            const float tl = length(tcr);
            const f3 tn = tl > 1e-12f ? tcr / tl : mk3(0.0f, 1.0f, 0.0f);
            const float h_old_raw = dot(rel_old - p_j, tn);
            if (ut && !pinned_i && layer_i > layer_j && tl > 1e-12f) {
                // a free outer vertex returning from behind the triangle to its front is let through
                const f3 ni = xyz(T.normals[index]);
                const float nil = length(ni);
                const bool allow = tri_pin || nil <= 0.5f || dot(ni / fmaxf(nil, EPS), tn) > 0.0f;
                if (allow && h_old_raw < 0.0f && dot(move, tn) > 0.0f) continue;
            }
            const float side = fabsf(h_old_raw) > 1e-6f ? h_old_raw : dot(-move, tn);
            const f3 n_side = side >= 0.0f ? tn : -tn;
            const float h_old = fmaxf(dot(rel_old - hit_pt, n_side), 0.0f);
            const float h_safe = fmaxf(h_old, eff * 0.25f);
            const f3 total = (hit_pt + n_side * h_safe) - p_i;
            const float pen_soft = fmaxf(0.0f, eff - h_safe);
            const f3 bc = mk3(clampf(1.0f - hu - hv, 0.0f, 1.0f), clampf(hu, 0.0f, 1.0f), clampf(hv, 0.0f, 1.0f));
            if (tri_pin) {
                if (!pinned_i) add_ccd(&acc[index], total + n_side * pen_soft);
            } else if (pinned_i) {
                const float wt = bc.x * bc.x * wj + bc.y * bc.y * w0 + bc.z * bc.z * w1;
                if (wt > EPS) {
                    const f3 full = total + n_side * pen_soft;
                    if (wj > 0.0f) add_ccd(&acc[j], (-full) * (bc.x * wj / wt));
                    if (w0 > 0.0f) add_ccd(&acc[v0], (-full) * (bc.y * w0 / wt));
                    if (w1 > 0.0f) add_ccd(&acc[v1], (-full) * (bc.z * w1 / wt));
                }
            } else {
                const float wt = bc.x * bc.x * wj + bc.y * bc.y * w0 + bc.z * bc.z * w1;
                const float wtot = wi + wt;
                if (wtot > EPS) {
                    add_ccd(&acc[index], total * (wi / wtot));
                    if (wj > 0.0f) add_ccd(&acc[j], (-total) * (bc.x * wj / wtot));
                    if (w0 > 0.0f) add_ccd(&acc[v0], (-total) * (bc.y * w0 / wtot));
                    if (w1 > 0.0f) add_ccd(&acc[v1], (-total) * (bc.z * w1 / wtot));
                    if (pen_soft > EPS) {
                        const f3 tp = n_side * (pen_soft * 0.6f);
                        add_disp(&acc[index], tp * (wi / wtot));
                        if (wj > 0.0f) add_disp(&acc[j], (-tp) * (bc.x * wj / wtot));
                        if (w0 > 0.0f) add_disp(&acc[v0], (-tp) * (bc.y * w0 / wtot));
                        if (w1 > 0.0f) add_disp(&acc[v1], (-tp) * (bc.z * w1 / wtot));
                    }
                }
            }
        } else {
            f3 bc;
            const f3 cp = closest_bary(p_i, p_j, p_v0, p_v1, bc);
            const f3 dvt = p_i - cp;
            const float dvsq = dot(dvt, dvt);
            if (!pinned_i && dvsq < eff_sq && dvsq > EPS * EPS) {
                const float dl = sqrtf(dvsq);
                if (ut && layer_i > layer_j) {  // behind the inner triangle: left to the untangle pass
                    const f3 tcr = cross(p_v0 - p_j, p_v1 - p_j);
                    const float tl = length(tcr);
                    if (tl > EPS) {
                        const f3 tn = tcr / tl;
                        if (dot(dvt, tn) < 0.0f) {
                            const f3 ni = xyz(T.normals[index]);
                            const float nil = length(ni);
                            if (tri_pin || nil <= 0.5f || dot(ni / fmaxf(nil, EPS), tn) > 0.0f) continue;
                        }
                    }
                }
                const f3 nvt = dvt / dl;
                const float pen = eff - dl;
                if (tri_pin) {
                    add_pin_disp(&acc[index], nvt * pen + tri_move * 0.5f);
                } else {
                    const float wt = bc.x * bc.x * wj + bc.y * bc.y * w0 + bc.z * bc.z * w1;
                    const float wtot = wi + wt;
                    if (wtot > EPS) {
                        const f3 tp = nvt * (pen * 0.6f);
                        add_disp(&acc[index], tp * (wi / wtot));
                        if (wj > 0.0f) add_disp(&acc[j], (-tp) * (bc.x * wj / wtot));
                        if (w0 > 0.0f) add_disp(&acc[v0], (-tp) * (bc.y * w0 / wtot));
                        if (w1 > 0.0f) add_disp(&acc[v1], (-tp) * (bc.z * w1 / wtot));
                    }
                }
            }
        }
    }
}

// sc_solve over the 27 neighbour cells of each vertex.  On a multi-layer cloth with untangling on,
// each free outer-layer vertex also records its nearest inner-layer vertex outside its two-hop
// ring within 2.5 effective thicknesses plus 1.3 local edge lengths (self_collision.wgsl:340-386,
// 678-681): every lane keeps the first minimum of its cell's bucket, and the half-wave reduction
// keeps the minimum by (distance^2, cell in the reference's x-outer order), which is the first
// minimum of the reference's sequential walk.
template <bool UT>
__global__ void __launch_bounds__(SC_BLOCK) k_sc_solve(const float4 *__restrict__ P, const float4 *__restrict__ X,
                                                       SelfColTopo T, const uint32_t *__restrict__ starts,
                                                       const uint32_t *__restrict__ sorted, Accum *__restrict__ acc,
                                                       SelfColArgs A) {
    __shared__ uint2 s_queue[SC_BLOCK / 64][SC_QCAP];
    const uint32_t lane = threadIdx.x & 63u, wv = threadIdx.x >> 6, ci = lane & 31u;
    const float EPS = 1e-7f;
    uint32_t qn = 0;
    for (uint32_t g = 0; g < SC_GROUPS; ++g) {
        const uint32_t index = ((blockIdx.x * (SC_BLOCK / 64) + wv) * SC_GROUPS + g) * 2u + (lane >> 5);
        uint32_t q = 0, qe = 0;
        f3 p_i = mk3(0.0f, 0.0f, 0.0f);
        float thick_i = 0.0f, llen_i = 0.0f;
        i3 nc = {0, 0, 0};
        bool track = false;           // untangle candidate search of this vertex
        uint32_t layer_i = 0u, best_j = SC_NONE;
        float best_dsq = 1e20f;
        if (index < A.nv && ci < 27u) {
            const float4 qi = P[index], xi = X[index];
            p_i = xyz(qi);
            const float mv_i = length(p_i - xyz(xi));
            if (!(qi.w <= 0.0f && mv_i < EPS)) {  // a pinned vertex that did not move takes no part
                thick_i = xi.w;
                llen_i = T.local_len[index];
                nc = neighbour_cell(cell_of(p_i, A.cell), ci);
                const uint32_t h = hash_cell(nc, A.table);
                q = starts[h];
                qe = starts[h + 1];
                if (UT) {
                    layer_i = T.layer[index];
                    track = !(qi.w <= 0.0f) && layer_i > 0u;
                }
            }
        }
        auto accept = [&](uint32_t k) -> uint32_t {
            const uint32_t j = sorted[k];
            if (j >= A.nv || j == index) return SC_NONE;
            const float4 qj = P[j];
            const f3 p_j = xyz(qj);
            const i3 jc = cell_of(p_j, A.cell);
            if (!(jc.x == nc.x && jc.y == nc.y && jc.z == nc.z)) return SC_NONE;
            const float4 xj = X[j];
            const f3 dvv = p_i - p_j;
            const float dsq = dot(dvv, dvv);
            const float thick_j = xj.w, llen_j = T.local_len[j];
            const float eff = fmaxf(thick_i + thick_j, fminf(llen_i, llen_j) * 0.06f);
            if (UT && track && dsq < best_dsq && T.layer[j] < layer_i) {
                const float r = eff * 2.5f + llen_j * 1.3f;
                if (dsq <= r * r && !within_two_hops(T, index, j)) { best_dsq = dsq; best_j = j; }
            }
            const float max_geom = eff + (llen_i + llen_j) * 1.3f;
            return dsq <= max_geom * max_geom ? j : SC_NONE;
        };
        sc_walk(s_queue[wv], qn, g + 1u == SC_GROUPS, lane, index, q, qe, accept,
                [&](uint32_t i, uint32_t j) { vv_pair<UT>(P, X, T, acc, A, i, j); });
        if (UT) {  // the vertex's nearest candidate over its 27 cells (lanes of one half)
            uint32_t kd = __float_as_uint(best_dsq);
            uint32_t ko = ci < 27u ? 9u * (ci % 3u) + 3u * ((ci / 3u) % 3u) + ci / 9u : 31u;
            for (uint32_t o = 16u; o != 0u; o >>= 1) {
                const uint32_t od = __shfl_xor(kd, int(o), 32), oo = __shfl_xor(ko, int(o), 32);
                const uint32_t oj = __shfl_xor(best_j, int(o), 32);
                if (od < kd || (od == kd && oo < ko)) { kd = od; ko = oo; best_j = oj; }
            }
            if (ci == 0u && track && best_j < 0x3FFFFEu) acc[index].untangle = best_j + 1u;
        }
    }
}

// compute_normals.wgsl:25-58: per vertex the normalised sum of (p_a - p_i) x (p_b - p_i) over its
// star pairs at the substep's start positions X; zero for a vertex without triangles.
__global__ void __launch_bounds__(256) k_compute_normals(const float4 *__restrict__ X, SelfColTopo T,
                                                         float4 *__restrict__ normals, uint32_t nv) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nv) return;
    const f3 p_i = xyz(X[i]);
    f3 sum = mk3(0.0f, 0.0f, 0.0f);
    for (uint32_t k = T.star_off[i]; k < T.star_off[i + 1]; ++k) {
        const uint2 pr = T.star[k];
        sum = sum + cross(xyz(X[pr.x]) - p_i, xyz(X[pr.y]) - p_i);
    }
    const float l = length(sum);
    normals[i] = l > 1e-7f ? make_float4(sum.x / l, sum.y / l, sum.z / l, 0.0f) : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

// closest point on a triangle, position only (self_collision_untangle.wgsl:85-129)
__device__ __forceinline__ f3 closest_pos(f3 p, f3 a, f3 b, f3 c) {
    const f3 ab = b - a, ac = c - a, ap = p - a;
    const float d1 = dot(ab, ap), d2 = dot(ac, ap);
    if (d1 <= 0.0f && d2 <= 0.0f) return a;
    const f3 bp = p - b;
    const float d3 = dot(ab, bp), d4 = dot(ac, bp);
    if (d3 >= 0.0f && d4 <= d3) return b;
    const float vc = d1 * d4 - d3 * d2;
    if (vc <= 0.0f && d1 >= 0.0f && d3 <= 0.0f) return a + ab * (d1 / (d1 - d3));
    const f3 cp = p - c;
    const float d5 = dot(ab, cp), d6 = dot(ac, cp);
    if (d6 >= 0.0f && d5 <= d6) return c;
    const float vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0f && d2 >= 0.0f && d6 <= 0.0f) return a + ac * (d2 / (d2 - d6));
    const float va = d3 * d6 - d5 * d4;
    if (va <= 0.0f && (d4 - d3) >= 0.0f && (d5 - d6) >= 0.0f) {
        const float w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        return b + (c - b) * w;
    }
    const float denom = 1.0f / (va + vb + vc);
    const float v = vb * denom, w = vc * denom;
    return a + ab * v + ac * w;
}

// self_collision_untangle.wgsl:163-394: a free outer-layer vertex with a recorded inner candidate j1
// takes j1's nearest 1-ring inner neighbour j2 outside its two-hop ring, picks the nearest triangle
// of j1's and j2's stars (ties: the front side, then the larger height), and if it lies behind it,
// inside a 45-degree cone under the triangle and within half a thickness of its normal line, with
// the vertex normal not opposed (or a pinned triangle) and behind the interpolated pseudo-normal too,
// it is pushed to the front, mass-weighted against the triangle unless the triangle holds a pin.
__global__ void __launch_bounds__(256) k_sc_untangle(const float4 *__restrict__ P, const float4 *__restrict__ X,
                                                     SelfColTopo T, Accum *__restrict__ acc, SelfColArgs A) {
    const uint32_t index = blockIdx.x * blockDim.x + threadIdx.x;
    if (index >= A.nv || !A.untangle) return;
    const float EPS = 1e-7f;
    const uint32_t cand = acc[index].untangle;
    if (cand == 0u) return;
    acc[index].untangle = 0u;
    const float4 qi = P[index];
    const uint32_t layer_i = T.layer[index];
    if (qi.w <= 0.0f || layer_i == 0u) return;
    const uint32_t j1e = cand & 0x3FFFFFu;
    if (j1e == 0u) return;
    const uint32_t j1 = j1e - 1u;
    if (j1 >= A.nv || j1 == index) return;
    const f3 p_i = xyz(qi);
    const float thick_i = X[index].w;
    const float llen_i = T.local_len[index];
    // j2: j1's nearest 1-ring neighbour on an inner layer, outside the vertex's two-hop ring
    uint32_t j2 = SC_NONE;
    float best_nb = 1e20f;
    for (uint32_t a = T.adj_off[j1]; a < T.adj_off[j1 + 1]; ++a) {
        const uint32_t nb = T.adj_idx[a];
        if (nb < A.nv && nb != index && T.layer[nb] < layer_i) {
            const f3 d = p_i - xyz(P[nb]);
            const float dsq = dot(d, d);
            if (dsq < best_nb && !within_two_hops(T, index, nb)) { best_nb = dsq; j2 = nb; }
        }
    }
    float best_dsq = 1e20f, best_h = 0.0f;
    uint32_t bj = SC_NONE, bv0 = 0u, bv1 = 0u;
    for (uint32_t c = 0; c < 2u; ++c) {
        const uint32_t j = c == 0u ? j1 : j2;
        if (j == SC_NONE) continue;
        const float4 qj = P[j];
        const f3 p_j = xyz(qj);
        const float eff = fmaxf(thick_i + X[j].w, fminf(llen_i, T.local_len[j]) * 0.06f);
        const float cap = eff * 2.5f;
        const float cap_sq = cap * cap;
        for (uint32_t st = T.star_off[j]; st < T.star_off[j + 1]; ++st) {
            const uint2 pr = T.star[st];
            const uint32_t v0 = pr.x, v1 = pr.y;
            if (v0 == index || v1 == index) continue;
            if (c == 1u && (v0 == j1 || v1 == j1)) continue;
            const f3 p_v0 = xyz(P[v0]), p_v1 = xyz(P[v1]);
            const f3 tmin = min3(min3(p_j, p_v0), p_v1), tmax = max3(max3(p_j, p_v0), p_v1);
            const f3 z = mk3(0.0f, 0.0f, 0.0f);
            const f3 ad = max3(tmin - p_i, z) + max3(p_i - tmax, z);
            const float lim = fminf(cap_sq, best_dsq + 1e-10f);
            if (dot(ad, ad) >= lim) continue;
            const f3 d = p_i - closest_pos(p_i, p_j, p_v0, p_v1);
            const float dsq = dot(d, d);
            if (dsq >= cap_sq) continue;
            if (bj != SC_NONE && dsq > best_dsq + 1e-10f) continue;
            const f3 tcr = cross(p_v0 - p_j, p_v1 - p_j);
            const float tl = length(tcr);
            if (tl <= EPS) continue;
            const float h = dot(d, tcr / tl);
            const bool same_d = fabsf(dsq - best_dsq) <= 1e-10f;
            const bool better_tie = (h >= 0.0f && best_h < 0.0f) || (((h >= 0.0f) == (best_h >= 0.0f)) && fabsf(h) > fabsf(best_h));
            if (bj == SC_NONE || dsq < best_dsq - 1e-10f || (same_d && better_tie)) {
                if (within_two_hops(T, index, j) || within_two_hops(T, index, v0) || within_two_hops(T, index, v1))
                    continue;
                best_dsq = dsq; best_h = h; bj = j; bv0 = v0; bv1 = v1;
            }
        }
    }
    if (bj == SC_NONE || !(best_h < 0.0f)) return;
    const float4 qj = P[bj], q0 = P[bv0], q1 = P[bv1];
    const f3 p_j = xyz(qj), p_v0 = xyz(q0), p_v1 = xyz(q1);
    f3 bc;
    const f3 d = p_i - closest_bary(p_i, p_j, p_v0, p_v1, bc);
    const float dist = length(d);
    const f3 tcr = cross(p_v0 - p_j, p_v1 - p_j);
    const float tl = length(tcr);
    if (!(tl > EPS && dist > EPS)) return;
    const f3 tn = tcr / tl;
    const float h = dot(d, tn);
    if (!(h < 0.0f)) return;
    const f3 ni = xyz(T.normals[index]);
    const float nil = length(ni);
    const bool tri_pin = qj.w <= 0.0f || q0.w <= 0.0f || q1.w <= 0.0f;
    const bool allow = tri_pin || nil <= 0.5f || dot(ni / fmaxf(nil, EPS), tn) > 0.0f;
    const float eff = fmaxf(thick_i + X[bj].w, fminf(llen_i, T.local_len[bj]) * 0.06f);
    const float cos_behind = -h / dist;
    const float r_tan = sqrtf(fmaxf(dist * dist - h * h, 0.0f));
    const float r_lim = fmaxf(eff * 0.5f, EPS);
    if (!(allow && cos_behind > 0.707f && r_tan < r_lim)) return;
    const f3 nint = xyz(T.normals[bj]) * bc.x + xyz(T.normals[bv0]) * bc.y + xyz(T.normals[bv1]) * bc.z;
    const float nl = length(nint);
    if (!(nl <= 0.5f || dot(d, nint) < 0.0f)) return;
    const float cone_w = clampf((cos_behind - 0.707f) / 0.2f, 0.0f, 1.0f);
    const float r_w = clampf(1.0f - r_tan / r_lim, 0.0f, 1.0f);
    const float pen = fminf(eff - h, eff * 2.0f) * cone_w * r_w;
    if (!(pen > EPS)) return;
    if (tri_pin) {
        add_disp(&acc[index], tn * pen);
        return;
    }
    const float wi = qi.w, wj = qj.w, w0 = q0.w, w1 = q1.w;
    const float wt = bc.x * bc.x * wj + bc.y * bc.y * w0 + bc.z * bc.z * w1;
    const float wtot = wi + wt;
    if (!(wtot > EPS)) return;
    const f3 disp = tn * (pen * 0.5f);
    add_disp(&acc[index], disp * (wi / wtot));
    if (wj > 0.0f) add_disp(&acc[bj], (-disp) * (bc.x * wj / wtot));
    if (w0 > 0.0f) add_disp(&acc[bv0], (-disp) * (bc.y * w0 / wtot));
    if (w1 > 0.0f) add_disp(&acc[bv1], (-disp) * (bc.z * w1 / wtot));
}

// ---------------------------------------------------------------------------
// E-E (self_collision_ee_direct.wgsl)
struct ST { float s, t; };
__device__ __forceinline__ ST closest_segs(f3 p1, f3 q1, f3 p2, f3 q2) {  // ee_direct.wgsl:155-193
    const f3 d1 = q1 - p1, d2 = q2 - p2, r = p1 - p2;
    const float a = dot(d1, d1), e = dot(d2, d2), f = dot(d2, r);
    if (a <= 1e-12f && e <= 1e-12f) return {0.0f, 0.0f};
    if (a <= 1e-12f) return {0.0f, clampf(f / e, 0.0f, 1.0f)};
    const float c = dot(d1, r);
    if (e <= 1e-12f) return {clampf(-c / a, 0.0f, 1.0f), 0.0f};
    const float b = dot(d1, d2);
    const float denom = a * e - b * b;
    float s = 0.0f;
    if (denom > a * e * 1e-7f) s = clampf((b * f - c * e) / denom, 0.0f, 1.0f);
    float t = (b * s + f) / e;
    if (t < 0.0f) { t = 0.0f; s = clampf(-c / a, 0.0f, 1.0f); }
    else if (t > 1.0f) { t = 1.0f; s = clampf((b - c) / a, 0.0f, 1.0f); }
    return {s, t};
}
struct Sweep { f3 a0o, a0, a1o, a1, b0o, b0, b1o, b1; };
__device__ __forceinline__ float sweep_dist_sq(const Sweep &w, float tau, ST &st) {
    const f3 A0 = mix3(w.a0o, w.a0, tau), A1 = mix3(w.a1o, w.a1, tau);
    const f3 B0 = mix3(w.b0o, w.b0, tau), B1 = mix3(w.b1o, w.b1, tau);
    st = closest_segs(A0, A1, B0, B1);
    const f3 d = mix3(A0, A1, st.s) - mix3(B0, B1, st.t);
    return dot(d, d);
}
__device__ __forceinline__ float sweep_volume(const Sweep &w, float tau) {
    const f3 A0 = mix3(w.a0o, w.a0, tau), A1 = mix3(w.a1o, w.a1, tau);
    const f3 B0 = mix3(w.b0o, w.b0, tau), B1 = mix3(w.b1o, w.b1, tau);
    return dot(cross(A1 - A0, B1 - B0), B0 - A0);
}
__device__ __forceinline__ bool edge_bit(const uint32_t *bits, uint32_t e) { return ((bits[e >> 5] >> (e & 31u)) & 1u) != 0u; }

// is_outer_edge_behind_inner (self_collision_ee_direct.wgsl:123-148): for edges of different layers,
// whether the outer one lies behind the inner one along the inner edge's summed vertex normals, with
// the outer normals not opposed (or the inner edge fixed); delta_ab points from edge b to edge a.
__device__ __forceinline__ bool outer_edge_behind_inner(const SelfColTopo &T, uint32_t la, uint32_t lb, uint32_t a0,
                                                        uint32_t a1, uint32_t b0, uint32_t b1, bool a_fixed,
                                                        bool b_fixed, f3 delta_ab) {
    const float EPS = 1e-7f;
    if (la == lb) return false;
    const f3 na = xyz(T.normals[a0]) + xyz(T.normals[a1]), nb = xyz(T.normals[b0]) + xyz(T.normals[b1]);
    const float nal = length(na), nbl = length(nb);
    if (la > lb) {
        if (nbl <= 0.5f) return false;
        const f3 ni = nb / nbl;
        const bool allow = b_fixed || nal <= 0.5f || dot(na / fmaxf(nal, EPS), ni) > 0.0f;
        return allow && dot(delta_ab, ni) < 0.0f;
    }
    if (nal <= 0.5f) return false;
    const f3 ni = na / nal;
    const bool allow = a_fixed || nbl <= 0.5f || dot(nb / fmaxf(nbl, EPS), ni) > 0.0f;
    return allow && dot(-delta_ab, ni) < 0.0f;
}
// The E-E CCD's untangling exemption (self_collision_ee_direct.wgsl:506-524, solve_ee.wgsl:386-408):
// an outer edge (not fixed) that starts behind the inner edge's normal and moves to its front.
__device__ __forceinline__ bool ee_ccd_untangling(const SelfColTopo &T, uint32_t la, uint32_t lb, uint32_t u0,
                                                  uint32_t j, bool a_fixed, bool b_fixed, f3 a0o, f3 a1o, f3 b0o,
                                                  f3 b1o, f3 a0, f3 a1, f3 b0, f3 b1) {
    if (la == lb) return false;
    const f3 d0 = ((a0o + a1o) - (b0o + b1o)) * 0.5f;
    const f3 d1 = ((a0 + a1) - (b0 + b1)) * 0.5f;
    const f3 mv = d1 - d0;
    if (la > lb && !a_fixed) {
        const f3 ni = xyz(T.normals[j]), no = xyz(T.normals[u0]);
        const bool allow = b_fixed || dot(no, no) <= 0.25f;
        return allow && dot(ni, ni) > 0.25f && dot(d0, ni) < 0.0f && dot(mv, ni) > 0.0f;
    }
    if (lb > la && !b_fixed) {
        const f3 ni = xyz(T.normals[u0]), no = xyz(T.normals[j]);
        const bool allow = a_fixed || dot(no, no) <= 0.25f;
        return allow && dot(ni, ni) > 0.25f && dot(-d0, ni) < 0.0f && dot(-mv, ni) > 0.0f;
    }
    return false;
}

// E-E response of edge ea against edge eb for a pair that passed the box test
// (self_collision_ee_direct.wgsl:266-776, the candidate loop's body after the
// search counter).
__device__ __forceinline__ void ee_pair(const float4 *__restrict__ P, const float4 *__restrict__ X,
                                        const SelfColTopo &T, Accum *__restrict__ acc, const SelfColArgs &A,
                                        uint32_t ea, uint32_t eb) {
    const float EPS = 1e-7f;
    const uint2 eA = T.edges[ea];
    const uint32_t u0 = eA.x, u1 = eA.y;
    const float4 qu0 = P[u0], qu1 = P[u1], xu0 = X[u0], xu1 = X[u1];
    const float wu0 = qu0.w, wu1 = qu1.w;
    const bool a_has_pin = wu0 <= 0.0f || wu1 <= 0.0f;
    const f3 p_u0 = xyz(qu0), p_u1 = xyz(qu1), p_u0o = xyz(xu0), p_u1o = xyz(xu1);
    const float th_u0 = xu0.w, th_u1 = xu1.w;
    const float max_th_a = fmaxf(th_u0, th_u1);
    const bool bnd_a = edge_bit(T.edge_bbits, ea);
    const uint2 eB = T.edges[eb];
    const uint32_t j = eB.x, vj = eB.y;
    const float4 qj = P[j], qvj = P[vj], xj = X[j], xvj = X[vj];
    const f3 p_j = xyz(qj), p_vj = xyz(qvj), p_jo = xyz(xj), p_vjo = xyz(xvj);
    const bool bnd_pair = bnd_a || edge_bit(T.edge_bbits, eb);
    bool run_ccd = bnd_pair, has_cached = false;
    ST st1c = {0.0f, 0.0f};
    f3 d1c = mk3(0.0f, 0.0f, 0.0f);
    float dsqc = 0.0f, effc = 0.0f;
    const float th_j = xj.w, th_vj = xvj.w;
    const float wj = qj.w, wvj = qvj.w;
    if (!run_ccd) {
        st1c = closest_segs(p_u0, p_u1, p_j, p_vj);
        d1c = mix3(p_u0, p_u1, st1c.s) - mix3(p_j, p_vj, st1c.t);
        dsqc = dot(d1c, d1c);
        effc = mixf(th_u0, th_u1, st1c.s) + mixf(th_j, th_vj, st1c.t);
        const float max_eff = max_th_a + fmaxf(th_j, th_vj);
        const f3 du0 = p_u0 - p_u0o, du1 = p_u1 - p_u1o, dj = p_j - p_jo, dvj = p_vj - p_vjo;
        const f3 va = (du0 + du1) * 0.5f, vb = (dj + dvj) * 0.5f;
        const float rel = length(va - vb) + 0.5f * (length(du1 - du0) + length(dvj - dj));
        const float reach = max_eff + rel;
        if (dsqc > reach * reach) return;
        has_cached = true;
        const f3 d0g = mix3(p_u0o, p_u1o, st1c.s) - mix3(p_jo, p_vjo, st1c.t);
        run_ccd = (rel > effc * 0.5f) || (dot(d0g, d1c) <= 0.0f);
    }
    if (run_ccd) {
        const Sweep W = {p_u0o, p_u0, p_u1o, p_u1, p_jo, p_j, p_vjo, p_vj};
        float min_sq = 1e20f, best_tau = 0.0f, half_span = 0.125f;
        ST min_st = {0.0f, 0.0f};
        float prev_vol = sweep_volume(W, 0.0f);
        float zc_tau = -1.0f;
#pragma unroll 1  // unrolled, the sweeps need 215 VGPRs (2 wavefronts per SIMD)
        for (uint32_t k = 0; k <= 8u; ++k) {
            const float tk = float(k) * 0.125f;
            ST stk;
            const float dk = sweep_dist_sq(W, tk, stk);
            if (dk < min_sq) { min_sq = dk; best_tau = tk; min_st = stk; }
            if (k > 0u) {
                const float vol = sweep_volume(W, tk);
                if (prev_vol * vol <= 0.0f && zc_tau < 0.0f) zc_tau = tk - 0.0625f;
                prev_vol = vol;
            }
        }
        if (zc_tau >= 0.0f) {
            ST stz;
            const float dzs = sweep_dist_sq(W, zc_tau, stz);
            if (dzs < min_sq) { min_sq = dzs; best_tau = zc_tau; half_span = 0.0625f; min_st = stz; }
        }
        float tlo = fmaxf(0.0f, best_tau - half_span), thi = fminf(1.0f, best_tau + half_span);
        float tau_star = best_tau, dsq_star = min_sq;
        ST st_star = min_st;
#pragma unroll 1
        for (int it = 0; it < 6; ++it) {
            const float m1 = tlo + (thi - tlo) * (1.0f / 3.0f);
            const float m2 = thi - (thi - tlo) * (1.0f / 3.0f);
            ST s1, s2;
            const float e1 = sweep_dist_sq(W, m1, s1);
            const float e2 = sweep_dist_sq(W, m2, s2);
            if (e1 < dsq_star) { dsq_star = e1; tau_star = m1; st_star = s1; }
            if (e2 < dsq_star) { dsq_star = e2; tau_star = m2; st_star = s2; }
            if (e1 < e2) thi = m2; else tlo = m1;
        }
        const float ss = st_star.s, ts = st_star.t;
        const float eff_star = mixf(th_u0, th_u1, ss) + mixf(th_j, th_vj, ts);
        if (dsq_star >= eff_star * eff_star) return;
        if (A.exclude_neighbors) {
            const f3 a0s = mix3(p_u0o, p_u0, tau_star), a1s = mix3(p_u1o, p_u1, tau_star);
            const f3 b0s = mix3(p_jo, p_j, tau_star), b1s = mix3(p_vjo, p_vj, tau_star);
            const f3 t0 = a0s - b0s, t1 = a0s - b1s, t2 = a1s - b0s, t3 = a1s - b1s;
            if (near_folded_sq(T, u0, j, dot(t0, t0)) || near_folded_sq(T, u0, vj, dot(t1, t1)) ||
                near_folded_sq(T, u1, j, dot(t2, t2)) || near_folded_sq(T, u1, vj, dot(t3, t3)))
                return;
        }
        if (A.untangle &&
            ee_ccd_untangling(T, T.layer[u0], T.layer[j], u0, j, wu0 == 0.0f && wu1 == 0.0f, wj == 0.0f && wvj == 0.0f,
                              p_u0o, p_u1o, p_jo, p_vjo, p_u0, p_u1, p_j, p_vj))
            return;
        const ST st0 = closest_segs(p_u0o, p_u1o, p_jo, p_vjo);
        const f3 delta0 = mix3(p_u0o, p_u1o, st0.s) - mix3(p_jo, p_vjo, st0.t);
        const float dist0 = length(delta0);
        if (dist0 > EPS) {
            const f3 n0 = delta0 / dist0;
            const ST st1 = closest_segs(p_u0, p_u1, p_j, p_vj);
            const f3 delta1 = mix3(p_u0, p_u1, st1.s) - mix3(p_j, p_vj, st1.t);
            const float dist1 = length(delta1);
            const float h_end = dot(mix3(p_u0, p_u1, ss) - mix3(p_j, p_vj, ts), n0);
            const bool crossover = (h_end <= 0.0f) || (dot(delta1, n0) <= 0.0f) ||
                                   (bnd_pair && dist1 >= eff_star && tau_star < 1.0f);
            if (h_end < eff_star && crossover) {
                const float pen = eff_star - h_end;
                const bool b_has_pin = wj <= 0.0f || wvj <= 0.0f;
                if (a_has_pin && b_has_pin) return;
                const bool a_all = wu0 == 0.0f && wu1 == 0.0f;
                const bool b_all = wj == 0.0f && wvj == 0.0f;
                const bool cnt_pin = !(a_all || b_all);
                if (b_has_pin) {
                    const f3 em = ((p_j - p_jo) + (p_vj - p_vjo)) * 0.5f;
                    const float a = 1.0f - ss, b = ss;
                    const float den = a * a * wu0 + b * b * wu1;
                    if (den > EPS) {
                        const f3 pb = n0 * pen + em * 0.5f;
                        if (wu0 > 0.0f) add_ccd_ee(&acc[u0], pb * fminf((a * wu0) / den, 2.0f), cnt_pin);
                        if (wu1 > 0.0f) add_ccd_ee(&acc[u1], pb * fminf((b * wu1) / den, 2.0f), cnt_pin);
                    }
                } else if (a_has_pin) {
                    const f3 em = ((p_u0 - p_u0o) + (p_u1 - p_u1o)) * 0.5f;
                    const float a = 1.0f - ts, b = ts;
                    const float den = a * a * wj + b * b * wvj;
                    if (den > EPS) {
                        const f3 pb = (-n0) * pen + em * 0.5f;
                        if (wj > 0.0f) add_ccd_ee(&acc[j], pb * fminf((a * wj) / den, 2.0f), cnt_pin);
                        if (wvj > 0.0f) add_ccd_ee(&acc[vj], pb * fminf((b * wvj) / den, 2.0f), cnt_pin);
                    }
                } else {
                    const float we1 = (1.0f - ss) * (1.0f - ss) * wu0 + ss * ss * wu1;
                    const float we2 = (1.0f - ts) * (1.0f - ts) * wj + ts * ts * wvj;
                    const float wt = we1 + we2;
                    if (wt > EPS) {
                        const f3 tot = n0 * pen;
                        if (wu0 > 0.0f) add_ccd_ee(&acc[u0], tot * ((1.0f - ss) * wu0 / wt), true);
                        if (wu1 > 0.0f) add_ccd_ee(&acc[u1], tot * (ss * wu1 / wt), true);
                        if (wj > 0.0f) add_ccd_ee(&acc[j], (-tot) * ((1.0f - ts) * wj / wt), true);
                        if (wvj > 0.0f) add_ccd_ee(&acc[vj], (-tot) * (ts * wvj / wt), true);
                    }
                }
                return;
            }
        }
    }
    // discrete proximity
    ST st;
    f3 delta_ee;
    float dsq, eff;
    if (has_cached) {
        st = st1c; delta_ee = d1c; dsq = dsqc; eff = effc;
    } else {
        st = closest_segs(p_u0, p_u1, p_j, p_vj);
        delta_ee = (p_u0 + (p_u1 - p_u0) * st.s) - (p_j + (p_vj - p_j) * st.t);
        dsq = dot(delta_ee, delta_ee);
        eff = mixf(th_u0, th_u1, st.s) + mixf(th_j, th_vj, st.t);
    }
    if (dsq >= eff * eff || dsq <= EPS * EPS) return;
    if (A.exclude_neighbors) {
        const f3 d0 = p_u0 - p_j, d1 = p_u0 - p_vj, d2 = p_u1 - p_j, d3 = p_u1 - p_vj;
        if (near_folded_sq(T, u0, j, dot(d0, d0)) || near_folded_sq(T, u0, vj, dot(d1, d1)) ||
            near_folded_sq(T, u1, j, dot(d2, d2)) || near_folded_sq(T, u1, vj, dot(d3, d3)))
            return;
    }
    const float dist = sqrtf(dsq);
    if (A.untangle && outer_edge_behind_inner(T, T.layer[u0], T.layer[j], u0, u1, j, vj, wu0 <= 0.0f && wu1 <= 0.0f,
                                              wj <= 0.0f && wvj <= 0.0f, delta_ee))
        return;
    const f3 n = delta_ee / dist;
    const float pen = eff - dist;
    const float sp = st.s, tp = st.t;
    const bool b_has_pin = wj <= 0.0f || wvj <= 0.0f;
    if (a_has_pin && b_has_pin) return;
    if (b_has_pin) {
        const f3 em = ((p_j - p_jo) + (p_vj - p_vjo)) * 0.5f;
        const float a = 1.0f - sp, b = sp;
        const float den = a * a * wu0 + b * b * wu1;
        if (den > EPS) {
            const f3 pb = n * pen + em * 0.5f;
            if (wu0 > 0.0f) add_pin_disp(&acc[u0], pb * fminf((a * wu0) / den, 2.0f));
            if (wu1 > 0.0f) add_pin_disp(&acc[u1], pb * fminf((b * wu1) / den, 2.0f));
        }
    } else if (a_has_pin) {
        const f3 em = ((p_u0 - p_u0o) + (p_u1 - p_u1o)) * 0.5f;
        const float a = 1.0f - tp, b = tp;
        const float den = a * a * wj + b * b * wvj;
        if (den > EPS) {
            const f3 pb = (-n) * pen + em * 0.5f;
            if (wj > 0.0f) add_pin_disp(&acc[j], pb * fminf((a * wj) / den, 2.0f));
            if (wvj > 0.0f) add_pin_disp(&acc[vj], pb * fminf((b * wvj) / den, 2.0f));
        }
    } else {
        const float we1 = (1.0f - sp) * (1.0f - sp) * wu0 + sp * sp * wu1;
        const float we2 = (1.0f - tp) * (1.0f - tp) * wj + tp * tp * wvj;
        const float wt = we1 + we2;
        if (wt > EPS) {
            const f3 td = n * (pen * 0.6f);
            if (wu0 > 0.0f) add_disp_ee(&acc[u0], td * ((1.0f - sp) * wu0 / wt));
            if (wu1 > 0.0f) add_disp_ee(&acc[u1], td * (sp * wu1 / wt));
            if (wj > 0.0f) add_disp_ee(&acc[j], (-td) * ((1.0f - tp) * wj / wt));
            if (wvj > 0.0f) add_disp_ee(&acc[vj], (-td) * (tp * wvj / wt));
        }
    }
}

__global__ void __launch_bounds__(SC_BLOCK, 4) k_sc_solve_ee(const float4 *__restrict__ P, const float4 *__restrict__ X,
                                                          SelfColTopo T, const uint32_t *__restrict__ starts,
                                                          const uint32_t *__restrict__ sorted,
                                                          Accum *__restrict__ acc, SelfColArgs A) {
    __shared__ uint2 s_queue[SC_BLOCK / 64][SC_QCAP];
    const uint32_t lane = threadIdx.x & 63u, wv = threadIdx.x >> 6, ci = lane & 31u;
    const uint32_t cap = A.max_search_iterations;
    uint32_t qn = 0;
    for (uint32_t g = 0; g < SC_GROUPS; ++g) {
        const uint32_t ea = ((blockIdx.x * (SC_BLOCK / 64) + wv) * SC_GROUPS + g) * 2u + (lane >> 5);
        uint32_t q = 0, qe = 0, u0 = 0, u1 = 0;
        f3 amin = mk3(0.0f, 0.0f, 0.0f), amax = amin;
        i3 nc = {0, 0, 0};
        if (ea < A.ne && ci < 27u) {
            const uint2 eA = T.edges[ea];
            u0 = eA.x;
            u1 = eA.y;
            const float4 qu0 = P[u0], qu1 = P[u1], xu0 = X[u0], xu1 = X[u1];
            const f3 p_u0 = xyz(qu0), p_u1 = xyz(qu1), p_u0o = xyz(xu0), p_u1o = xyz(xu1);
            const float max_th_a = fmaxf(xu0.w, xu1.w);
            const f3 mta = mk3(max_th_a * 2.0f, max_th_a * 2.0f, max_th_a * 2.0f);
            amin = min3(min3(p_u0, p_u1), min3(p_u0o, p_u1o)) - mta;
            amax = max3(max3(p_u0, p_u1), max3(p_u0o, p_u1o)) + mta;
            nc = neighbour_cell(cell_of((p_u0 + p_u1) * 0.5f, A.cell), ci);
            const uint32_t h = hash_cell(nc, A.table);
            q = starts[h];
            qe = starts[h + 1];
        }
        // candidate tests ahead of the reference's search counter
        auto boxed = [&](uint32_t eb) -> bool {
            if (ea >= eb) return false;
            const uint2 eB = T.edges[eb];
            const uint32_t j = eB.x, vj = eB.y;
            if (j == u0 || j == u1 || vj == u0 || vj == u1) return false;
            const float4 qj = P[j], qvj = P[vj];
            const f3 p_j = xyz(qj), p_vj = xyz(qvj);
            const i3 bcell = cell_of((p_j + p_vj) * 0.5f, A.cell);
            if (!(bcell.x == nc.x && bcell.y == nc.y && bcell.z == nc.z)) return false;
            const float4 xj = X[j], xvj = X[vj];
            const f3 p_jo = xyz(xj), p_vjo = xyz(xvj);
            const f3 bmin = min3(min3(p_j, p_vj), min3(p_jo, p_vjo));
            const f3 bmax = max3(max3(p_j, p_vj), max3(p_jo, p_vjo));
            return !(amin.x > bmax.x || amin.y > bmax.y || amin.z > bmax.z || amax.x < bmin.x ||
                     amax.y < bmin.y || amax.z < bmin.z);
        };
        // The reference ends an edge's search once max_search_iterations candidates
        // have passed the box test, counting through the cells in order.  When the 27
        // buckets of an edge hold more entries than that, each lane first counts its
        // cell's passes so that it knows how many the cells before it contribute.
        uint32_t ord = 0;  // box-test passes of this edge before the current candidate
        uint32_t n = qe - q;
        for (uint32_t o = 16u; o != 0u; o >>= 1) n += __shfl_xor(n, int(o), 32);
        if (__any(n > cap)) {
            uint32_t cnt = 0;
            for (uint32_t k = q; k < qe; ++k) cnt += boxed(sorted[k]) ? 1u : 0u;
            uint32_t incl = cnt;
            for (uint32_t o = 1u; o < 32u; o <<= 1) {
                const uint32_t y = __shfl_up(incl, o, 32);
                if (ci >= o) incl += y;
            }
            ord = incl - cnt;
        }
        auto accept = [&](uint32_t k) -> uint32_t {
            const uint32_t eb = sorted[k];
            if (!boxed(eb)) return SC_NONE;
            return ++ord <= cap ? eb : SC_STOP;
        };
        sc_walk(s_queue[wv], qn, g + 1u == SC_GROUPS, lane, ea, q, qe, accept,
                [&](uint32_t a, uint32_t b) { ee_pair(P, X, T, acc, A, a, b); });
    }
}

// self_collision_apply.wgsl:57-177; clears the accumulator it reads.
__global__ void __launch_bounds__(256) k_sc_apply(float4 *__restrict__ P, const float *__restrict__ local_len,
                                                  Accum *__restrict__ acc, SelfColArgs A) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= A.nv) return;
    const float EPS = 1e-7f;
    const float INV = 1.0f / 1000000.0f;
    Accum a = acc[i];
    const float4 q = P[i];
    const bool pinned = q.w <= 0.0f;
    f3 fin = mk3(0.0f, 0.0f, 0.0f);
    if (a.count > 0u && !pinned) {
        const f3 tot = mk3(float(a.dx), float(a.dy), float(a.dz)) * INV;
        f3 eff = tot / fmaxf(1.0f, sqrtf(float(a.count)));
        if (A.enable_relief) eff = eff * A.relief_factor;
        const float pl = length(eff);
        const float ratio = A.max_disp_ratio > 0.0f ? A.max_disp_ratio : 0.5f;
        const float ms = fminf(local_len[i] * ratio, 0.10f);
        if (pl > ms && pl > EPS) eff = (eff / pl) * ms;
        fin = fin + eff;
    }
    f3 avg_pin = mk3(0.0f, 0.0f, 0.0f);
    if (a.pcount > 0u && !pinned) avg_pin = (mk3(float(a.pdx), float(a.pdy), float(a.pdz)) * INV) / float(a.pcount);
    if (!pinned) {
        f3 env = mk3(float(a.cpx), float(a.cpy), float(a.cpz)) * INV + mk3(float(a.cnx), float(a.cny), float(a.cnz)) * INV;
        if (a.ee_count >= 4u) env = env * fmaxf(0.5f, sqrtf(3.0f / float(a.ee_count)));
        const f3 z = mk3(0.0f, 0.0f, 0.0f);
        fin = fin + (max3(max3(env, z), max3(avg_pin, z)) + min3(min3(env, z), min3(avg_pin, z)));
    }
    if (length(fin) > EPS) P[i] = make_float4(q.x + fin.x, q.y + fin.y, q.z + fin.z, q.w);
    acc[i] = Accum{};
}

hipError_t launch_sc_solve(hipStream_t s, const float4 *P, const float4 *X, const SelfColTopo &t, const HashBufs &vh,
                           Accum *acc, const SelfColArgs &a) {
    if (!a.nv) return hipSuccess;
    if (a.untangle)
        hipLaunchKernelGGL(k_sc_solve<true>, dim3(grid_for(a.nv, SC_PER_BLOCK)), dim3(SC_BLOCK), 0, s, P, X, t,
                           vh.starts, vh.sorted, acc, a);
    else
        hipLaunchKernelGGL(k_sc_solve<false>, dim3(grid_for(a.nv, SC_PER_BLOCK)), dim3(SC_BLOCK), 0, s, P, X, t,
                           vh.starts, vh.sorted, acc, a);
    return hipGetLastError();
}

hipError_t launch_sc_solve_ee(hipStream_t s, const float4 *P, const float4 *X, const SelfColTopo &t,
                              const HashBufs &eh, Accum *acc, const SelfColArgs &a) {
    if (!a.ne) return hipSuccess;
    hipLaunchKernelGGL(k_sc_solve_ee, dim3(grid_for(a.ne, SC_PER_BLOCK)), dim3(SC_BLOCK), 0, s, P, X, t, eh.starts,
                       eh.sorted, acc, a);
    return hipGetLastError();
}

hipError_t launch_sc_apply(hipStream_t s, float4 *P, const float *local_len, Accum *acc, const SelfColArgs &a,
                           uint32_t block) {
    if (!a.nv) return hipSuccess;
    hipLaunchKernelGGL(k_sc_apply, dim3(grid_for(a.nv, block)), dim3(block), 0, s, P, local_len, acc, a);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Pair cache (self_collision_algorithm 1): a broadphase that keeps, per vertex, the `quota`
// closest V-T triples and E-E quadruples (self_collision_collect_pairs.wgsl:294-661), and
// narrowphase kernels that resolve the cached pairs with the current positions
// (self_collision_solve_vt.wgsl:227-414, self_collision_solve_ee.wgsl:231-589).  The slots are
// vertex-major (vertex i owns slots [i * quota, (i + 1) * quota)); unused slots hold the sentinel
// 0xFFFFFFFF in their first word.  One thread per vertex walks the 27 cells in the reference's
// order (x outer, z inner; buckets in ascending index), so the evaluation counters match too; the
// per-thread top-K lists are unrolled so they stay in registers.

constexpr uint32_t PAIR_SENTINEL = 0xFFFFFFFFu;

// self_collision_collect_pairs.wgsl:122-156: compares the distance itself (not its square)
__device__ __forceinline__ bool near_folded_dist(const SelfColTopo &t, uint32_t a, uint32_t b, float cur) {
    const int32_t s = two_hop_slot(t, a, b);
    if (s < 0) return false;
    return !(cur < t.th_rest[s] * 0.5f);
}
__device__ __forceinline__ int32_t iabs_wrap(int32_t d) { return d < 0 ? int32_t(0u - uint32_t(d)) : d; }
__device__ __forceinline__ int32_t isub_wrap(int32_t a, int32_t b) { return int32_t(uint32_t(a) - uint32_t(b)); }
__device__ __forceinline__ float pair_horizon(const float4 *__restrict__ V, uint32_t v, const PairArgs &A) {
    const float4 u = V[v];
    return fminf(length(mk3(u.x, u.y, u.z)) * A.dt_frame * A.horizon_scale, A.max_horizon);
}

// can_ee_anchor_reach (self_collision_collect_pairs.wgsl:250-279)
__device__ __forceinline__ bool ee_anchor_reach(const float4 *__restrict__ P, const float4 *__restrict__ X,
                                                const float4 *__restrict__ V, const SelfColTopo &T,
                                                const PairTopo &PT, const PairArgs &A, uint32_t u, uint32_t w) {
    const float4 qu = P[u], qw = P[w];
    if (qu.w <= 0.0f) return false;
    const f3 pu = xyz(qu), pw = xyz(qw);
    const i3 cu = cell_of(pu, A.cell), cw = cell_of(pw, A.cell);
    if (iabs_wrap(isub_wrap(cu.x, cw.x)) > 1 || iabs_wrap(isub_wrap(cu.y, cw.y)) > 1 ||
        iabs_wrap(isub_wrap(cu.z, cw.z)) > 1)
        return false;
    const f3 d = pu - pw;
    const float dsq = dot(d, d);
    const float hu = pair_horizon(V, u, A) * A.mode_f, hw = pair_horizon(V, w, A) * A.mode_f;
    const float bound = X[u].w + X[w].w + (T.local_len[u] + T.local_len[w]) * 1.3f + A.safety_margin + hu + hw;
    if (dsq > bound * bound) return false;
    if (PT.island[u] == PT.island[w] && A.exclude_neighbors) {
        if (near_folded_dist(T, u, w, sqrtf(fmaxf(dsq, 0.0f)))) return false;
    }
    return true;
}
__device__ __forceinline__ bool smaller_anchor(const float4 *__restrict__ P, const float4 *__restrict__ X,
                                               const float4 *__restrict__ V, const SelfColTopo &T,
                                               const PairTopo &PT, const PairArgs &A, uint32_t u, uint32_t w,
                                               uint32_t cu, uint32_t cw) {
    if (u > cu || (u == cu && w >= cw)) return false;
    return ee_anchor_reach(P, X, V, T, PT, A, u, w);
}

// Vertex-local top-K list ordered by (distance^2, key), keys compared lexicographically.
struct TopK {
    float d[8];
    uint4 e[8];
    uint32_t n;
};
// Inserts (dist, entry) unless an entry with the same key (words y, z, w) is held; counts the
// evaluation and, when the list is full of closer entries, the drop (collect_pairs.wgsl:427-476).
__device__ __forceinline__ void topk_offer(TopK &L, uint32_t quota, float dist, uint4 ent, uint32_t &eval,
                                           uint32_t &drop) {
    bool dup = false;
#pragma unroll
    for (uint32_t q = 0; q < 8u; ++q)
        if (q < L.n && L.e[q].y == ent.y && L.e[q].z == ent.z && L.e[q].w == ent.w) dup = true;
    if (dup) return;
    ++eval;
    uint32_t pos = L.n;
#pragma unroll
    for (int q = 7; q >= 0; --q) {
        if (uint32_t(q) < L.n) {
            const uint4 o = L.e[q];
            const bool tie_lt = dist == L.d[q] &&
                                (ent.y < o.y || (ent.y == o.y && (ent.z < o.z || (ent.z == o.z && ent.w < o.w))));
            if (dist < L.d[q] || tie_lt) pos = uint32_t(q);  // the first such slot wins: scanned from the top
        }
    }
    if (pos >= quota) { ++drop; return; }
    const uint32_t up = min(L.n, quota - 1u);
#pragma unroll
    for (int q = 7; q >= 1; --q) {
        if (uint32_t(q) <= up && uint32_t(q) > pos) { L.d[q] = L.d[q - 1]; L.e[q] = L.e[q - 1]; }
    }
#pragma unroll
    for (uint32_t q = 0; q < 8u; ++q) {
        if (q == pos) { L.d[q] = dist; L.e[q] = ent; }
    }
    if (L.n < quota) ++L.n;
}

// One list per launch (VT: the V-T triples, else the E-E quadruples): both walk the same cells
// and candidates, and one list's state per thread keeps the top-K in registers.
template <bool VT>
__global__ void __launch_bounds__(64) k_pair_collect(const float4 *__restrict__ P, const float4 *__restrict__ X,
                                                     const float4 *__restrict__ V, SelfColTopo T, PairTopo PT,
                                                     const uint32_t *__restrict__ starts,
                                                     const uint32_t *__restrict__ sorted, uint4 *__restrict__ out,
                                                     uint32_t *__restrict__ counters, PairArgs A) {
    const uint32_t index = blockIdx.x * blockDim.x + threadIdx.x;
    if (index >= A.nv) return;
    const float EPS = 1e-7f;
    const float4 qi = P[index];
    const bool pinned_i = qi.w <= 0.0f;
    const f3 p_i = xyz(qi);
    const float thick_i = X[index].w, lli = T.local_len[index];
    const i3 base = cell_of(p_i, A.cell);
    TopK L;
    L.n = 0;
    uint32_t eval = 0, drop = 0;
    const uint32_t quota = VT ? A.quota_vt : A.quota_ee;
    const float hi_eff = pair_horizon(V, index, A) * A.mode_f;
    const uint32_t island_i = PT.island[index];
    for (int gx = -1; gx <= 1; ++gx)
        for (int gy = -1; gy <= 1; ++gy)
            for (int gz = -1; gz <= 1; ++gz) {
                const i3 cc = {int32_t(uint32_t(base.x) + uint32_t(gx)), int32_t(uint32_t(base.y) + uint32_t(gy)),
                               int32_t(uint32_t(base.z) + uint32_t(gz))};
                const uint32_t h = hash_cell(cc, A.table);
                const uint32_t start = starts[h], end = starts[h + 1];
                for (uint32_t k = start; k < end; ++k) {
                    const uint32_t j = sorted[k];
                    if (j >= A.nv || j == index) continue;
                    const float4 qj = P[j];
                    const f3 p_j = xyz(qj);
                    const i3 jc = cell_of(p_j, A.cell);
                    if (!(jc.x == cc.x && jc.y == cc.y && jc.z == cc.z)) continue;
                    const bool pinned_j = qj.w <= 0.0f;
                    if (pinned_i && pinned_j) continue;
                    const float hj = pair_horizon(V, j, A) * A.mode_f;
                    const f3 dvv = p_i - p_j;
                    const float dsq = dot(dvv, dvv);
                    const float eff = thick_i + X[j].w;
                    const float llj = T.local_len[j];
                    const float broad = eff + (lli + llj) * 1.3f + A.safety_margin + hi_eff + hj;
                    if (!(dsq <= broad * broad)) continue;
                    const bool same_island = island_i == PT.island[j];
                    const float dvv_len = sqrtf(fmaxf(dsq, 0.0f));
                    if (same_island && A.exclude_neighbors && near_folded_dist(T, index, j, dvv_len)) continue;
                    const float check = eff + A.safety_margin + hi_eff + hj;
                    if (VT) {  // 1. V-T triples over the first 8 star triangles of j
                    const float vt_bound = eff + llj * 1.3f + A.safety_margin + hi_eff + hj;
                    const bool near2 = A.exclude_neighbors && within_two_hops(T, index, j);
                    if (dsq <= vt_bound * vt_bound && !near2) {
                        const uint32_t ss = T.star_off[j], se = T.star_off[j + 1];
                        const uint32_t ns = se > ss ? min(se - ss, 8u) : 0u;
                        for (uint32_t st = 0; st < ns; ++st) {
                            const uint2 pr = T.star[ss + st];
                            const uint32_t v0 = pr.x, v1 = pr.y;
                            if (index == v0 || index == v1 || v0 >= A.nv || v1 >= A.nv) continue;
                            uint32_t c0 = j, c1 = v0, c2 = v1;  // rotate the smallest index first (orientation kept)
                            if (v0 < j && v0 < v1) { c0 = v0; c1 = v1; c2 = j; }
                            else if (v1 < j && v1 < v0) { c0 = v1; c1 = j; c2 = v0; }
                            f3 bc;
                            const f3 cp = closest_bary(p_i, xyz(P[c0]), xyz(P[c1]), xyz(P[c2]), bc);
                            const f3 dvt = p_i - cp;
                            const float dsq_vt = dot(dvt, dvt);
                            if (dsq_vt < check * check) {
                                if (A.exclude_neighbors &&
                                    (within_two_hops(T, index, v0) || within_two_hops(T, index, v1)))
                                    continue;
                                topk_offer(L, quota, dsq_vt, make_uint4(index, c0, c1, c2), eval, drop);
                            }
                        }
                    }
                    continue;
                    }
                    // 2. E-E quadruples over the first 8 neighbours of index and of j (the E-E bound
                    // equals the broad bound)
                    const uint32_t ai0 = PT.adj_off[index], ai1 = PT.adj_off[index + 1];
                    const uint32_t aj0 = PT.adj_off[j], aj1 = PT.adj_off[j + 1];
                    if (!(ai0 < ai1 && aj0 < aj1)) continue;
                    const uint32_t nei = min(ai1 - ai0, 8u), nej = min(aj1 - aj0, 8u);
                    for (uint32_t ei = 0; ei < nei; ++ei) {
                        const uint32_t ui = PT.adj_idx[ai0 + ei];
                        if (ui >= A.nv) continue;
                        const f3 p_ui = xyz(P[ui]);
                        const float llui = T.local_len[ui];
                        const f3 dui = p_ui - p_i;
                        const float len_sq_i = dot(dui, dui);
                        const float max_len_i = (lli + llui) * 1.5f;
                        if (len_sq_i > max_len_i * max_len_i || len_sq_i <= 1e-12f) continue;
                        for (uint32_t ej = 0; ej < nej; ++ej) {
                            const uint32_t vj = PT.adj_idx[aj0 + ej];
                            if (vj >= A.nv) continue;
                            if (ui == j || ui == vj || index == vj) continue;
                            const uint32_t c_j = min(j, vj), c_vj = max(j, vj);
                            const bool smaller =
                                smaller_anchor(P, X, V, T, PT, A, index, vj, index, j) ||
                                smaller_anchor(P, X, V, T, PT, A, ui, j, index, j) ||
                                smaller_anchor(P, X, V, T, PT, A, ui, vj, index, j) ||
                                smaller_anchor(P, X, V, T, PT, A, j, index, index, j) ||
                                smaller_anchor(P, X, V, T, PT, A, j, ui, index, j) ||
                                smaller_anchor(P, X, V, T, PT, A, vj, index, index, j) ||
                                smaller_anchor(P, X, V, T, PT, A, vj, ui, index, j);
                            if (pinned_i || smaller) continue;
                            const f3 p_vj = xyz(P[vj]);
                            const float llvj = T.local_len[vj];
                            const f3 djv = p_vj - p_j;
                            const float len_sq_j = dot(djv, djv);
                            const float max_len_j = (llj + llvj) * 1.5f;
                            if (len_sq_j > max_len_j * max_len_j || len_sq_j <= 1e-12f) continue;
                            const f3 p_cj = j < vj ? p_j : p_vj, p_cvj = j < vj ? p_vj : p_j;
                            const ST stp = closest_segs(p_i, p_ui, p_cj, p_cvj);
                            const f3 pt1 = p_i + (p_ui - p_i) * stp.s;
                            const f3 pt2 = p_cj + (p_cvj - p_cj) * stp.t;
                            const f3 dee = pt1 - pt2;
                            const float dsq_ee = dot(dee, dee);
                            if (!(dsq_ee < check * check && dsq_ee > EPS * EPS)) continue;
                            if (A.exclude_neighbors) {
                                if (near_folded_dist(T, index, vj, length(p_i - p_vj)) ||
                                    near_folded_dist(T, ui, j, length(p_ui - p_j)) ||
                                    near_folded_dist(T, ui, vj, length(p_ui - p_vj)))
                                    continue;
                            }
                            topk_offer(L, quota, dsq_ee, make_uint4(index, ui, c_j, c_vj), eval, drop);
                        }
                    }
                }
            }
    const uint4 none = make_uint4(PAIR_SENTINEL, 0u, 0u, 0u);
#pragma unroll
    for (uint32_t k = 0; k < 8u; ++k)
        if (k < quota) out[size_t(index) * quota + k] = k < L.n ? L.e[k] : none;
    if (eval) atomicAdd(&counters[VT ? 0 : 1], eval);
    if (drop) atomicAdd(&counters[VT ? 2 : 3], drop);
}

// record_untangle_cand (self_collision_solve_vt.wgsl:97-107): the nearest candidate wins through an
// atomic max of (1023 - distance rank) << 22 | (j + 1)
__device__ __forceinline__ void record_untangle(Accum *__restrict__ acc, uint32_t nv, uint32_t vi, uint32_t j, float dist,
                                                float cap) {
    if (vi >= nv || j >= 0x3FFFFEu || dist >= cap) return;
    const float ratio = clampf(dist / fmaxf(cap, 1e-7f), 0.0f, 1.0f);
    const uint32_t rank = 1023u - uint32_t(ratio * 1023.0f);
    atomicMax(&acc[vi].untangle, (rank << 22) | ((j + 1u) & 0x3FFFFFu));
}

// self_collision_solve_vt.wgsl:227-414 for one cached triple
__global__ void __launch_bounds__(256) k_pair_solve_vt(const float4 *__restrict__ P, const float4 *__restrict__ X,
                                                       const uint4 *__restrict__ pairs, uint32_t slots,
                                                       Accum *__restrict__ acc, uint32_t nv, SelfColTopo T,
                                                       int untangle) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= slots) return;
    const float EPS = 1e-7f;
    const uint4 pr = pairs[t];
    if (pr.x == PAIR_SENTINEL) return;
    const uint32_t index = pr.x, j = pr.y, v0 = pr.z, v1 = pr.w;
    if (index >= nv || j >= nv || v0 >= nv || v1 >= nv) return;
    if (index == j || index == v0 || index == v1 || v0 == v1 || j == v0 || j == v1) return;
    const float4 qi = P[index], xi = X[index], qj = P[j], xj = X[j], q0 = P[v0], x0 = X[v0], q1 = P[v1], x1 = X[v1];
    const bool pinned_i = qi.w <= 0.0f;
    const f3 p_i = xyz(qi), xo_i = xyz(xi);
    const f3 p_j = xyz(qj), p_v0 = xyz(q0), p_v1 = xyz(q1);
    const f3 pj_old = xyz(xj), p_v0o = xyz(x0), p_v1o = xyz(x1);
    const float eff = xi.w + xj.w;
    const f3 tri_move = (((p_j - pj_old) + (p_v0 - p_v0o)) + (p_v1 - p_v1o)) / 3.0f;
    const float wi = qi.w, wj = qj.w, w0 = q0.w, w1 = q1.w;
    const bool tri_pin = wj <= 0.0f || w0 <= 0.0f || w1 <= 0.0f;
    const bool ut = untangle && !pinned_i && T.layer[index] > T.layer[j];
    f3 ni_unit = mk3(0.0f, 0.0f, 0.0f);
    float nil = 0.0f;
    if (ut) {
        const f3 ni = xyz(T.normals[index]);
        nil = length(ni);
        if (nil > 0.5f) ni_unit = ni / fmaxf(nil, EPS);
    }
    const f3 rel_old = xo_i + tri_move;
    float ht = 0.0f, hu = 0.0f, hv = 0.0f;
    if (seg_tri_uv(rel_old, p_i, p_j, p_v0, p_v1, ht, hu, hv)) {
        const f3 move = p_i - rel_old;
        const f3 hit_pt = rel_old + move * ht;
        const f3 tcr = cross(p_v0 - p_j, p_v1 - p_j);
        const float tl = length(tcr);
        const f3 tn = tl > 1e-12f ? tcr / tl : mk3(0.0f, 1.0f, 0.0f);
        const float h_old_raw = dot(rel_old - p_j, tn);
        if (ut && tl > 1e-12f && (tri_pin || nil <= 0.5f || dot(ni_unit, tn) > 0.0f) && h_old_raw < 0.0f &&
            dot(move, tn) > 0.0f) {  // returning to the front: no response, the candidate is kept
            record_untangle(acc, nv, index, j, length(p_i - hit_pt), eff * 2.5f);
            return;
        }
        const float side = fabsf(h_old_raw) > 1e-6f ? h_old_raw : dot(-move, tn);
        const f3 n_side = side >= 0.0f ? tn : -tn;
        const float h_old = fmaxf(dot(rel_old - hit_pt, n_side), 0.0f);
        const float h_safe = fmaxf(h_old, eff * 0.25f);
        const f3 total = (hit_pt + n_side * h_safe) - p_i;
        const float pen_soft = fmaxf(0.0f, eff - h_safe);
        const f3 bc = mk3(clampf(1.0f - hu - hv, 0.0f, 1.0f), clampf(hu, 0.0f, 1.0f), clampf(hv, 0.0f, 1.0f));
        if (tri_pin) {
            if (!pinned_i) add_ccd(&acc[index], total + n_side * pen_soft);
        } else if (pinned_i) {
            const float wt = bc.x * bc.x * wj + bc.y * bc.y * w0 + bc.z * bc.z * w1;
            if (wt > EPS) {
                const f3 full = total + n_side * pen_soft;
                if (wj > 0.0f) add_ccd(&acc[j], (-full) * (bc.x * wj / wt));
                if (w0 > 0.0f) add_ccd(&acc[v0], (-full) * (bc.y * w0 / wt));
                if (w1 > 0.0f) add_ccd(&acc[v1], (-full) * (bc.z * w1 / wt));
            }
        } else {
            const float wt = bc.x * bc.x * wj + bc.y * bc.y * w0 + bc.z * bc.z * w1;
            const float wtot = wi + wt;
            if (wtot > EPS) {
                add_ccd(&acc[index], total * (wi / wtot));
                if (wj > 0.0f) add_ccd(&acc[j], (-total) * (bc.x * wj / wtot));
                if (w0 > 0.0f) add_ccd(&acc[v0], (-total) * (bc.y * w0 / wtot));
                if (w1 > 0.0f) add_ccd(&acc[v1], (-total) * (bc.z * w1 / wtot));
                if (pen_soft > EPS) {
                    const f3 tp = n_side * (pen_soft * 0.6f);
                    add_disp(&acc[index], tp * (wi / wtot));
                    if (wj > 0.0f) add_disp(&acc[j], (-tp) * (bc.x * wj / wtot));
                    if (w0 > 0.0f) add_disp(&acc[v0], (-tp) * (bc.y * w0 / wtot));
                    if (w1 > 0.0f) add_disp(&acc[v1], (-tp) * (bc.z * w1 / wtot));
                }
            }
        }
        return;
    }
    f3 bc;
    const f3 cp = closest_bary(p_i, p_j, p_v0, p_v1, bc);
    const f3 dvt = p_i - cp;
    const float dist = length(dvt);
    if (ut) {  // behind an inner triangle: the candidate goes to the untangle pass, no response
        const f3 tcr = cross(p_v0 - p_j, p_v1 - p_j);
        const float tl = length(tcr);
        const f3 tn = tl > EPS ? tcr / fmaxf(tl, EPS) : mk3(0.0f, 0.0f, 0.0f);
        const float h = tl > EPS ? dot(dvt, tn) : 0.0f;
        if (tl > EPS && (tri_pin || nil <= 0.5f || dot(ni_unit, tn) > 0.0f) && h < 0.0f) {
            record_untangle(acc, nv, index, j, dist, eff * 2.5f);
            return;
        }
    }
    if (pinned_i || !(dist < eff) || !(dist > EPS)) return;
    const f3 nvt = dvt / dist;
    const float pen = eff - dist;
    if (tri_pin) {
        add_pin_disp(&acc[index], nvt * pen + tri_move * 0.5f);
        return;
    }
    const float wt = bc.x * bc.x * wj + bc.y * bc.y * w0 + bc.z * bc.z * w1;
    const float wtot = wi + wt;
    if (wtot > EPS) {
        const f3 tp = nvt * (pen * 0.6f);
        add_disp(&acc[index], tp * (wi / wtot));
        if (wj > 0.0f) add_disp(&acc[j], (-tp) * (bc.x * wj / wtot));
        if (w0 > 0.0f) add_disp(&acc[v0], (-tp) * (bc.y * w0 / wtot));
        if (w1 > 0.0f) add_disp(&acc[v1], (-tp) * (bc.z * w1 / wtot));
    }
}

// self_collision_solve_ee.wgsl:231-589 for one cached quadruple
__global__ void __launch_bounds__(256) k_pair_solve_ee(const float4 *__restrict__ P, const float4 *__restrict__ X,
                                                       const uint4 *__restrict__ pairs, uint32_t slots,
                                                       Accum *__restrict__ acc, uint32_t nv, SelfColTopo T,
                                                       int untangle) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= slots) return;
    const float EPS = 1e-7f;
    const uint4 pr = pairs[t];
    if (pr.x == PAIR_SENTINEL) return;
    const uint32_t u0 = pr.x, u1 = pr.y, j = pr.z, vj = pr.w;
    if (u0 >= nv || u1 >= nv || j >= nv || vj >= nv) return;
    if (u0 == u1 || j == vj || u0 == j || u0 == vj || u1 == j || u1 == vj) return;
    const float4 qu0 = P[u0], qu1 = P[u1], qj = P[j], qvj = P[vj];
    const float wu0 = qu0.w, wu1 = qu1.w, wj = qj.w, wvj = qvj.w;
    const bool a_has_pin = wu0 <= 0.0f || wu1 <= 0.0f, b_has_pin = wj <= 0.0f || wvj <= 0.0f;
    if (a_has_pin && b_has_pin) return;
    const float4 xu0 = X[u0], xu1 = X[u1], xj = X[j], xvj = X[vj];
    const f3 p_u0 = xyz(qu0), p_u1 = xyz(qu1), p_j = xyz(qj), p_vj = xyz(qvj);
    const f3 p_u0o = xyz(xu0), p_u1o = xyz(xu1), p_jo = xyz(xj), p_vjo = xyz(xvj);
    const float th_u0 = xu0.w, th_u1 = xu1.w, th_j = xj.w, th_vj = xvj.w;
    const ST st1c = closest_segs(p_u0, p_u1, p_j, p_vj);
    const f3 d1c = mix3(p_u0, p_u1, st1c.s) - mix3(p_j, p_vj, st1c.t);
    const float dsqc = dot(d1c, d1c);
    const float effc = mixf(th_u0, th_u1, st1c.s) + mixf(th_j, th_vj, st1c.t);
    const float max_eff = fmaxf(th_u0, th_u1) + fmaxf(th_j, th_vj);
    const f3 du0 = p_u0 - p_u0o, du1 = p_u1 - p_u1o, dj = p_j - p_jo, dvj = p_vj - p_vjo;
    const f3 va = (du0 + du1) * 0.5f, vb = (dj + dvj) * 0.5f;
    const float rel = length(va - vb) + 0.5f * (length(du1 - du0) + length(dvj - dj));
    const float reach = max_eff + rel;
    if (dsqc > reach * reach) return;
    const f3 d0g = mix3(p_u0o, p_u1o, st1c.s) - mix3(p_jo, p_vjo, st1c.t);
    if ((rel > effc * 0.5f) || (dot(d0g, d1c) <= 0.0f)) {
        const Sweep W = {p_u0o, p_u0, p_u1o, p_u1, p_jo, p_j, p_vjo, p_vj};
        float min_sq = 1e20f, best_tau = 0.0f, half_span = 0.125f;
        ST min_st = {0.0f, 0.0f};
        float prev_vol = sweep_volume(W, 0.0f);
        float zc_tau = -1.0f;
#pragma unroll 1
        for (uint32_t k = 0; k <= 8u; ++k) {
            const float tk = float(k) * 0.125f;
            ST stk;
            const float dk = sweep_dist_sq(W, tk, stk);
            if (dk < min_sq) { min_sq = dk; best_tau = tk; min_st = stk; }
            if (k > 0u) {
                const float vol = sweep_volume(W, tk);
                if (prev_vol * vol <= 0.0f && zc_tau < 0.0f) zc_tau = tk - 0.0625f;
                prev_vol = vol;
            }
        }
        if (zc_tau >= 0.0f) {
            ST stz;
            const float dzs = sweep_dist_sq(W, zc_tau, stz);
            if (dzs < min_sq) { min_sq = dzs; best_tau = zc_tau; half_span = 0.0625f; min_st = stz; }
        }
        float tlo = fmaxf(0.0f, best_tau - half_span), thi = fminf(1.0f, best_tau + half_span);
        float dsq_star = min_sq;
        ST st_star = min_st;
#pragma unroll 1
        for (int it = 0; it < 6; ++it) {
            const float m1 = tlo + (thi - tlo) * (1.0f / 3.0f);
            const float m2 = thi - (thi - tlo) * (1.0f / 3.0f);
            ST s1, s2;
            const float e1 = sweep_dist_sq(W, m1, s1);
            const float e2 = sweep_dist_sq(W, m2, s2);
            if (e1 < dsq_star) { dsq_star = e1; st_star = s1; }
            if (e2 < dsq_star) { dsq_star = e2; st_star = s2; }
            if (e1 < e2) thi = m2; else tlo = m1;
        }
        const float ss = st_star.s, ts = st_star.t;
        const float eff_star = mixf(th_u0, th_u1, ss) + mixf(th_j, th_vj, ts);
        if (dsq_star >= eff_star * eff_star) return;
        if (untangle &&
            ee_ccd_untangling(T, T.layer[u0], T.layer[j], u0, j, wu0 == 0.0f && wu1 == 0.0f, wj == 0.0f && wvj == 0.0f,
                              p_u0o, p_u1o, p_jo, p_vjo, p_u0, p_u1, p_j, p_vj))
            return;
        const ST st0 = closest_segs(p_u0o, p_u1o, p_jo, p_vjo);
        const f3 delta0 = mix3(p_u0o, p_u1o, st0.s) - mix3(p_jo, p_vjo, st0.t);
        const float dist0 = length(delta0);
        if (dist0 > EPS) {
            const f3 n0 = delta0 / dist0;
            const float h_end = dot(mix3(p_u0, p_u1, ss) - mix3(p_j, p_vj, ts), n0);
            const bool crossover = (h_end <= 0.0f) || (dot(d1c, n0) <= 0.0f);
            if (h_end < eff_star && crossover) {
                const float pen = eff_star - h_end;
                const bool a_all = wu0 == 0.0f && wu1 == 0.0f;
                const bool b_all = wj == 0.0f && wvj == 0.0f;
                const bool cnt_pin = !(a_all || b_all);
                if (b_has_pin) {
                    const f3 em = ((p_j - p_jo) + (p_vj - p_vjo)) * 0.5f;
                    const float a = 1.0f - ss, b = ss;
                    const float den = a * a * wu0 + b * b * wu1;
                    if (den > EPS) {
                        const f3 pb = n0 * pen + em * 0.5f;
                        if (wu0 > 0.0f) add_ccd_ee(&acc[u0], pb * fminf((a * wu0) / den, 2.0f), cnt_pin);
                        if (wu1 > 0.0f) add_ccd_ee(&acc[u1], pb * fminf((b * wu1) / den, 2.0f), cnt_pin);
                    }
                } else if (a_has_pin) {
                    const f3 em = ((p_u0 - p_u0o) + (p_u1 - p_u1o)) * 0.5f;
                    const float a = 1.0f - ts, b = ts;
                    const float den = a * a * wj + b * b * wvj;
                    if (den > EPS) {
                        const f3 pb = (-n0) * pen + em * 0.5f;
                        if (wj > 0.0f) add_ccd_ee(&acc[j], pb * fminf((a * wj) / den, 2.0f), cnt_pin);
                        if (wvj > 0.0f) add_ccd_ee(&acc[vj], pb * fminf((b * wvj) / den, 2.0f), cnt_pin);
                    }
                } else {
                    const float we1 = (1.0f - ss) * (1.0f - ss) * wu0 + ss * ss * wu1;
                    const float we2 = (1.0f - ts) * (1.0f - ts) * wj + ts * ts * wvj;
                    const float wt = we1 + we2;
                    if (wt > EPS) {
                        const f3 tot = n0 * pen;
                        if (wu0 > 0.0f) add_ccd_ee(&acc[u0], tot * ((1.0f - ss) * wu0 / wt), true);
                        if (wu1 > 0.0f) add_ccd_ee(&acc[u1], tot * (ss * wu1 / wt), true);
                        if (wj > 0.0f) add_ccd_ee(&acc[j], (-tot) * ((1.0f - ts) * wj / wt), true);
                        if (wvj > 0.0f) add_ccd_ee(&acc[vj], (-tot) * (ts * wvj / wt), true);
                    }
                }
                return;
            }
        }
    }
    if (dsqc >= effc * effc || dsqc <= EPS * EPS) return;
    const float dist = sqrtf(dsqc);
    if (untangle && outer_edge_behind_inner(T, T.layer[u0], T.layer[j], u0, u1, j, vj, wu0 <= 0.0f && wu1 <= 0.0f,
                                            wj <= 0.0f && wvj <= 0.0f, d1c))
        return;
    const f3 n = d1c / dist;
    const float pen = effc - dist;
    const float sp = st1c.s, tp = st1c.t;
    if (b_has_pin) {
        const f3 em = ((p_j - p_jo) + (p_vj - p_vjo)) * 0.5f;
        const float a = 1.0f - sp, b = sp;
        const float den = a * a * wu0 + b * b * wu1;
        if (den > EPS) {
            const f3 pb = n * pen + em * 0.5f;
            if (wu0 > 0.0f) add_pin_disp(&acc[u0], pb * fminf((a * wu0) / den, 2.0f));
            if (wu1 > 0.0f) add_pin_disp(&acc[u1], pb * fminf((b * wu1) / den, 2.0f));
        }
    } else if (a_has_pin) {
        const f3 em = ((p_u0 - p_u0o) + (p_u1 - p_u1o)) * 0.5f;
        const float a = 1.0f - tp, b = tp;
        const float den = a * a * wj + b * b * wvj;
        if (den > EPS) {
            const f3 pb = (-n) * pen + em * 0.5f;
            if (wj > 0.0f) add_pin_disp(&acc[j], pb * fminf((a * wj) / den, 2.0f));
            if (wvj > 0.0f) add_pin_disp(&acc[vj], pb * fminf((b * wvj) / den, 2.0f));
        }
    } else {
        const float we1 = (1.0f - sp) * (1.0f - sp) * wu0 + sp * sp * wu1;
        const float we2 = (1.0f - tp) * (1.0f - tp) * wj + tp * tp * wvj;
        const float wt = we1 + we2;
        if (wt > EPS) {
            const f3 td = n * (pen * 0.6f);
            if (wu0 > 0.0f) add_disp(&acc[u0], td * ((1.0f - sp) * wu0 / wt));
            if (wu1 > 0.0f) add_disp(&acc[u1], td * (sp * wu1 / wt));
            if (wj > 0.0f) add_disp(&acc[j], (-td) * ((1.0f - tp) * wj / wt));
            if (wvj > 0.0f) add_disp(&acc[vj], (-td) * (tp * wvj / wt));
        }
    }
}

hipError_t launch_pair_collect(hipStream_t s, const float4 *P, const float4 *X, const float4 *V, const SelfColTopo &T,
                               const PairTopo &PT, const HashBufs &vh, uint4 *vt, uint4 *ee, uint32_t *counters,
                               const PairArgs &a) {
    if (!a.nv) return hipSuccess;
    hipLaunchKernelGGL(k_pair_collect<true>, dim3(grid_for(a.nv, 64)), dim3(64), 0, s, P, X, V, T, PT, vh.starts,
                       vh.sorted, vt, counters, a);
    hipLaunchKernelGGL(k_pair_collect<false>, dim3(grid_for(a.nv, 64)), dim3(64), 0, s, P, X, V, T, PT, vh.starts,
                       vh.sorted, ee, counters, a);
    return hipGetLastError();
}

hipError_t launch_pair_solve_vt(hipStream_t s, const float4 *P, const float4 *X, const uint4 *vt, uint32_t slots,
                                Accum *acc, uint32_t nv, const SelfColTopo &T, int untangle) {
    if (!slots) return hipSuccess;
    hipLaunchKernelGGL(k_pair_solve_vt, dim3(grid_for(slots, 256)), dim3(256), 0, s, P, X, vt, slots, acc, nv, T,
                       untangle);
    return hipGetLastError();
}

hipError_t launch_pair_solve_ee(hipStream_t s, const float4 *P, const float4 *X, const uint4 *ee, uint32_t slots,
                                Accum *acc, uint32_t nv, const SelfColTopo &T, int untangle) {
    if (!slots) return hipSuccess;
    hipLaunchKernelGGL(k_pair_solve_ee, dim3(grid_for(slots, 256)), dim3(256), 0, s, P, X, ee, slots, acc, nv, T,
                       untangle);
    return hipGetLastError();
}

hipError_t launch_compute_normals(hipStream_t s, const float4 *X, const SelfColTopo &t, float4 *normals, uint32_t nv) {
    if (!nv) return hipSuccess;
    hipLaunchKernelGGL(k_compute_normals, dim3(grid_for(nv, 256)), dim3(256), 0, s, X, t, normals, nv);
    return hipGetLastError();
}

hipError_t launch_sc_untangle(hipStream_t s, const float4 *P, const float4 *X, const SelfColTopo &t, Accum *acc,
                              const SelfColArgs &a) {
    if (!a.nv || !a.untangle) return hipSuccess;
    hipLaunchKernelGGL(k_sc_untangle, dim3(grid_for(a.nv, 256)), dim3(256), 0, s, P, X, t, acc, a);
    return hipGetLastError();
}

}  // namespace tc
