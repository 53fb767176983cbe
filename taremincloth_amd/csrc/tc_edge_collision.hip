// Collider-edge collision: cloth edges against the unique edges of the mesh collider, after the
// collider pass and the final pin, before self-collision.  Contacts go to the accumulators' pin
// (hard) channel and the self-collision apply kernel moves the vertices.
//
// Reference semantics (paths relative to the reference repository root):
//   edge extraction   crates/cloth_core/src/mesh.rs:1492-1567 (extract_collider_edges)
//   edge grid         crates/cloth_core/src/spatial_hash.rs:473-529 (build_collider_edge_grid),
//                     32768 buckets, the simulator's initial hash cell
//   kernel            crates/cloth_core/src/shaders/edge_collision.wgsl:153-276
//   dispatch          crates/cloth_core/src/simulation/dispatch.rs:252-278
//
// One thread per cloth edge walks the 27 cells around the edge's midpoint.  A collider edge sits
// in every bucket its half-cell samples touch, so one pair can be met in several of those cells;
// each meeting adds its push again, as in the reference (the integer sums make the order
// irrelevant, the multiplicity is part of the result).
#include "tc_sim_impl.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <map>

namespace tc {

struct EdgeSeg { float s, t, dist; };

// edge_collision.wgsl:98-150 (the distance comes out of the segment points p0 + d1 s, q0 + d2 t)
__device__ __forceinline__ EdgeSeg edge_closest(f3 p0, f3 p1, f3 q0, f3 q1) {
    const f3 d1 = p1 - p0, d2 = q1 - q0, r = p0 - q0;
    const float a = dot(d1, d1), e = dot(d2, d2), f = dot(d2, r);
    float s = 0.0f, t = 0.0f;
    if (a <= 1e-12f && e <= 1e-12f) return {0.0f, 0.0f, length(p0 - q0)};
    if (a <= 1e-12f) {
        t = clampf(f / e, 0.0f, 1.0f);
    } else {
        const float c = dot(d1, r);
        if (e <= 1e-12f) {
            s = clampf(-c / a, 0.0f, 1.0f);
        } else {
            const float b = dot(d1, d2);
            const float denom = a * e - b * b;
            s = denom > a * e * 1e-7f ? clampf((b * f - c * e) / denom, 0.0f, 1.0f) : 0.0f;
            t = (b * s + f) / e;
            if (t < 0.0f) { t = 0.0f; s = clampf(-c / a, 0.0f, 1.0f); }
            else if (t > 1.0f) { t = 1.0f; s = clampf((b - c) / a, 0.0f, 1.0f); }
        }
    }
    return {s, t, length((p0 + d1 * s) - (q0 + d2 * t))};
}

__global__ void __launch_bounds__(256) k_edge_collision(const float4 *__restrict__ P, const float4 *__restrict__ X,
                                                        const uint2 *__restrict__ edges, uint32_t ne,
                                                        const ColliderEdge *__restrict__ ce, uint32_t nce,
                                                        const uint32_t *__restrict__ starts,
                                                        const uint32_t *__restrict__ sorted, EdgeColArgs A,
                                                        Accum *__restrict__ acc) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= ne || nce == 0u) return;
    const float EPS = 1e-7f;
    const uint2 ed = edges[k];
    const uint32_t u0 = ed.x, u1 = ed.y;
    const float4 q0v = P[u0], q1v = P[u1];
    const float w0 = q0v.w, w1 = q1v.w;
    if (w0 + w1 <= EPS) return;
    const f3 p0 = xyz(q0v), p1 = xyz(q1v);
    const f3 mid = (p0 + p1) * 0.5f;
    const float cloth_th = fmaxf(X[u0].w, X[u1].w);
    const i3 cc = cell_of(mid, A.cell);
    const f3 amin = min3(p0, p1), amax = max3(p0, p1);
    for (int dz = -1; dz <= 1; ++dz)
        for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx) {
                const i3 nc = {int32_t(uint32_t(cc.x) + uint32_t(dx)), int32_t(uint32_t(cc.y) + uint32_t(dy)),
                               int32_t(uint32_t(cc.z) + uint32_t(dz))};
                const uint32_t h = hash_cell(nc, A.table);
                const uint32_t start = starts[h], end = starts[h + 1];
                for (uint32_t i = start; i < end; ++i) {
                    const ColliderEdge c = ce[sorted[i]];
                    const f3 q0 = mk3(c.p0x, c.p0y, c.p0z), q1 = mk3(c.p1x, c.p1y, c.p1z);
                    const float target = (cloth_th + c.thickness) * A.margin_scale + A.margin_offset;
                    const f3 cmin = min3(q0, q1), cmax = max3(q0, q1);
                    if (amax.x + target < cmin.x || amin.x - target > cmax.x || amax.y + target < cmin.y ||
                        amin.y - target > cmax.y || amax.z + target < cmin.z || amin.z - target > cmax.z)
                        continue;
                    const EdgeSeg r = edge_closest(p0, p1, q0, q1);
                    if (!(r.dist < target)) continue;
                    const f3 diff = (p0 + (p1 - p0) * r.s) - (q0 + (q1 - q0) * r.t);
                    const f3 en = mk3(c.nx, c.ny, c.nz);
                    const bool single = (c.flags & 1u) != 0u, has_n = dot(en, en) > 0.25f;
                    f3 n = mk3(0.0f, 0.0f, 1.0f);
                    float pen = target - r.dist;
                    if (single && has_n) {
                        const float sp = dot(diff, en);
                        if (sp >= 0.0f) {
                            n = r.dist > EPS ? diff / r.dist : en;
                            pen = target - r.dist;
                        } else {  // behind a one-sided edge: back out along its outward normal
                            n = en;
                            pen = fminf(target - sp, target * 2.0f);
                        }
                    } else if (r.dist > EPS) {
                        n = diff / r.dist;
                    } else {
                        const f3 cr = cross(p1 - p0, q1 - q0);
                        const float cl = length(cr);
                        if (cl > EPS) n = cr / cl;
                    }
                    const f3 push = n * pen;
                    const float a = 1.0f - r.s, b = r.s;
                    const float denom = a * a * w0 + b * b * w1;
                    if (denom > EPS) {
                        const float f0 = fminf((a * w0) / denom, 2.0f), f1 = fminf((b * w1) / denom, 2.0f);
                        auto pin_add = [&](uint32_t v, f3 d) {
                            auto fx = [](float x) { return f2i_sat(clampf(x * 1000000.0f, -50000000.0f, 50000000.0f)); };
                            atomicAdd(&acc[v].pdx, fx(d.x));
                            atomicAdd(&acc[v].pdy, fx(d.y));
                            atomicAdd(&acc[v].pdz, fx(d.z));
                            atomicAdd(&acc[v].pcount, 1u);
                        };
                        if (w0 > 0.0f) pin_add(u0, push * f0);
                        if (w1 > 0.0f) pin_add(u1, push * f1);
                    }
                }
            }
}

hipError_t launch_edge_collision(hipStream_t s, const float4 *P, const float4 *X, const uint2 *edges, uint32_t ne,
                                 const ColliderEdge *ce, uint32_t nce, const uint32_t *starts, const uint32_t *sorted,
                                 const EdgeColArgs &a, Accum *acc) {
    if (!ne || !nce) return hipSuccess;
    hipLaunchKernelGGL(k_edge_collision, dim3((ne + 255) / 256), dim3(256), 0, s, P, X, edges, ne, ce, nce, starts,
                       sorted, a, acc);
    return hipGetLastError();
}

// The unique edges of the (Morton-ordered) collider triangles, keyed by the bits of their two end
// points in ascending order, listed in order of first appearance: thickness the largest of the
// triangles sharing the edge, friction averaged pairwise in triangle order, normal the normalised
// sum of the triangle normals, single-sided only if every triangle is (mesh.rs:1503-1567).
std::vector<ColliderEdge> extract_collider_edges(const std::vector<MeshTri> &tris) {
    typedef std::array<uint32_t, 6> Key;
    std::map<Key, size_t> where;
    std::vector<ColliderEdge> out;
    std::vector<std::array<float, 3>> nsum;
    auto bits = [](float x) { uint32_t u; std::memcpy(&u, &x, 4); return u; };
    for (const MeshTri &t : tris) {
        const float e1[3] = {t.p1x - t.p0x, t.p1y - t.p0y, t.p1z - t.p0z};
        const float e2[3] = {t.p2x - t.p0x, t.p2y - t.p0y, t.p2z - t.p0z};
        const float cr[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
        const float cl = std::sqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]);
        float tn[3] = {0.0f, 0.0f, 0.0f};
        if (cl > 1e-12f) { tn[0] = cr[0] / cl; tn[1] = cr[1] / cl; tn[2] = cr[2] / cl; }
        const uint32_t single = t.flags & 1u;
        const float P3[3][3] = {{t.p0x, t.p0y, t.p0z}, {t.p1x, t.p1y, t.p1z}, {t.p2x, t.p2y, t.p2z}};
        for (int k = 0; k < 3; ++k) {
            const float *a = P3[k], *b = P3[(k + 1) % 3];
            const std::array<uint32_t, 3> ab = {bits(a[0]), bits(a[1]), bits(a[2])}, bb = {bits(b[0]), bits(b[1]), bits(b[2])};
            Key key;
            if (ab < bb) { std::copy(ab.begin(), ab.end(), key.begin()); std::copy(bb.begin(), bb.end(), key.begin() + 3); }
            else { std::copy(bb.begin(), bb.end(), key.begin()); std::copy(ab.begin(), ab.end(), key.begin() + 3); }
            auto it = where.find(key);
            if (it == where.end()) {
                where.emplace(key, out.size());
                ColliderEdge e{};
                e.p0x = a[0]; e.p0y = a[1]; e.p0z = a[2]; e.thickness = t.thickness;
                e.p1x = b[0]; e.p1y = b[1]; e.p1z = b[2]; e.friction = t.friction;
                e.flags = single;
                out.push_back(e);
                nsum.push_back({tn[0], tn[1], tn[2]});
            } else {
                ColliderEdge &e = out[it->second];
                e.thickness = std::max(e.thickness, t.thickness);
                e.friction = (e.friction + t.friction) * 0.5f;
                std::array<float, 3> &ns = nsum[it->second];
                ns[0] += tn[0]; ns[1] += tn[1]; ns[2] += tn[2];
                e.flags &= single;
            }
        }
    }
    for (size_t i = 0; i < out.size(); ++i) {
        const std::array<float, 3> &ns = nsum[i];
        const float nl = std::sqrt(ns[0] * ns[0] + ns[1] * ns[1] + ns[2] * ns[2]);
        if (nl > 1e-6f) { out[i].nx = ns[0] / nl; out[i].ny = ns[1] / nl; out[i].nz = ns[2] / nl; }
        else { out[i].nx = out[i].ny = out[i].nz = 0.0f; }
    }
    return out;
}

// build_collider_edge_grid (spatial_hash.rs:473-529): each edge in every bucket its samples at
// half-cell pitch fall into (once per bucket), buckets filled in edge order.
void collider_edge_grid(const std::vector<ColliderEdge> &edges, float cell, uint32_t table, std::vector<uint32_t> &starts,
                        std::vector<uint32_t> &sorted) {
    const float safe = std::max(cell, 1e-4f), pitch = safe * 0.5f;
    std::vector<uint32_t> counts(table, 0u);
    std::vector<std::pair<uint32_t, uint32_t>> entries;
    std::vector<uint32_t> hs;
    auto sat = [](float x) -> int32_t {
        if (x != x) return 0;
        if (x >= 2147483520.0f) return 2147483647;
        if (x <= -2147483648.0f) return INT32_MIN;
        return int32_t(x);
    };
    for (size_t i = 0; i < edges.size(); ++i) {
        const ColliderEdge &e = edges[i];
        const float dx = e.p1x - e.p0x, dy = e.p1y - e.p0y, dz = e.p1z - e.p0z;
        const float len = std::sqrt(dx * dx + dy * dy + dz * dz);
        const float cf = std::ceil(len / pitch);
        size_t steps = cf != cf || cf <= 0.0f ? 0 : (cf >= 1.8446744e19f ? SIZE_MAX : size_t(cf));
        steps = std::min<size_t>(std::max<size_t>(steps, 1), 4096);
        hs.clear();
        for (size_t k = 0; k <= steps; ++k) {
            const float t = float(k) / float(steps);
            const float px = e.p0x + dx * t, py = e.p0y + dy * t, pz = e.p0z + dz * t;
            const uint32_t c0 = uint32_t(sat(std::floor(px / safe))), c1 = uint32_t(sat(std::floor(py / safe))),
                           c2 = uint32_t(sat(std::floor(pz / safe)));
            hs.push_back(((c0 * 73856093u) ^ (c1 * 19349663u) ^ (c2 * 83492791u)) % table);
        }
        std::sort(hs.begin(), hs.end());
        hs.erase(std::unique(hs.begin(), hs.end()), hs.end());
        for (uint32_t h : hs) { ++counts[h]; entries.emplace_back(h, uint32_t(i)); }
    }
    starts.assign(size_t(table) + 1, 0u);
    uint32_t sum = 0;
    for (uint32_t h = 0; h < table; ++h) { starts[h] = sum; sum += counts[h]; }
    starts[table] = sum;
    std::vector<uint32_t> cur(starts.begin(), starts.end() - 1);
    sorted.assign(entries.size(), 0u);
    for (const auto &en : entries) sorted[cur[en.first]++] = en.second;
}

}  // namespace tc

// Rebuilds the collider edges and their grid from the simulator's triangles (upload_colliders,
// simulation/mod.rs:2730-2743).
int upload_collider_edges(tc_sim *s) {
    const std::vector<tc::ColliderEdge> ce = tc::extract_collider_edges(s->tris_h);
    std::vector<uint32_t> starts, sorted;
    tc::collider_edge_grid(ce, s->mesh.cell_size, tc::COLLIDER_EDGE_TABLE, starts, sorted);
    TC_HIP(hipStreamSynchronize(s->stream));
    TC_HIP(s->cedges.upload(ce.data(), ce.size(), s->stream));
    TC_HIP(s->cedge_starts.upload(starts.data(), starts.size(), s->stream));
    TC_HIP(s->cedge_sorted.upload(sorted.data(), sorted.size(), s->stream));
    TC_HIP(hipStreamSynchronize(s->stream));
    s->ncedges = uint32_t(ce.size());
    s->touch();
    return TC_OK;
}

hipError_t enqueue_edge_collision(tc_sim *s, float4 *P) {
    tc::EdgeColArgs a;
    a.cell = s->mesh.cell_size;
    a.table = tc::COLLIDER_EDGE_TABLE;
    a.margin_scale = s->edge_margin_scale;
    a.margin_offset = s->edge_margin_offset;
    hipError_t e = tc::launch_edge_collision(s->stream, P, s->X.p, s->edges.p, uint32_t(s->mesh.e_v0.size()), s->cedges.p,
                                             s->ncedges, s->cedge_starts.p, s->cedge_sorted.p, a, s->accum.p);
    if (e != hipSuccess) return e;
    return tc::launch_sc_apply(s->stream, P, s->local_len.p, s->accum.p, s->sc_args(), s->block);
}

extern "C" {

int tc_sim_set_enable_edge_collision(tc_sim *s, int enable) {
    if (!s) return fail(TC_ERR_VALUE, "null simulator");
    s->edge_collision = enable != 0;
    s->touch();
    return TC_OK;
}

int tc_sim_set_edge_margin(tc_sim *s, float scale, float offset) {
    if (!s) return fail(TC_ERR_VALUE, "null simulator");
    s->edge_margin_scale = std::max(scale, 1.0f);  // simulation/mod.rs:1007-1016
    s->edge_margin_offset = std::max(offset, 0.0f);
    s->touch();
    return TC_OK;
}

int tc_sim_collider_edge_count(const tc_sim *s, uint32_t *n) {
    if (!s || !n) return fail(TC_ERR_VALUE, "null argument");
    *n = s->ncedges;
    return TC_OK;
}

int tc_sim_get_collider_edges(tc_sim *s, float *out, uint32_t cap) {
    if (!s) return fail(TC_ERR_VALUE, "null simulator");
    if (cap < s->ncedges || (!out && s->ncedges)) return fail(TC_ERR_VALUE, "output must hold 12 floats per collider edge");
    if (!s->ncedges) return TC_OK;
    TC_HIP(hipStreamSynchronize(s->stream));
    TC_HIP(hipMemcpy(out, s->cedges.p, sizeof(tc::ColliderEdge) * s->ncedges, hipMemcpyDeviceToHost));
    return TC_OK;
}

}  // extern "C"
