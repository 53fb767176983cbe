// Mesh SDF bakes for gfx950 and their host-side preparation.
//
// Dense bake (bake_mesh_sdf_gpu_ext, crates/cloth_core/src/sdf_baker.rs:1114-1323 with
// shaders/bake_mesh_sdf.wgsl:80-121): every voxel centre takes the signed distance to the
// nearest triangle (Ericson's closest point, sign from the face normal, first strict minimum
// in triangle order wins), stored as pack2x16float(d, 1.0) in Z, Y, X order.
//
// Here the triangles are sorted along a Morton curve and grouped into clusters of 32 with a box
// each.  Work items are (brick of voxels, segment of 128 clusters): a wavefront bounds its
// voxels' nearest distances from above, then evaluates only the clusters and triangles whose
// boxes can still hold a nearer triangle for one of its voxels, and merges per voxel through a
// 64-bit key (squared distance bits, original triangle index).  The key orders like (d_sq,
// index), so the result is the same argmin the reference's in-order scan returns, whatever
// order the work is done in.
//
// The coarse pass of the hierarchical bake (bake_mesh_sdf_hier_coarse_main.wgsl) keeps the 16
// nearest per cell the same way and rebuilds the reference's eight-entry list from them; the
// fine pass (bake_mesh_sdf_hier_fine_main.wgsl) is restated voxel by voxel.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <numeric>
#include <set>
#include <unordered_map>

#include "tc_device.h"
#include "tc_sdf.h"

namespace tc {
namespace sdf {

// ------------------------------------------------------------------ host preparation
namespace {
inline void cross3(const float *a, const float *b, float *o) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}
inline void normalize3(const float *v, float *o) {
    const float len_sq = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
    if (len_sq > 1e-12f) {
        const float inv = 1.0f / std::sqrt(len_sq);
        o[0] = v[0] * inv; o[1] = v[1] * inv; o[2] = v[2] * inv;
    } else {
        o[0] = 0.0f; o[1] = 1.0f; o[2] = 0.0f;
    }
}
inline bool valid_tri(const int32_t *t, size_t n) {
    return t[0] >= 0 && t[1] >= 0 && t[2] >= 0 && size_t(t[0]) < n && size_t(t[1]) < n && size_t(t[2]) < n &&
           t[0] != t[1] && t[1] != t[2] && t[2] != t[0];
}
inline uint32_t spread10(uint32_t v) {
    v &= 1023u;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}
}  // namespace

size_t close_open_boundary_loops(std::vector<float> &verts, std::vector<int32_t> &tris,
                                 std::vector<std::vector<float>> *weights) {
    const size_t n = verts.size() / 3, m = tris.size() / 3;
    if (m < 4 || n < 4) return 0;
    std::map<std::pair<int32_t, int32_t>, uint32_t> count;
    std::set<std::array<int32_t, 3>> existing;
    for (size_t t = 0; t < m; ++t) {
        const int32_t *tr = &tris[3 * t];
        if (!valid_tri(tr, n)) continue;
        std::array<int32_t, 3> s = {tr[0], tr[1], tr[2]};
        std::sort(s.begin(), s.end());
        existing.insert(s);
        for (int e = 0; e < 3; ++e) {
            const int32_t u = tr[e], v = tr[(e + 1) % 3];
            ++count[{std::min(u, v), std::max(u, v)}];
        }
    }
    std::vector<std::pair<int32_t, int32_t>> bedge;
    std::vector<std::array<float, 3>> bnormal;
    std::unordered_map<int32_t, std::vector<size_t>> next;
    for (size_t t = 0; t < m; ++t) {
        const int32_t *tr = &tris[3 * t];
        if (!valid_tri(tr, n)) continue;
        const float *p0 = &verts[3 * size_t(tr[0])], *p1 = &verts[3 * size_t(tr[1])], *p2 = &verts[3 * size_t(tr[2])];
        const float e01[3] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
        const float e02[3] = {p2[0] - p0[0], p2[1] - p0[1], p2[2] - p0[2]};
        float c[3], nrm[3];
        cross3(e01, e02, c);
        normalize3(c, nrm);
        for (int e = 0; e < 3; ++e) {
            const int32_t u = tr[e], v = tr[(e + 1) % 3];
            if (count[{std::min(u, v), std::max(u, v)}] == 1u) {
                next[u].push_back(bedge.size());
                bedge.push_back({u, v});
                bnormal.push_back({nrm[0], nrm[1], nrm[2]});
            }
        }
    }
    if (bedge.empty()) return 0;
    std::vector<char> used(bedge.size(), 0);
    size_t closed = 0;
    for (size_t start = 0; start < bedge.size(); ++start) {
        if (used[start]) continue;
        used[start] = 1;
        const int32_t su = bedge[start].first, sv = bedge[start].second;
        std::vector<int32_t> loop = {su, sv};
        std::vector<size_t> loop_edges = {start};
        int32_t curr = sv;
        bool is_closed = false;
        while (loop.size() <= n) {
            if (curr == su) {
                loop.pop_back();
                is_closed = true;
                break;
            }
            bool found = false;
            int32_t nxt = 0;
            auto it = next.find(curr);
            if (it != next.end())
                for (size_t e : it->second)
                    if (!used[e]) {
                        used[e] = 1;
                        loop_edges.push_back(e);
                        nxt = bedge[e].second;
                        found = true;
                        break;
                    }
            if (!found) break;
            loop.push_back(nxt);
            curr = nxt;
        }
        const size_t k = loop.size();
        if (!is_closed || k < 3 || k >= n) continue;
        const float inv_k = 1.0f / float(k);
        float center[3] = {0.0f, 0.0f, 0.0f};
        for (int32_t vi : loop)
            for (int a = 0; a < 3; ++a) center[a] += verts[3 * size_t(vi) + a];
        for (int a = 0; a < 3; ++a) center[a] *= inv_k;
        float cap_sum[3] = {0.0f, 0.0f, 0.0f};
        for (size_t i = 0; i < k; ++i) {
            const float *pu = &verts[3 * size_t(loop[i])], *pv = &verts[3 * size_t(loop[(i + 1) % k])];
            const float d1[3] = {pu[0] - pv[0], pu[1] - pv[1], pu[2] - pv[2]};
            const float d2[3] = {center[0] - pv[0], center[1] - pv[1], center[2] - pv[2]};
            float cn[3];
            cross3(d1, d2, cn);
            for (int a = 0; a < 3; ++a) cap_sum[a] += cn[a];
        }
        float cap_n[3];
        normalize3(cap_sum, cap_n);
        float dot_sum = 0.0f;
        for (size_t e : loop_edges) {
            const auto &an = bnormal[e];
            dot_sum += an[0] * cap_n[0] + an[1] * cap_n[1] + an[2] * cap_n[2];
        }
        const float mean_dot = dot_sum / float(std::max<size_t>(loop_edges.size(), 1));
        if (mean_dot <= -0.85f) continue;  // the outer rim of a flat sheet stays open
        if (k == 3) {
            std::array<int32_t, 3> s = {loop[0], loop[1], loop[2]};
            std::sort(s.begin(), s.end());
            if (existing.count(s)) continue;
            tris.insert(tris.end(), {loop[0], loop[2], loop[1]});
            existing.insert(s);
            ++closed;
            continue;
        }
        const int32_t ci = int32_t(verts.size() / 3);
        verts.insert(verts.end(), center, center + 3);
        if (weights)  // each bone's weight at the cap centre: the loop's mean (sdf_baker.rs:427-439)
            for (std::vector<float> &w : *weights) {
                float ws = 0.0f;
                for (int32_t vi : loop)
                    if (size_t(vi) < w.size()) ws += w[size_t(vi)];
                if (w.size() < size_t(ci)) w.resize(size_t(ci), 0.0f);
                w.push_back(ws * inv_k);
            }
        for (size_t i = 0; i < k; ++i) tris.insert(tris.end(), {loop[(i + 1) % k], loop[i], ci});
        ++closed;
    }
    return closed;
}

Domain mesh_sdf_domain(const std::vector<float> &verts, float voxel_size, float margin, float thickness) {
    Domain dm;
    for (int a = 0; a < 3; ++a) { dm.min[a] = INFINITY; dm.max[a] = -INFINITY; }
    for (size_t i = 0; i + 2 < verts.size(); i += 3)
        for (int a = 0; a < 3; ++a) {
            dm.min[a] = std::fmin(dm.min[a], verts[i + a]);
            dm.max[a] = std::fmax(dm.max[a], verts[i + a]);
        }
    for (int a = 0; a < 3; ++a) dm.size[a] = std::fmax(dm.max[a] - dm.min[a], 0.05f);
    for (int a = 0; a < 3; ++a) {
        const float pad = std::fmax(dm.size[a] * margin, 0.04f) + std::fmax(thickness, 0.0f) + 0.02f;
        dm.min[a] -= pad;
        dm.max[a] += pad;
        dm.size[a] = dm.max[a] - dm.min[a];
    }
    const float v = std::fmax(voxel_size, 0.001f);
    auto dim = [&](float s) {
        const double raw = std::ceil(double(s / v));
        return uint32_t(std::min(2048.0, std::max(8.0, raw)));
    };
    dm.w = dim(dm.size[0]);
    dm.h = dim(dm.size[1]);
    dm.d = dim(dm.size[2]);
    return dm;
}

void mesh_sdf_bone_info(const Domain &dm, float friction, float thickness, float restitution, float out[20]) {
    for (int k = 0; k < 20; ++k) out[k] = 0.0f;
    for (int a = 0; a < 3; ++a) {
        out[a] = dm.min[a];
        out[4 + a] = dm.max[a];
        out[8 + a] = 1.0f / dm.size[a];
        out[12 + a] = -dm.min[a] / dm.size[a];
    }
    out[7] = float(dm.w);
    out[16] = friction;
    out[17] = thickness;
    out[18] = restitution;
}

std::vector<BakeTri> collect_bake_triangles(const std::vector<float> &verts, const std::vector<int32_t> &tris) {
    std::vector<BakeTri> out;
    const size_t n = verts.size() / 3;
    for (size_t t = 0; t + 2 < tris.size(); t += 3) {
        const int32_t i0 = tris[t], i1 = tris[t + 1], i2 = tris[t + 2];
        if (i0 < 0 || i1 < 0 || i2 < 0 || size_t(i0) >= n || size_t(i1) >= n || size_t(i2) >= n) continue;
        const float *p0 = &verts[3 * size_t(i0)], *p1 = &verts[3 * size_t(i1)], *p2 = &verts[3 * size_t(i2)];
        const float e01[3] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
        const float e02[3] = {p2[0] - p0[0], p2[1] - p0[1], p2[2] - p0[2]};
        float c[3], nrm[3];
        cross3(e01, e02, c);
        normalize3(c, nrm);
        out.push_back({p0[0], p0[1], p0[2], 0.0f, p1[0], p1[1], p1[2], 0.0f, p2[0], p2[1], p2[2], 0.0f, nrm[0], nrm[1],
                       nrm[2], 0.0f});
    }
    if (out.empty()) out.push_back({0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0.0f, 1.0f, 0.0f, 0.0f});
    return out;
}

Clusters build_clusters(const std::vector<BakeTri> &tris, const Domain &dm) {
    Clusters c;
    const float ext = std::max(dm.size[0], std::max(dm.size[1], dm.size[2]));
    const float scale = 1023.0f / std::max(ext, 1e-6f);
    std::vector<uint32_t> code(tris.size());
    for (size_t t = 0; t < tris.size(); ++t) {
        const BakeTri &b = tris[t];
        const float cx = (b.p0x + b.p1x + b.p2x) / 3.0f, cy = (b.p0y + b.p1y + b.p2y) / 3.0f,
                    cz = (b.p0z + b.p1z + b.p2z) / 3.0f;
        auto q = [&](float v, int a) {
            const float f = (v - dm.min[a]) * scale;
            return uint32_t(std::min(1023.0f, std::max(0.0f, f)));
        };
        code[t] = spread10(q(cx, 0)) | (spread10(q(cy, 1)) << 1) | (spread10(q(cz, 2)) << 2);
    }
    std::vector<uint32_t> order(tris.size());
    std::iota(order.begin(), order.end(), 0u);
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return code[a] < code[b]; });
    c.tris.resize(tris.size());
    c.index = order;
    for (size_t k = 0; k < order.size(); ++k) c.tris[k] = tris[order[k]];
    const size_t nc = (tris.size() + CLUSTER - 1) / CLUSTER;
    c.lo.resize(nc);
    c.hi.resize(nc);
    for (size_t ci = 0; ci < nc; ++ci) {
        float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (size_t k = ci * CLUSTER; k < std::min(tris.size(), (ci + 1) * CLUSTER); ++k) {
            const BakeTri &b = c.tris[k];
            const float px[3] = {b.p0x, b.p1x, b.p2x}, py[3] = {b.p0y, b.p1y, b.p2y}, pz[3] = {b.p0z, b.p1z, b.p2z};
            for (int j = 0; j < 3; ++j) {
                lo[0] = std::min(lo[0], px[j]); hi[0] = std::max(hi[0], px[j]);
                lo[1] = std::min(lo[1], py[j]); hi[1] = std::max(hi[1], py[j]);
                lo[2] = std::min(lo[2], pz[j]); hi[2] = std::max(hi[2], pz[j]);
            }
        }
        c.lo[ci] = make_float4(lo[0], lo[1], lo[2], 0.0f);
        c.hi[ci] = make_float4(hi[0], hi[1], hi[2], 0.0f);
    }
    return c;
}

// ------------------------------------------------------------------ device
// Ericson, Real-Time Collision Detection, 5.1.5 (operation order of the reference shader)
__device__ __forceinline__ f3 closest_on_tri(f3 p, f3 a, f3 b, f3 c) {
    const f3 ab = b - a, ac = c - a, ap = p - a;
    const float d1 = dot(ab, ap), d2 = dot(ac, ap);
    if (d1 <= 0.0f && d2 <= 0.0f) return a;
    const f3 bp = p - b;
    const float d3 = dot(ab, bp), d4 = dot(ac, bp);
    if (d3 >= 0.0f && d4 <= d3) return b;
    const float vc = d1 * d4 - d3 * d2;
    if (vc <= 0.0f && d1 >= 0.0f && d3 <= 0.0f) {
        const float v = d1 / (d1 - d3);
        return a + ab * v;
    }
    const f3 cp = p - c;
    const float d5 = dot(ab, cp), d6 = dot(ac, cp);
    if (d6 >= 0.0f && d5 <= d6) return c;
    const float vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0f && d2 >= 0.0f && d6 <= 0.0f) {
        const float w = d2 / (d2 - d6);
        return a + ac * w;
    }
    const float va = d3 * d6 - d5 * d4;
    if (va <= 0.0f && (d4 - d3) >= 0.0f && (d5 - d6) >= 0.0f) {
        const float w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        return b + (c - b) * w;
    }
    const float denom = 1.0f / (va + vb + vc);
    const float v = vb * denom, w = vc * denom;
    return a + ab * v + ac * w;
}

__device__ __forceinline__ uint32_t pack_dist(float d) {
    const _Float16 h = (_Float16)d;
    uint16_t bits;
    __builtin_memcpy(&bits, &h, 2);
    return uint32_t(bits) | 0x3C000000u;  // second half: 1.0
}
__device__ __forceinline__ float unpack_dist(uint32_t v) {
    const uint16_t bits = uint16_t(v & 0xFFFFu);
    _Float16 h;
    __builtin_memcpy(&h, &bits, 2);
    return (float)h;
}

__device__ __forceinline__ float voxel_coord(uint32_t i, uint32_t dim, float lo, float hi) {
    const float uvw = (float(i) + 0.5f) * (1.0f / float(dim));
    return lo + uvw * (hi - lo);
}

__device__ __forceinline__ float wave_min_f(float v) {
    for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ float wave_max_f(float v) {
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
    return v;
}

// (d_sq, index) lexicographic "less": the in-order first strict minimum
__device__ __forceinline__ bool better(float d, uint32_t i, float bd, uint32_t bi) { return d < bd || (d == bd && i < bi); }

constexpr uint32_t NO_TRI = 0xFFFFFFFFu;
constexpr float FAR = 3.0e38f;

// Lower bound of the squared distance from p to a box, shortened by 1e-5 m per axis and 1e-4
// relative so that the rounding of an exact point-triangle distance can never fall below it.
__device__ __forceinline__ float lb_box_point(f3 lo, f3 hi, f3 p) {
    const float gx = fmaxf(0.0f, fmaxf(lo.x - p.x, p.x - hi.x) - 1e-5f);
    const float gy = fmaxf(0.0f, fmaxf(lo.y - p.y, p.y - hi.y) - 1e-5f);
    const float gz = fmaxf(0.0f, fmaxf(lo.z - p.z, p.z - hi.z) - 1e-5f);
    return (gx * gx + gy * gy + gz * gz) * 0.9999f;
}
__device__ __forceinline__ float lb_box_box(f3 lo, f3 hi, f3 bmin, f3 bmax) {
    const float gx = fmaxf(0.0f, fmaxf(lo.x - bmax.x, bmin.x - hi.x) - 1e-5f);
    const float gy = fmaxf(0.0f, fmaxf(lo.y - bmax.y, bmin.y - hi.y) - 1e-5f);
    const float gz = fmaxf(0.0f, fmaxf(lo.z - bmax.z, bmin.z - hi.z) - 1e-5f);
    return (gx * gx + gy * gy + gz * gz) * 0.9999f;
}
// Squared distance from the farthest point of the box [bmin, bmax] to q (to any point of box
// [lo, hi] for the second form): an upper bound on the distance to whatever lies there.
__device__ __forceinline__ float ub_box_point(f3 bmin, f3 bmax, f3 q) {
    const float dx = fmaxf(fabsf(q.x - bmin.x), fabsf(q.x - bmax.x));
    const float dy = fmaxf(fabsf(q.y - bmin.y), fabsf(q.y - bmax.y));
    const float dz = fmaxf(fabsf(q.z - bmin.z), fabsf(q.z - bmax.z));
    return dx * dx + dy * dy + dz * dz;
}
__device__ __forceinline__ float ub_box_box(f3 bmin, f3 bmax, f3 lo, f3 hi) {
    const float dx = fmaxf(fabsf(hi.x - bmin.x), fabsf(bmax.x - lo.x));
    const float dy = fmaxf(fabsf(hi.y - bmin.y), fabsf(bmax.y - lo.y));
    const float dz = fmaxf(fabsf(hi.z - bmin.z), fabsf(bmax.z - lo.z));
    return dx * dx + dy * dy + dz * dz;
}

// ------------------------------------------------------------------ dense bake, segmented
// The nearest triangle of a voxel is the minimum of the 64-bit key (squared distance bits << 32 |
// original index): squared distances are non-negative, so their bits order like the values and
// the key orders like (d_sq, index), the first strict minimum of the in-order scan.  Work items
// are (brick of 8 x 8 x DZ voxels, segment of SEG clusters) so that bricks whose voxels see many
// nearly equidistant triangles (the middle of a closed surface) spread over many wavefronts;
// each keeps the current key minima of its voxels as culling bounds and merges with atomicMin.
constexpr uint32_t DZ = 2, SEG = SEG_CLUSTERS;

__device__ __forceinline__ unsigned long long dense_key(float dsq, uint32_t idx) {
    return (static_cast<unsigned long long>(__float_as_uint(dsq)) << 32) | idx;
}

struct Brick {
    uint32_t vx, vy, vz0;
    f3 p[DZ];
    bool ok[DZ];
    f3 bmin, bmax;
};

__device__ __forceinline__ Brick dense_brick(const DenseArgs &a, uint32_t brick, uint32_t lane) {
    Brick b;
    const uint32_t nbx = (a.w + 7) / 8, nby = (a.h + 7) / 8;
    const uint32_t bx = brick % nbx, by = (brick / nbx) % nby, bz = brick / (nbx * nby);
    b.vx = bx * 8 + (lane & 7u);
    b.vy = by * 8 + (lane >> 3);
    b.vz0 = bz * DZ;
    const bool col_ok = b.vx < a.w && b.vy < a.h;
    for (uint32_t j = 0; j < DZ; ++j) {
        const uint32_t vz = b.vz0 + j;
        b.ok[j] = col_ok && vz < a.d;
        b.p[j] = mk3(voxel_coord(b.vx, a.w, a.lmin[0], a.lmax[0]), voxel_coord(b.vy, a.h, a.lmin[1], a.lmax[1]),
                     voxel_coord(vz, a.d, a.lmin[2], a.lmax[2]));
    }
    const uint32_t x0 = bx * 8, y0 = by * 8, z0 = bz * DZ;
    const uint32_t x1 = min(a.w, x0 + 8) - 1, y1 = min(a.h, y0 + 8) - 1, z1 = min(a.d, z0 + DZ) - 1;
    b.bmin = mk3(voxel_coord(x0, a.w, a.lmin[0], a.lmax[0]), voxel_coord(y0, a.h, a.lmin[1], a.lmax[1]),
                 voxel_coord(z0, a.d, a.lmin[2], a.lmax[2]));
    b.bmax = mk3(voxel_coord(x1, a.w, a.lmin[0], a.lmax[0]), voxel_coord(y1, a.h, a.lmin[1], a.lmax[1]),
                 voxel_coord(z1, a.d, a.lmin[2], a.lmax[2]));
    return b;
}

__device__ __forceinline__ uint32_t dense_bricks(const DenseArgs &a) {
    return ((a.w + 7) / 8) * ((a.h + 7) / 8) * ((a.d + DZ - 1) / DZ);
}

// Upper bound for every voxel of a brick: the farthest brick point from the first vertex of the
// nearest cluster; the keys start there with no triangle.
__global__ void __launch_bounds__(256) k_bake_dense_bound(DenseArgs a) {
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t brick = blockIdx.x * 4u + (threadIdx.x >> 6);
    if (brick >= dense_bricks(a)) return;  // whole wavefront
    const Brick b = dense_brick(a, brick, lane);
    float ub = FAR;
    for (uint32_t c = lane; c < a.nclusters; c += 64u) {
        const BakeTri &t0 = a.tris[size_t(c) * CLUSTER];
        ub = fminf(ub, ub_box_point(b.bmin, b.bmax, mk3(t0.p0x, t0.p0y, t0.p0z)));
    }
    const unsigned long long key = dense_key(wave_min_f(ub) * 1.0001f + 1e-12f, NO_TRI);
    for (uint32_t j = 0; j < DZ; ++j)
        if (b.ok[j]) a.keys[(size_t(b.vz0 + j) * a.h + b.vy) * a.w + b.vx] = key;
}

__global__ void __launch_bounds__(256) k_bake_dense_seg(DenseArgs a) {
    const uint32_t lane = threadIdx.x & 63u;
    const unsigned long long nseg = (a.nclusters + SEG - 1) / SEG;
    const unsigned long long item = a.item0 + blockIdx.x * 4ull + (threadIdx.x >> 6);
    if (item >= (unsigned long long)dense_bricks(a) * nseg) return;  // whole wavefront
    const uint32_t brick = uint32_t(item / nseg), seg = uint32_t(item % nseg);
    const Brick b = dense_brick(a, brick, lane);
    size_t flat[DZ];
    float bsq[DZ], bsg_unused[DZ];
    uint32_t bid[DZ];
    unsigned long long start[DZ];
    (void)bsg_unused;
    for (uint32_t j = 0; j < DZ; ++j) {
        flat[j] = (size_t(b.vz0 + j) * a.h + b.vy) * a.w + b.vx;
        start[j] = b.ok[j] ? a.keys[flat[j]] : ~0ull;
        bsq[j] = __uint_as_float(uint32_t(start[j] >> 32));
        bid[j] = uint32_t(start[j]);
    }
    __shared__ float4 s_tri[4][CLUSTER * 4];
    __shared__ uint32_t s_idx[4][CLUSTER];
    const uint32_t wv = threadIdx.x >> 6;
    const uint32_t c0 = seg * SEG, c1 = min(a.nclusters, c0 + SEG);
    for (uint32_t base = c0; base < c1; base += 64u) {
        float worst = -1.0f;
        for (uint32_t j = 0; j < DZ; ++j)
            if (b.ok[j]) worst = fmaxf(worst, bsq[j]);
        const float bound = wave_max_f(worst);
        const uint32_t c = base + lane;
        bool cand = false;
        if (c < c1) {
            const float4 lo = a.clo[c], hi = a.chi[c];
            cand = !(lb_box_box(mk3(lo.x, lo.y, lo.z), mk3(hi.x, hi.y, hi.z), b.bmin, b.bmax) > bound);
        }
        uint64_t m = __ballot(cand);
        while (m) {
            const uint32_t cc = base + uint32_t(__ffsll((unsigned long long)m) - 1);
            m &= m - 1u;
            const float4 cl4 = a.clo[cc], ch4 = a.chi[cc];
            const f3 cl = mk3(cl4.x, cl4.y, cl4.z), ch = mk3(ch4.x, ch4.y, ch4.z);
            bool need = false;
            for (uint32_t j = 0; j < DZ; ++j) need = need || (b.ok[j] && !(lb_box_point(cl, ch, b.p[j]) > bsq[j]));
            if (!__any(need)) continue;
            const uint32_t t0 = cc * CLUSTER, cnt = min(a.ntris - t0, CLUSTER);
            const float4 *src = reinterpret_cast<const float4 *>(a.tris + t0);
            __builtin_amdgcn_wave_barrier();  // the previous cluster's reads come first
            if (lane < cnt * 4u) s_tri[wv][lane] = src[lane];
            if (lane + 64u < cnt * 4u) s_tri[wv][lane + 64u] = src[lane + 64u];
            if (lane < cnt) s_idx[wv][lane] = a.index[t0 + lane];
            __builtin_amdgcn_wave_barrier();
            for (uint32_t t = 0; t < cnt; ++t) {
                const float4 q0 = s_tri[wv][4 * t], q1 = s_tri[wv][4 * t + 1], q2 = s_tri[wv][4 * t + 2];
                const f3 A = mk3(q0.x, q0.y, q0.z), B = mk3(q1.x, q1.y, q1.z), C = mk3(q2.x, q2.y, q2.z);
                const f3 tmin = min3(A, min3(B, C)), tmax = max3(A, max3(B, C));
                bool tn[DZ];
                bool any = false;
                for (uint32_t j = 0; j < DZ; ++j) {
                    tn[j] = b.ok[j] && !(lb_box_point(tmin, tmax, b.p[j]) > bsq[j]);
                    any = any || tn[j];
                }
                if (!__any(any)) continue;
                const uint32_t idx = s_idx[wv][t];
                for (uint32_t j = 0; j < DZ; ++j) {
                    if (!tn[j]) continue;
                    const f3 q = closest_on_tri(b.p[j], A, B, C);
                    const f3 diff = b.p[j] - q;
                    const float dsq = dot(diff, diff);
                    if (better(dsq, idx, bsq[j], bid[j])) { bsq[j] = dsq; bid[j] = idx; }
                }
            }
        }
    }
    for (uint32_t j = 0; j < DZ; ++j) {
        if (!b.ok[j]) continue;
        const unsigned long long k = dense_key(bsq[j], bid[j]);
        if (k < start[j]) atomicMin(&a.keys[flat[j]], k);
    }
}

// The winning triangle of each voxel, evaluated once more for the sign of its distance.
__global__ void __launch_bounds__(256) k_bake_dense_finish(DenseArgs a) {
    const size_t i = size_t(blockIdx.x) * 256u + threadIdx.x;
    const size_t n = size_t(a.w) * a.h * a.d;
    if (i >= n) return;
    const uint32_t vx = uint32_t(i % a.w), vy = uint32_t((i / a.w) % a.h), vz = uint32_t(i / (size_t(a.w) * a.h));
    const uint32_t idx = uint32_t(a.keys[i]);
    if (idx >= a.ntris) {  // unreachable: the bound comes from a triangle that is always evaluated
        a.out[i] = pack_dist(10.0f);
        return;
    }
    const f3 p = mk3(voxel_coord(vx, a.w, a.lmin[0], a.lmax[0]), voxel_coord(vy, a.h, a.lmin[1], a.lmax[1]),
                     voxel_coord(vz, a.d, a.lmin[2], a.lmax[2]));
    const BakeTri tr = a.otris[idx];
    const f3 q = closest_on_tri(p, mk3(tr.p0x, tr.p0y, tr.p0z), mk3(tr.p1x, tr.p1y, tr.p1z), mk3(tr.p2x, tr.p2y, tr.p2z));
    const f3 diff = p - q;
    const float dsq = dot(diff, diff);
    a.out[i] = pack_dist((dot(diff, mk3(tr.nx, tr.ny, tr.nz)) >= 0.0f ? 1.0f : -1.0f) * sqrtf(dsq));
}

// The dense bake's nearest triangles as an fp32 brick atlas: one wavefront per 8^3 brick, each
// lane eight voxels (lane + 64 k), so every store of a wavefront covers 64 consecutive floats.
__global__ void __launch_bounds__(256) k_bake_atlas(DenseArgs a, float *__restrict__ atlas, uint32_t bx_n,
                                                    uint32_t by_n, uint32_t nbricks) {
    const uint32_t brick = blockIdx.x * 4u + (threadIdx.x >> 6);
    if (brick >= nbricks) return;
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t bx = brick % bx_n, by = (brick / bx_n) % by_n, bz = brick / (bx_n * by_n);
    for (uint32_t k = 0; k < 8; ++k) {
        const uint32_t v = lane + 64u * k;  // (lz * 8 + ly) * 8 + lx
        const uint32_t lx = v & 7u, ly = (v >> 3) & 7u, lz = v >> 6;
        const uint32_t vx = min(bx * ATLAS_BRICK + lx, a.w - 1), vy = min(by * ATLAS_BRICK + ly, a.h - 1),
                       vz = min(bz * ATLAS_BRICK + lz, a.d - 1);
        const size_t i = (size_t(vz) * a.h + vy) * a.w + vx;
        const uint32_t idx = uint32_t(a.keys[i]);
        float val = 10.0f;
        if (idx < a.ntris) {
            const f3 p = mk3(voxel_coord(vx, a.w, a.lmin[0], a.lmax[0]), voxel_coord(vy, a.h, a.lmin[1], a.lmax[1]),
                             voxel_coord(vz, a.d, a.lmin[2], a.lmax[2]));
            const BakeTri tr = a.otris[idx];
            const f3 q = closest_on_tri(p, mk3(tr.p0x, tr.p0y, tr.p0z), mk3(tr.p1x, tr.p1y, tr.p1z),
                                        mk3(tr.p2x, tr.p2y, tr.p2z));
            const f3 diff = p - q;
            val = (dot(diff, mk3(tr.nx, tr.ny, tr.nz)) >= 0.0f ? 1.0f : -1.0f) * sqrtf(dot(diff, diff));
        }
        atlas[size_t(brick) * 512u + v] = val;
    }
}

hipError_t launch_bake_atlas(hipStream_t st, const DenseArgs &a, float *atlas) {
    const uint32_t bx = (a.w + 7) / 8, by = (a.h + 7) / 8, bz = (a.d + 7) / 8;
    const uint32_t nb = bx * by * bz;
    if (!nb) return hipSuccess;
    hipLaunchKernelGGL(k_bake_atlas, dim3((nb + 3) / 4), dim3(256), 0, st, a, atlas, bx, by, nb);
    return hipGetLastError();
}

// ------------------------------------------------------------------ coarse pass, segmented
// bake_mesh_sdf_hier_coarse_main.wgsl keeps each cell's eight nearest triangles with its own
// update rule (strict comparisons, so an entry pushed down the list passes the entries of equal
// distance below it), applied in triangle order.  Here the 16 nearest of each cell are found in
// (d_sq, index) order by (brick of 4 x 4 x 4 cells, segment of SEG clusters) work items, each
// keeping a sorted list of 64-bit keys, and merged per cell.  Only triangles at or below the
// eighth distance D8 can be in the reference's list or move an entry of it, so replaying exactly
// those in index order gives the reference's list; with no equal distances among them the
// replay is the identity and is skipped.  More than 16 at or below D8 replays every triangle.
constexpr uint32_t KC = 16;
constexpr unsigned long long NO_KEY = ~0ull;

__device__ __forceinline__ float key_dsq(unsigned long long k) {
    return k == NO_KEY ? FAR : __uint_as_float(uint32_t(k >> 32));
}

struct CBrick {
    uint32_t vx, vy, vz;
    bool ok;
    f3 p, bmin, bmax;
};

__device__ __forceinline__ uint32_t coarse_bricks(const DenseArgs &a) {
    return ((a.w + 3) / 4) * ((a.h + 3) / 4) * ((a.d + 3) / 4);
}

__device__ __forceinline__ CBrick coarse_brick(const DenseArgs &a, uint32_t brick, uint32_t lane) {
    CBrick b;
    const uint32_t nbx = (a.w + 3) / 4, nby = (a.h + 3) / 4;
    const uint32_t bx = brick % nbx, by = (brick / nbx) % nby, bz = brick / (nbx * nby);
    b.vx = bx * 4 + (lane & 3u);
    b.vy = by * 4 + ((lane >> 2) & 3u);
    b.vz = bz * 4 + (lane >> 4);
    b.ok = b.vx < a.w && b.vy < a.h && b.vz < a.d;
    b.p = mk3(voxel_coord(b.vx, a.w, a.lmin[0], a.lmax[0]), voxel_coord(b.vy, a.h, a.lmin[1], a.lmax[1]),
              voxel_coord(b.vz, a.d, a.lmin[2], a.lmax[2]));
    const uint32_t x0 = bx * 4, y0 = by * 4, z0 = bz * 4;
    const uint32_t x1 = min(a.w, x0 + 4) - 1, y1 = min(a.h, y0 + 4) - 1, z1 = min(a.d, z0 + 4) - 1;
    b.bmin = mk3(voxel_coord(x0, a.w, a.lmin[0], a.lmax[0]), voxel_coord(y0, a.h, a.lmin[1], a.lmax[1]),
                 voxel_coord(z0, a.d, a.lmin[2], a.lmax[2]));
    b.bmax = mk3(voxel_coord(x1, a.w, a.lmin[0], a.lmax[0]), voxel_coord(y1, a.h, a.lmin[1], a.lmax[1]),
                 voxel_coord(z1, a.d, a.lmin[2], a.lmax[2]));
    return b;
}

// Upper bound on every cell's 16th nearest distance: the farthest pair of points between the
// brick and a cluster holding at least 16 triangles.  The cells' bound keys start there.
__global__ void __launch_bounds__(256) k_bake_coarse_bound(DenseArgs a) {
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t brick = blockIdx.x * 4u + (threadIdx.x >> 6);
    if (brick >= coarse_bricks(a)) return;  // whole wavefront
    const CBrick b = coarse_brick(a, brick, lane);
    float ub = FAR;
    for (uint32_t c = lane; c < a.nclusters; c += 64u)
        if (min(a.ntris - c * CLUSTER, CLUSTER) >= KC) {
            const float4 lo = a.clo[c], hi = a.chi[c];
            ub = fminf(ub, ub_box_box(b.bmin, b.bmax, mk3(lo.x, lo.y, lo.z), mk3(hi.x, hi.y, hi.z)));
        }
    ub = wave_min_f(ub);
    const unsigned long long key = ub >= FAR ? NO_KEY : dense_key(ub * 1.0001f + 1e-12f, NO_TRI);
    if (b.ok) a.keys[(size_t(b.vz) * a.h + b.vy) * a.w + b.vx] = key;
}

// Insert `key` into the ascending list `l` (the largest falls off).
__device__ __forceinline__ void insert_key(unsigned long long (&l)[KC], unsigned long long key) {
#pragma unroll
    for (uint32_t k = 0; k < KC; ++k)
        if (key < l[k]) {
            const unsigned long long t = l[k];
            l[k] = key;
            key = t;
        }
}

__global__ void __launch_bounds__(256) k_bake_coarse_seg(DenseArgs a) {
    const uint32_t lane = threadIdx.x & 63u, wv = threadIdx.x >> 6;
    const uint32_t nseg = (a.nclusters + SEG - 1) / SEG;
    const unsigned long long item = blockIdx.x * 4ull + wv;
    if (item >= (unsigned long long)a.brick_count * nseg) return;  // whole wavefront
    const uint32_t lb = uint32_t(item / nseg), seg = uint32_t(item % nseg);
    const CBrick b = coarse_brick(a, a.brick0 + lb, lane);
    const size_t flat = (size_t(b.vz) * a.h + b.vy) * a.w + b.vx;
    unsigned long long l[KC];
    for (uint32_t k = 0; k < KC; ++k) l[k] = NO_KEY;
    unsigned long long gb = b.ok ? a.keys[flat] : NO_KEY;
    float th = key_dsq(gb);
    __shared__ float4 s_tri[4][CLUSTER * 4];
    __shared__ uint32_t s_idx[4][CLUSTER];
    const uint32_t c0 = seg * SEG, c1 = min(a.nclusters, c0 + SEG);
    for (uint32_t base = c0; base < c1; base += 64u) {
        if (b.ok && base != c0) {  // other work items of this cell may have tightened its bound
            gb = min(gb, a.keys[flat]);
            th = fminf(key_dsq(l[KC - 1]), key_dsq(gb));
        }
        const float bound = wave_max_f(b.ok ? th : -1.0f);
        const uint32_t c = base + lane;
        bool cand = false;
        if (c < c1) {
            const float4 lo = a.clo[c], hi = a.chi[c];
            cand = !(lb_box_box(mk3(lo.x, lo.y, lo.z), mk3(hi.x, hi.y, hi.z), b.bmin, b.bmax) > bound);
        }
        uint64_t m = __ballot(cand);
        while (m) {
            const uint32_t cc = base + uint32_t(__ffsll((unsigned long long)m) - 1);
            m &= m - 1u;
            const float4 cl4 = a.clo[cc], ch4 = a.chi[cc];
            const bool need = b.ok && !(lb_box_point(mk3(cl4.x, cl4.y, cl4.z), mk3(ch4.x, ch4.y, ch4.z), b.p) > th);
            if (!__any(need)) continue;
            const uint32_t t0 = cc * CLUSTER, cnt = min(a.ntris - t0, CLUSTER);
            const float4 *src = reinterpret_cast<const float4 *>(a.tris + t0);
            __builtin_amdgcn_wave_barrier();  // the previous cluster's reads come first
            if (lane < cnt * 4u) s_tri[wv][lane] = src[lane];
            if (lane + 64u < cnt * 4u) s_tri[wv][lane + 64u] = src[lane + 64u];
            if (lane < cnt) s_idx[wv][lane] = a.index[t0 + lane];
            __builtin_amdgcn_wave_barrier();
            for (uint32_t t = 0; t < cnt; ++t) {
                const float4 q0 = s_tri[wv][4 * t], q1 = s_tri[wv][4 * t + 1], q2 = s_tri[wv][4 * t + 2];
                const f3 A = mk3(q0.x, q0.y, q0.z), B = mk3(q1.x, q1.y, q1.z), C = mk3(q2.x, q2.y, q2.z);
                const bool tn = b.ok && !(lb_box_point(min3(A, min3(B, C)), max3(A, max3(B, C)), b.p) > th);
                if (!__any(tn)) continue;
                if (!tn) continue;
                const f3 q = closest_on_tri(b.p, A, B, C);
                const f3 diff = b.p - q;
                const unsigned long long key = dense_key(dot(diff, diff), s_idx[wv][t]);
                if (key < l[KC - 1]) {
                    insert_key(l, key);
                    th = fminf(key_dsq(l[KC - 1]), key_dsq(gb));
                }
            }
        }
    }
    if (!b.ok) return;
    unsigned long long *dst = a.part + (size_t(lb) * 64u + lane) * nseg * KC + size_t(seg) * KC;
    for (uint32_t k = 0; k < KC; ++k) dst[k] = l[k];
    if (l[KC - 1] != NO_KEY && l[KC - 1] < gb) atomicMin(&a.keys[flat], l[KC - 1]);
}

// The reference's list update on (squared distance, index) pairs (signs are not needed: the
// cell's texel comes from its nearest triangle, the first entry either way).
__device__ __forceinline__ void ref_insert_pair(float (&sq)[8], uint32_t (&id)[8], float d, uint32_t i) {
    if (!(d < sq[7])) return;
    float is = d;
    uint32_t ii = i;
#pragma unroll
    for (int k = 0; k < 8; ++k)
        if (is < sq[k]) {
            const float ts = sq[k];
            const uint32_t ti = id[k];
            sq[k] = is; id[k] = ii;
            is = ts; ii = ti;
        }
}

// One thread per cell: merge the work items' lists, write the cell's texel and the reference's
// eight-entry list.
__global__ void __launch_bounds__(256) k_bake_coarse_merge(DenseArgs a) {
    const uint32_t g = blockIdx.x * 256u + threadIdx.x;
    if (g >= a.brick_count * 64u) return;
    const uint32_t lb = g / 64u, lane = g % 64u;
    const CBrick b = coarse_brick(a, a.brick0 + lb, lane);
    if (!b.ok) return;
    const size_t flat = (size_t(b.vz) * a.h + b.vy) * a.w + b.vx;
    const uint32_t nseg = (a.nclusters + SEG - 1) / SEG;
    unsigned long long l[KC];
    for (uint32_t k = 0; k < KC; ++k) l[k] = NO_KEY;
    const unsigned long long *src = a.part + size_t(g) * nseg * KC;
    for (uint32_t sg = 0; sg < nseg; ++sg)
        for (uint32_t k = 0; k < KC; ++k) {
            const unsigned long long key = src[size_t(sg) * KC + k];
            if (key >= l[KC - 1]) break;  // each list is ascending
            insert_key(l, key);
        }
    // texel: the nearest triangle, evaluated again for its sign
    const uint32_t i0 = uint32_t(l[0]);
    if (l[0] == NO_KEY || i0 >= a.ntris) {
        a.out[flat] = pack_dist(10.0f);
    } else {
        const BakeTri tr = a.otris[i0];
        const f3 q = closest_on_tri(b.p, mk3(tr.p0x, tr.p0y, tr.p0z), mk3(tr.p1x, tr.p1y, tr.p1z),
                                    mk3(tr.p2x, tr.p2y, tr.p2z));
        const f3 diff = b.p - q;
        a.out[flat] = pack_dist((dot(diff, mk3(tr.nx, tr.ny, tr.nz)) >= 0.0f ? 1.0f : -1.0f) * sqrtf(dot(diff, diff)));
    }
    uint32_t rid[8];
    bool tie = false;
    for (int k = 0; k < 8; ++k) tie = tie || key_dsq(l[k]) == key_dsq(l[k + 1]);
    if (!tie) {
        for (int k = 0; k < 8; ++k) rid[k] = l[k] == NO_KEY ? 0u : uint32_t(l[k]);
    } else {
        float rsq[8];
        for (int k = 0; k < 8; ++k) { rsq[k] = 1e10f; rid[k] = 0u; }
        const float d8 = key_dsq(l[7]);
        if (l[KC - 1] == NO_KEY || key_dsq(l[KC - 1]) > d8) {
            // every triangle at or below D8 is in the list: replay those in index order
            uint32_t used = 0u;
            for (uint32_t n = 0; n < KC; ++n) {
                uint32_t pick = NO_TRI;
                float pd = 0.0f;
                int pk = -1;
                for (uint32_t k = 0; k < KC; ++k)
                    if (!((used >> k) & 1u) && l[k] != NO_KEY && key_dsq(l[k]) <= d8 && uint32_t(l[k]) < pick) {
                        pick = uint32_t(l[k]);
                        pd = key_dsq(l[k]);
                        pk = int(k);
                    }
                if (pk < 0) break;
                used |= 1u << pk;
                ref_insert_pair(rsq, rid, pd, pick);
            }
        } else {
            for (uint32_t i = 0; i < a.ntris; ++i) {
                const BakeTri tr = a.otris[i];
                const f3 q = closest_on_tri(b.p, mk3(tr.p0x, tr.p0y, tr.p0z), mk3(tr.p1x, tr.p1y, tr.p1z),
                                            mk3(tr.p2x, tr.p2y, tr.p2z));
                const f3 diff = b.p - q;
                ref_insert_pair(rsq, rid, dot(diff, diff), i);
            }
        }
    }
    for (int k = 0; k < 8; ++k) a.nearest[flat * 8 + k] = rid[k];
}

__device__ __forceinline__ void fine_candidate(f3 p, const BakeTri *tris, uint32_t ntris, uint32_t ti, uint32_t &best_i,
                                               float &best_sq, float &best_sg, float &second_sq, float &second_sg) {
    if (ti >= ntris || ti == best_i) return;
    const BakeTri tr = tris[ti];
    const f3 A = mk3(tr.p0x, tr.p0y, tr.p0z), B = mk3(tr.p1x, tr.p1y, tr.p1z), C = mk3(tr.p2x, tr.p2y, tr.p2z);
    const f3 tmin = min3(A, min3(B, C)), tmax = max3(A, max3(B, C));
    const f3 z = mk3(0.0f, 0.0f, 0.0f);
    const f3 dv = max3(tmin - p, z) + max3(p - tmax, z);
    if (dot(dv, dv) >= second_sq) return;
    const f3 q = closest_on_tri(p, A, B, C);
    const f3 diff = p - q;
    const float dsq = dot(diff, diff);
    const float sg = sqrtf(dsq) * (dot(diff, mk3(tr.nx, tr.ny, tr.nz)) >= 0.0f ? 1.0f : -1.0f);
    if (dsq < best_sq) {
        second_sq = best_sq; second_sg = best_sg;
        best_sq = dsq; best_sg = sg; best_i = ti;
    } else if (dsq < second_sq) {
        second_sq = dsq; second_sg = sg;
    }
}

// bake_mesh_sdf_hier_fine_main.wgsl restated: far voxels copy the parent coarse value, near
// voxels evaluate the eight nearest triangles of the 5x5x5 coarse cells around the parent.
__global__ void __launch_bounds__(256) k_bake_fine(FineArgs a) {
    const uint32_t vx = blockIdx.x * 8 + (threadIdx.x & 7u);
    const uint32_t vy = blockIdx.y * 8 + ((threadIdx.x >> 3) & 7u);
    const uint32_t vz = blockIdx.z * 4 + (threadIdx.x >> 6);
    if (vx >= a.w || vy >= a.h || vz >= a.d) return;
    const f3 lmin = mk3(a.lmin[0], a.lmin[1], a.lmin[2]), lmax = mk3(a.lmax[0], a.lmax[1], a.lmax[2]);
    const f3 size = lmax - lmin;
    const f3 uvw = mk3((float(vx) + 0.5f) * (1.0f / float(a.w)), (float(vy) + 0.5f) * (1.0f / float(a.h)),
                       (float(vz) + 0.5f) * (1.0f / float(a.d)));
    const f3 p = mk3(lmin.x + uvw.x * size.x, lmin.y + uvw.y * size.y, lmin.z + uvw.z * size.z);
    const f3 spacing = mk3(size.x / float(a.cw), size.y / float(a.ch), size.z / float(a.cd));
    const float halfdiag = 0.5f * length(spacing);
    const f3 rel = p - lmin;
    const int32_t px = min(max(f2i_sat(floorf(rel.x / spacing.x)), 0), int32_t(a.cw) - 1);
    const int32_t py = min(max(f2i_sat(floorf(rel.y / spacing.y)), 0), int32_t(a.ch) - 1);
    const int32_t pz = min(max(f2i_sat(floorf(rel.z / spacing.z)), 0), int32_t(a.cd) - 1);
    const uint32_t pidx = (uint32_t(pz) * a.ch + uint32_t(py)) * a.cw + uint32_t(px);
    const uint32_t dc_packed = a.coarse_dist[pidx];
    const float dc = unpack_dist(dc_packed);
    const size_t flat = (size_t(vz) * a.h + vy) * a.w + vx;
    if (fabsf(dc) - halfdiag >= a.band) {
        a.out[flat] = dc_packed;
        return;
    }
    uint32_t best_i = 0xFFFFFFFFu;
    float best_sq = 1e10f, best_sg = 10.0f, second_sq = 1e10f, second_sg = 10.0f;
    for (int dz = -2; dz <= 2; ++dz)
        for (int dy = -2; dy <= 2; ++dy)
            for (int dx = -2; dx <= 2; ++dx) {
                const int32_t cx = min(max(px + dx, 0), int32_t(a.cw) - 1);
                const int32_t cy = min(max(py + dy, 0), int32_t(a.ch) - 1);
                const int32_t cz = min(max(pz + dz, 0), int32_t(a.cd) - 1);
                const size_t base = ((size_t(cz) * a.ch + uint32_t(cy)) * a.cw + uint32_t(cx)) * 8u;
                for (int k = 0; k < 8; ++k)
                    fine_candidate(p, a.tris, a.ntris, a.coarse_tri[base + k], best_i, best_sq, best_sg, second_sq,
                                   second_sg);
            }
    bool ambiguous = sqrtf(second_sq) - sqrtf(best_sq) <= 0.0002f;
    if (ambiguous) ambiguous = (best_sg >= 0.0f) != (second_sg >= 0.0f);
    if (ambiguous) {  // exhaustive scan in triangle order: k_bake_resolve
        a.amb[atomicAdd(a.amb_count, 1u)] = uint32_t(flat);
        return;
    }
    a.out[flat] = pack_dist(best_sg);
}


__global__ void __launch_bounds__(256) k_bake_resolve(ResolveArgs a) {
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t item = blockIdx.x * 4u + (threadIdx.x >> 6);
    if (item >= a.count) return;  // whole wavefront
    const uint32_t flat = a.list[item];
    const uint32_t vx = flat % a.w, vy = (flat / a.w) % a.h, vz = flat / (a.w * a.h);
    // the voxel centre as the fine pass computes it
    const f3 lmin = mk3(a.lmin[0], a.lmin[1], a.lmin[2]), lmax = mk3(a.lmax[0], a.lmax[1], a.lmax[2]);
    const f3 size = lmax - lmin;
    const f3 uvw = mk3((float(vx) + 0.5f) * (1.0f / float(a.w)), (float(vy) + 0.5f) * (1.0f / float(a.h)),
                       (float(vz) + 0.5f) * (1.0f / float(a.d)));
    const f3 p = mk3(lmin.x + uvw.x * size.x, lmin.y + uvw.y * size.y, lmin.z + uvw.z * size.z);
    // upper bound: the nearest first vertex of a cluster
    float ub = FAR;
    for (uint32_t c = lane; c < a.nclusters; c += 64u) {
        const BakeTri &t0 = a.tris[size_t(c) * CLUSTER];
        const f3 d = p - mk3(t0.p0x, t0.p0y, t0.p0z);
        ub = fminf(ub, dot(d, d));
    }
    float bound = wave_min_f(ub) * 1.0001f + 1e-12f;
    float bsq = 1e10f, bsg = 10.0f;
    uint32_t bid = NO_TRI;
    const uint32_t rounds = (a.nclusters + 63u) / 64u;
    for (uint32_t r = 0; r < rounds; ++r) {
        if ((r & 3u) == 0u) bound = fminf(bound, wave_min_f(bsq));
        const uint32_t c = r * 64u + lane;
        if (c >= a.nclusters) continue;
        const float4 lo = a.clo[c], hi = a.chi[c];
        if (lb_box_point(mk3(lo.x, lo.y, lo.z), mk3(hi.x, hi.y, hi.z), p) > bound) continue;
        const uint32_t t1 = min(a.ntris, (c + 1u) * CLUSTER);
        for (uint32_t t = c * CLUSTER; t < t1; ++t) {
            const BakeTri tr = a.tris[t];
            const f3 A = mk3(tr.p0x, tr.p0y, tr.p0z), B = mk3(tr.p1x, tr.p1y, tr.p1z), C = mk3(tr.p2x, tr.p2y, tr.p2z);
            if (lb_box_point(min3(A, min3(B, C)), max3(A, max3(B, C)), p) > fminf(bound, bsq)) continue;
            const f3 q = closest_on_tri(p, A, B, C);
            const f3 diff = p - q;
            const float dsq = dot(diff, diff);
            const uint32_t idx = a.index[t];
            if (!better(dsq, idx, bsq, bid)) continue;
            bsq = dsq;
            bid = idx;
            bsg = sqrtf(dsq) * (dot(diff, mk3(tr.nx, tr.ny, tr.nz)) >= 0.0f ? 1.0f : -1.0f);
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        const float os = __shfl_xor(bsq, o, 64), og = __shfl_xor(bsg, o, 64);
        const uint32_t oi = __shfl_xor(bid, o, 64);
        if (better(os, oi, bsq, bid)) { bsq = os; bsg = og; bid = oi; }
    }
    if (lane == 0) a.out[flat] = pack_dist(bsg);
}

hipError_t launch_bake_dense(hipStream_t st, const DenseArgs &a) {
    if (a.topk == 1) {
        const size_t nb = size_t((a.w + 7) / 8) * ((a.h + 7) / 8) * ((a.d + DZ - 1) / DZ);
        const size_t nseg = (a.nclusters + SEG - 1) / SEG, nvox = size_t(a.w) * a.h * a.d;
        hipLaunchKernelGGL(k_bake_dense_bound, dim3(uint32_t((nb + 3) / 4)), dim3(256), 0, st, a);
        const size_t items = nb * nseg, chunk = size_t(1) << 30;  // items per launch
        for (size_t i0 = 0; i0 < items; i0 += chunk) {
            DenseArgs c = a;
            c.item0 = i0;
            const size_t n = std::min(chunk, items - i0);
            hipLaunchKernelGGL(k_bake_dense_seg, dim3(uint32_t((n + 3) / 4)), dim3(256), 0, st, c);
        }
        hipLaunchKernelGGL(k_bake_dense_finish, dim3(uint32_t((nvox + 255) / 256)), dim3(256), 0, st, a);
    } else {
        // coarse pass: bounds, then bricks in chunks whose per-segment lists fit `part`
        const size_t nb = size_t((a.w + 3) / 4) * ((a.h + 3) / 4) * ((a.d + 3) / 4);
        const size_t nseg = (a.nclusters + SEG - 1) / SEG;
        if (a.part_bricks == 0) return hipErrorInvalidValue;
        hipLaunchKernelGGL(k_bake_coarse_bound, dim3(uint32_t((nb + 3) / 4)), dim3(256), 0, st, a);
        for (size_t b0 = 0; b0 < nb; b0 += a.part_bricks) {
            DenseArgs c = a;
            c.brick0 = uint32_t(b0);
            c.brick_count = uint32_t(std::min<size_t>(a.part_bricks, nb - b0));
            hipLaunchKernelGGL(k_bake_coarse_seg, dim3(uint32_t((size_t(c.brick_count) * nseg + 3) / 4)), dim3(256), 0,
                               st, c);
            hipLaunchKernelGGL(k_bake_coarse_merge, dim3((c.brick_count * 64u + 255u) / 256u), dim3(256), 0, st, c);
        }
    }
    return hipGetLastError();
}

hipError_t launch_bake_fine(hipStream_t st, const FineArgs &a) {
    hipLaunchKernelGGL(k_bake_fine, dim3((a.w + 7) / 8, (a.h + 7) / 8, (a.d + 3) / 4), dim3(256), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_bake_resolve(hipStream_t st, const ResolveArgs &a) {
    if (a.count == 0) return hipSuccess;
    hipLaunchKernelGGL(k_bake_resolve, dim3((a.count + 3u) / 4u), dim3(256), 0, st, a);
    return hipGetLastError();
}

}  // namespace sdf
}  // namespace tc

// ------------------------------------------------------------------ C ABI
#include "tc_sim_impl.h"

struct tc_sdf_result {
    uint32_t w = 0, h = 0, d = 0;
    float bone_info[20] = {0};
    float coarse_voxel = 0.0f;
    float kernel_ms[2] = {0.0f, 0.0f};  // device time of the walk (dense or coarse) and of the fine pass
    uint32_t fallback_voxels = 0;       // voxels the fine pass left to the exhaustive fallback
    std::vector<uint32_t> texels, nearest;
    std::vector<float> atlas;           // method 3: fp32 brick atlas
    uint32_t brick_dims[3] = {0, 0, 0};
};

namespace {

using tc::sdf::BakeTri;

struct EventPair {  // device-side timing of one launch on the null stream
    hipEvent_t a = nullptr, b = nullptr;
    ~EventPair() {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
    hipError_t init() {
        hipError_t e = hipEventCreate(&a);
        return e == hipSuccess ? hipEventCreate(&b) : e;
    }
    float ms() const {
        float t = 0.0f;
        return hipEventElapsedTime(&t, a, b) == hipSuccess ? t : 0.0f;
    }
};

struct DevAlloc {  // device buffers of one bake, freed on every exit path
    std::vector<void *> ptrs;
    ~DevAlloc() { for (void *p : ptrs) (void)hipFree(p); }
    template <typename T>
    hipError_t get(T **p, size_t count) {
        hipError_t e = hipMalloc(reinterpret_cast<void **>(p), sizeof(T) * std::max<size_t>(count, 1));
        if (e == hipSuccess) ptrs.push_back(*p);
        return e;
    }
};

// Bake triangles on the device: the Morton-ordered clusters the traversals walk and the original
// order the fine pass indexes.
struct DevTris {
    DevAlloc mem;
    BakeTri *sorted = nullptr, *orig = nullptr;
    uint32_t *index = nullptr;
    float4 *lo = nullptr, *hi = nullptr;
    uint32_t n = 0, ncl = 0;
};

int upload_tris(const std::vector<BakeTri> &tris, const tc::sdf::Domain &dm, DevTris &d) {
    const tc::sdf::Clusters cl = tc::sdf::build_clusters(tris, dm);
    d.n = uint32_t(tris.size());
    d.ncl = uint32_t(cl.lo.size());
    TC_HIP(d.mem.get(&d.sorted, tris.size()));
    TC_HIP(d.mem.get(&d.orig, tris.size()));
    TC_HIP(d.mem.get(&d.index, tris.size()));
    TC_HIP(d.mem.get(&d.lo, cl.lo.size()));
    TC_HIP(d.mem.get(&d.hi, cl.hi.size()));
    TC_HIP(hipMemcpy(d.sorted, cl.tris.data(), sizeof(BakeTri) * tris.size(), hipMemcpyHostToDevice));
    TC_HIP(hipMemcpy(d.orig, tris.data(), sizeof(BakeTri) * tris.size(), hipMemcpyHostToDevice));
    TC_HIP(hipMemcpy(d.index, cl.index.data(), 4 * tris.size(), hipMemcpyHostToDevice));
    TC_HIP(hipMemcpy(d.lo, cl.lo.data(), 16 * cl.lo.size(), hipMemcpyHostToDevice));
    TC_HIP(hipMemcpy(d.hi, cl.hi.data(), 16 * cl.hi.size(), hipMemcpyHostToDevice));
    return TC_OK;
}

int check_memory(const tc::sdf::Domain &dm, size_t bytes_per_voxel, size_t ntris) {
    const size_t nvox = size_t(dm.w) * dm.h * dm.d;
    size_t free_b = 0, total_b = 0;
    TC_HIP(hipMemGetInfo(&free_b, &total_b));
    const size_t need = nvox * bytes_per_voxel + ntris * 160;
    if (need > free_b / 10 * 9)
        return fail(TC_ERR_RUNTIME, "SDF resolution " + std::to_string(dm.w) + "x" + std::to_string(dm.h) + "x" +
                                        std::to_string(dm.d) + " needs " + std::to_string(need >> 20) +
                                        " MiB of device memory; use a larger voxel size");
    return TC_OK;
}

// The brick traversal (topk 1: dense bake, 8: coarse pass) over domain dm into device buffers.
// Per-segment candidate lists of the coarse pass: 16 keys per cell and cluster segment, for as
// many 4x4x4 bricks at a time as fit in 1 GiB.
int alloc_part(DevAlloc &mem, const tc::sdf::Domain &cdm, uint32_t ncl, unsigned long long *&part,
               uint32_t &part_bricks) {
    const size_t nb = size_t((cdm.w + 3) / 4) * ((cdm.h + 3) / 4) * ((cdm.d + 3) / 4);
    const size_t nseg = std::max<size_t>(1, (ncl + tc::sdf::SEG_CLUSTERS - 1) / tc::sdf::SEG_CLUSTERS);
    const size_t per_brick = size_t(64) * nseg * 16;  // keys
    const size_t budget = (size_t(1) << 30) / 8;
    part_bricks = uint32_t(std::max<size_t>(1, std::min(nb, budget / per_brick)));
    TC_HIP(mem.get(&part, size_t(part_bricks) * per_brick));
    return TC_OK;
}

int launch_walk(const DevTris &t, const tc::sdf::Domain &dm, uint32_t topk, uint32_t *dout, uint32_t *dnear,
                unsigned long long *keys, unsigned long long *part, uint32_t part_bricks, float &ms,
                float *atlas = nullptr) {
    tc::sdf::DenseArgs a;
    a.tris = t.sorted; a.otris = t.orig; a.index = t.index; a.clo = t.lo; a.chi = t.hi;
    a.ntris = t.n; a.nclusters = t.ncl;
    for (int k = 0; k < 3; ++k) { a.lmin[k] = dm.min[k]; a.lmax[k] = dm.max[k]; }
    a.w = dm.w; a.h = dm.h; a.d = dm.d;
    a.topk = topk;
    a.out = dout;
    a.nearest = dnear;
    a.keys = keys;
    a.item0 = 0;
    a.part = part;
    a.part_bricks = part_bricks;
    a.brick0 = 0;
    a.brick_count = 0;
    EventPair ev;
    TC_HIP(ev.init());
    TC_HIP(hipEventRecord(ev.a, nullptr));
    TC_HIP(tc::sdf::launch_bake_dense(nullptr, a));
    if (atlas) TC_HIP(tc::sdf::launch_bake_atlas(nullptr, a, atlas));
    TC_HIP(hipEventRecord(ev.b, nullptr));
    TC_HIP(hipDeviceSynchronize());
    ms = ev.ms();
    return TC_OK;
}

int download(tc_sdf_result &r, const tc::sdf::Domain &dm, const uint32_t *dout, const uint32_t *dnear) {
    const size_t nvox = size_t(dm.w) * dm.h * dm.d;
    r.w = dm.w; r.h = dm.h; r.d = dm.d;
    r.texels.resize(nvox);
    TC_HIP(hipMemcpy(r.texels.data(), dout, 4 * nvox, hipMemcpyDeviceToHost));
    if (dnear) {
        r.nearest.resize(nvox * 8);
        TC_HIP(hipMemcpy(r.nearest.data(), dnear, 4 * nvox * 8, hipMemcpyDeviceToHost));
    }
    return TC_OK;
}

int bake_dense(const std::vector<BakeTri> &bt, const tc::sdf::Domain &dm, uint32_t topk, tc_sdf_result &r,
               bool with_atlas = false) {
    const size_t nvox = size_t(dm.w) * dm.h * dm.d;
    const uint32_t bx = (dm.w + 7) / 8, by = (dm.h + 7) / 8, bz = (dm.d + 7) / 8;
    const size_t natlas = with_atlas ? size_t(bx) * by * bz * 512 : 0;
    if (int rc = check_memory(dm, (topk == 8 ? 44 : 12) + (with_atlas ? 8 : 0), bt.size())) return rc;
    DevTris t;
    if (int rc = upload_tris(bt, dm, t)) return rc;
    DevAlloc mem;
    uint32_t *dout = nullptr, *dnear = nullptr;
    unsigned long long *keys = nullptr, *part = nullptr;
    float *datlas = nullptr;
    uint32_t part_bricks = 0;
    TC_HIP(mem.get(&dout, nvox));
    TC_HIP(mem.get(&keys, nvox));
    if (with_atlas) TC_HIP(mem.get(&datlas, natlas));
    if (topk == 8) {
        TC_HIP(mem.get(&dnear, nvox * 8));
        if (int rc = alloc_part(mem, dm, t.ncl, part, part_bricks)) return rc;
    }
    if (int rc = launch_walk(t, dm, topk, dout, dnear, keys, part, part_bricks, r.kernel_ms[0], datlas)) return rc;
    if (with_atlas) {
        r.atlas.resize(natlas);
        r.brick_dims[0] = bx; r.brick_dims[1] = by; r.brick_dims[2] = bz;
        TC_HIP(hipMemcpy(r.atlas.data(), datlas, sizeof(float) * natlas, hipMemcpyDeviceToHost));
    }
    return download(r, dm, dout, dnear);
}

// bake_mesh_sdf_hierarchical_ext (sdf_baker.rs:1625-1882): coarse top-8 pass, fine pass, and the
// fine pass's exhaustive fallback for the voxels it leaves.
int bake_hierarchical(const std::vector<BakeTri> &bt, const tc::sdf::Domain &dm, const tc::sdf::Domain &cdm,
                      tc_sdf_result &r) {
    const float h_fine = std::fmax(std::fmax(dm.size[0] / float(dm.w), dm.size[1] / float(dm.h)), dm.size[2] / float(dm.d));
    const float band = 4.0f * h_fine;
    const float h_coarse_min =
        std::fmin(std::fmin(cdm.size[0] / float(cdm.w), cdm.size[1] / float(cdm.h)), cdm.size[2] / float(cdm.d));
    if (band > 2.0f * h_coarse_min)
        return fail(TC_ERR_RUNTIME, "hierarchical bake: band " + std::to_string(band) +
                                        " m is too wide for coarse cells of " + std::to_string(h_coarse_min) + " m");
    const size_t nvox = size_t(dm.w) * dm.h * dm.d, ncvox = size_t(cdm.w) * cdm.h * cdm.d;
    if (int rc = check_memory(dm, 9, bt.size())) return rc;  // out + fallback list, coarse grid ~1/64
    DevTris t;
    if (int rc = upload_tris(bt, dm, t)) return rc;
    DevAlloc mem;
    uint32_t *dcd = nullptr, *dct = nullptr, *dout = nullptr, *damb = nullptr, *dcount = nullptr;
    unsigned long long *ckeys = nullptr, *part = nullptr;
    uint32_t part_bricks = 0;
    TC_HIP(mem.get(&dcd, ncvox));
    TC_HIP(mem.get(&dct, ncvox * 8));
    TC_HIP(mem.get(&ckeys, ncvox));
    if (int rc = alloc_part(mem, cdm, t.ncl, part, part_bricks)) return rc;
    TC_HIP(mem.get(&dout, nvox));
    TC_HIP(mem.get(&damb, nvox));
    TC_HIP(mem.get(&dcount, 1));
    TC_HIP(hipMemset(dcount, 0, 4));
    if (int rc = launch_walk(t, cdm, 8, dcd, dct, ckeys, part, part_bricks, r.kernel_ms[0])) return rc;
    tc::sdf::FineArgs fa;
    fa.tris = t.orig; fa.ntris = t.n;
    fa.coarse_dist = dcd; fa.coarse_tri = dct;
    fa.cw = cdm.w; fa.ch = cdm.h; fa.cd = cdm.d;
    fa.band = band;
    for (int k = 0; k < 3; ++k) { fa.lmin[k] = dm.min[k]; fa.lmax[k] = dm.max[k]; }
    fa.w = dm.w; fa.h = dm.h; fa.d = dm.d;
    fa.out = dout;
    fa.amb = damb;
    fa.amb_count = dcount;
    EventPair ev;
    TC_HIP(ev.init());
    TC_HIP(hipEventRecord(ev.a, nullptr));
    TC_HIP(tc::sdf::launch_bake_fine(nullptr, fa));
    uint32_t count = 0;
    TC_HIP(hipMemcpy(&count, dcount, 4, hipMemcpyDeviceToHost));
    tc::sdf::ResolveArgs ra;
    ra.tris = t.sorted; ra.index = t.index; ra.clo = t.lo; ra.chi = t.hi;
    ra.ntris = t.n; ra.nclusters = t.ncl;
    ra.list = damb; ra.count = count;
    for (int k = 0; k < 3; ++k) { ra.lmin[k] = dm.min[k]; ra.lmax[k] = dm.max[k]; }
    ra.w = dm.w; ra.h = dm.h; ra.d = dm.d;
    ra.out = dout;
    TC_HIP(tc::sdf::launch_bake_resolve(nullptr, ra));
    TC_HIP(hipEventRecord(ev.b, nullptr));
    TC_HIP(hipDeviceSynchronize());
    r.kernel_ms[1] = ev.ms();
    r.fallback_voxels = count;
    return download(r, dm, dout, nullptr);
}

}  // namespace

int tc_bake_mesh_sdf(const float *verts, uint32_t nv, const int32_t *tris, uint32_t nt, float voxel_size,
                     float margin, float friction, float thickness, float restitution, int close_open_boundaries,
                     int method, tc_sdf_result **out) {
    if (!out) return fail(TC_ERR_VALUE, "null output pointer");
    *out = nullptr;
    if (method < 0 || method > 3)
        return fail(TC_ERR_VALUE, "method must be 0 (dense), 1 (hierarchical), 2 (coarse) or 3 (dense + fp32 atlas)");
    if (nv == 0 || !verts) return fail(TC_ERR_RUNTIME, "the mesh has no vertices");
    if (nt && !tris) return fail(TC_ERR_VALUE, "null triangle array");
    std::vector<float> v(verts, verts + 3 * size_t(nv));
    std::vector<int32_t> t(tris, tris + 3 * size_t(nt));
    for (float x : v)
        if (!std::isfinite(x)) return fail(TC_ERR_VALUE, "vertex coordinates must be finite");
    if (close_open_boundaries) tc::sdf::close_open_boundary_loops(v, t);
    std::unique_ptr<tc_sdf_result> r(new tc_sdf_result());
    const std::vector<BakeTri> bt = tc::sdf::collect_bake_triangles(v, t);
    const float cv = std::fmax(voxel_size * 4.0f, 0.001f);  // coarse cell: 4 voxels
    int rc = TC_OK;
    if (method == 0 || method == 3) {
        const tc::sdf::Domain dm = tc::sdf::mesh_sdf_domain(v, voxel_size, margin, thickness);
        tc::sdf::mesh_sdf_bone_info(dm, friction, thickness, restitution, r->bone_info);
        rc = bake_dense(bt, dm, 1, *r, method == 3);
    } else if (method == 2) {
        const tc::sdf::Domain dm = tc::sdf::mesh_sdf_domain(v, cv, margin, thickness);
        tc::sdf::mesh_sdf_bone_info(dm, friction, thickness, restitution, r->bone_info);
        r->coarse_voxel = cv;
        rc = bake_dense(bt, dm, 8, *r);
    } else {
        const tc::sdf::Domain dm = tc::sdf::mesh_sdf_domain(v, voxel_size, margin, thickness);
        tc::sdf::mesh_sdf_bone_info(dm, friction, thickness, restitution, r->bone_info);
        rc = bake_hierarchical(bt, dm, tc::sdf::mesh_sdf_domain(v, cv, margin, thickness), *r);
    }
    if (rc != TC_OK) return rc;
    *out = r.release();
    return TC_OK;
}

const float *tc_sdf_result_atlas(const tc_sdf_result *r, uint32_t *brick_dims) {
    if (!r || r->atlas.empty()) return nullptr;
    if (brick_dims) { brick_dims[0] = r->brick_dims[0]; brick_dims[1] = r->brick_dims[1]; brick_dims[2] = r->brick_dims[2]; }
    return r->atlas.data();
}

int tc_sdf_result_info(const tc_sdf_result *r, uint32_t *dims, float *bone_info, float *coarse_voxel) {
    if (!r) return fail(TC_ERR_VALUE, "null result");
    if (dims) { dims[0] = r->w; dims[1] = r->h; dims[2] = r->d; }
    if (bone_info) std::memcpy(bone_info, r->bone_info, sizeof(r->bone_info));
    if (coarse_voxel) *coarse_voxel = r->coarse_voxel;
    return TC_OK;
}

int tc_sdf_result_timing(const tc_sdf_result *r, float *kernel_ms, uint32_t *fallback_voxels) {
    if (!r) return fail(TC_ERR_VALUE, "null result");
    if (kernel_ms) {
        kernel_ms[0] = r->kernel_ms[0];
        kernel_ms[1] = r->kernel_ms[1];
    }
    if (fallback_voxels) *fallback_voxels = r->fallback_voxels;
    return TC_OK;
}

const uint32_t *tc_sdf_result_texels(const tc_sdf_result *r) { return r && !r->texels.empty() ? r->texels.data() : nullptr; }
const uint32_t *tc_sdf_result_nearest(const tc_sdf_result *r) {
    return r && !r->nearest.empty() ? r->nearest.data() : nullptr;
}
void tc_sdf_result_free(tc_sdf_result *r) { delete r; }

int tc_close_open_boundary_loops(const float *verts, uint32_t nv, const int32_t *tris, uint32_t nt, float *out_verts,
                                 size_t cap_verts, int32_t *out_tris, size_t cap_tris, uint32_t *nv_out,
                                 uint32_t *nt_out, uint32_t *loops) {
    if ((!verts && nv) || (!tris && nt)) return fail(TC_ERR_VALUE, "null input");
    std::vector<float> v(verts, verts + 3 * size_t(nv));
    std::vector<int32_t> t(tris, tris + 3 * size_t(nt));
    const size_t closed = tc::sdf::close_open_boundary_loops(v, t);
    if (nv_out) *nv_out = uint32_t(v.size() / 3);
    if (nt_out) *nt_out = uint32_t(t.size() / 3);
    if (loops) *loops = uint32_t(closed);
    if (out_verts) {
        if (cap_verts < v.size()) return fail(TC_ERR_VALUE, "vertex output too small");
        std::memcpy(out_verts, v.data(), 4 * v.size());
    }
    if (out_tris) {
        if (cap_tris < t.size()) return fail(TC_ERR_VALUE, "triangle output too small");
        std::memcpy(out_tris, t.data(), 4 * t.size());
    }
    return TC_OK;
}
