// Mesh SDF bakes for gfx950 and their host-side preparation.
//
// Dense bake (bake_mesh_sdf_gpu_ext, crates/cloth_core/src/sdf_baker.rs:1114-1323 with
// shaders/bake_mesh_sdf.wgsl:80-121): every voxel centre takes the signed distance to the
// nearest triangle (Ericson's closest point, sign from the face normal, first strict minimum
// in triangle order wins), stored as pack2x16float(d, 1.0) in Z, Y, X order.
//
// Here one wavefront owns a brick of 8x8x8 voxels (8 per lane).  Triangles are sorted along
// a Morton curve and grouped into clusters of 32 with a box each; a wavefront walks the
// clusters outward from its brick's Morton position and skips every cluster whose box lies
// farther from the brick than the worst current best of its voxels.  Candidates compare by
// (squared distance, original triangle index), so the result is the same argmin the
// reference's in-order scan returns, whatever order the clusters are visited in.
//
// The coarse pass of the hierarchical bake (bake_mesh_sdf_hier_coarse_main.wgsl) is the same
// walk with one voxel per lane (4x4x4 bricks) and the eight nearest triangles per voxel;
// the fine pass (bake_mesh_sdf_hier_fine_main.wgsl) is restated voxel by voxel.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <map>
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

size_t close_open_boundary_loops(std::vector<float> &verts, std::vector<int32_t> &tris) {
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
    for (int a = 0; a < 3; ++a) c.origin[a] = dm.min[a];
    c.scale = 1023.0f / std::max(ext, 1e-6f);
    std::vector<uint32_t> code(tris.size());
    for (size_t t = 0; t < tris.size(); ++t) {
        const BakeTri &b = tris[t];
        const float cx = (b.p0x + b.p1x + b.p2x) / 3.0f, cy = (b.p0y + b.p1y + b.p2y) / 3.0f,
                    cz = (b.p0z + b.p1z + b.p2z) / 3.0f;
        auto q = [&](float v, int a) {
            const float f = (v - c.origin[a]) * c.scale;
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
    c.code.resize(nc);
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
        c.code[ci] = code[order[ci * CLUSTER]];
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

__device__ __forceinline__ float wave_max_f(float v) {
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
    return v;
}

__device__ __forceinline__ uint32_t morton_of(f3 p, const DenseArgs &a) {
    auto q = [&](float v, int ax) {
        const float f = (v - a.mo[ax]) * a.ms;
        return uint32_t(fminf(1023.0f, fmaxf(0.0f, f)));
    };
    auto spread = [](uint32_t v) {
        v &= 1023u;
        v = (v | (v << 16)) & 0x030000FFu;
        v = (v | (v << 8)) & 0x0300F00Fu;
        v = (v | (v << 4)) & 0x030C30C3u;
        v = (v | (v << 2)) & 0x09249249u;
        return v;
    };
    return spread(q(p.x, 0)) | (spread(q(p.y, 1)) << 1) | (spread(q(p.z, 2)) << 2);
}

// (d_sq, index) lexicographic "less": the in-order first strict minimum
__device__ __forceinline__ bool better(float d, uint32_t i, float bd, uint32_t bi) { return d < bd || (d == bd && i < bi); }

// The coarse pass's list update (bake_mesh_sdf_hier_coarse_main.wgsl:41-60): strict comparisons,
// so an entry pushed down the list passes over the entries of equal distance below it.
__device__ __forceinline__ void ref_insert(float (&sq)[8], float (&sg)[8], uint32_t (&id)[8], float d, float s,
                                           uint32_t i) {
    if (!(d < sq[7])) return;
    float is = d, ig = s;
    uint32_t ii = i;
#pragma unroll
    for (int k = 0; k < 8; ++k)
        if (is < sq[k]) {
            const float ts = sq[k], tg = sg[k];
            const uint32_t ti = id[k];
            sq[k] = is; sg[k] = ig; id[k] = ii;
            is = ts; ig = tg; ii = ti;
        }
}

constexpr uint32_t NO_TRI = 0xFFFFFFFFu;

// VPL voxels per lane (brick = 8 x 8 x VPL for VPL = 8, 4 x 4 x 4 for VPL = 1), K nearest kept.
// K = 1: the dense bake.  K = 16: the coarse pass; the walk keeps the 16 nearest in (d_sq, index)
// order, which holds every triangle at or below the eighth distance D8 unless more than 16 do.
// The reference's list is the result of its update rule applied in triangle order, and only
// triangles at or below D8 can be in it or move an entry of it, so replaying exactly those in
// index order gives the reference's list; with no equal distances among them the replay is
// the identity and is skipped.  More than 16 at or below D8 replays every triangle.
template <int VPL, int K>
__global__ void __launch_bounds__(256) k_bake_walk(DenseArgs a) {
    constexpr int BX = VPL == 8 ? 8 : 4, BY = VPL == 8 ? 8 : 4, BZ = VPL == 8 ? 8 : 4;
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t brick = blockIdx.x * 4u + (threadIdx.x >> 6);
    const uint32_t nbx = (a.w + BX - 1) / BX, nby = (a.h + BY - 1) / BY, nbz = (a.d + BZ - 1) / BZ;
    if (brick >= nbx * nby * nbz) return;  // whole wavefront
    const uint32_t bx = brick % nbx, by = (brick / nbx) % nby, bz = brick / (nbx * nby);
    // lane's voxels
    uint32_t vx, vy, vz0;
    if (VPL == 8) { vx = bx * 8 + (lane & 7u); vy = by * 8 + (lane >> 3); vz0 = bz * 8; }
    else { vx = bx * 4 + (lane & 3u); vy = by * 4 + ((lane >> 2) & 3u); vz0 = bz * 4 + (lane >> 4); }
    const bool col_ok = vx < a.w && vy < a.h;
    f3 p[VPL];
    bool ok[VPL];
    for (int j = 0; j < VPL; ++j) {
        const uint32_t vz = vz0 + j;
        ok[j] = col_ok && vz < a.d;
        p[j] = mk3(voxel_coord(vx, a.w, a.lmin[0], a.lmax[0]), voxel_coord(vy, a.h, a.lmin[1], a.lmax[1]),
                   voxel_coord(vz, a.d, a.lmin[2], a.lmax[2]));
    }
    // brick box over its valid voxel centres
    const uint32_t x0 = bx * BX, y0 = by * BY, z0 = bz * BZ;
    const uint32_t x1 = min(a.w, x0 + BX) - 1, y1 = min(a.h, y0 + BY) - 1, z1 = min(a.d, z0 + BZ) - 1;
    const f3 bmin = mk3(voxel_coord(x0, a.w, a.lmin[0], a.lmax[0]), voxel_coord(y0, a.h, a.lmin[1], a.lmax[1]),
                        voxel_coord(z0, a.d, a.lmin[2], a.lmax[2]));
    const f3 bmax = mk3(voxel_coord(x1, a.w, a.lmin[0], a.lmax[0]), voxel_coord(y1, a.h, a.lmin[1], a.lmax[1]),
                        voxel_coord(z1, a.d, a.lmin[2], a.lmax[2]));
    float bsq[VPL][K], bsg[VPL][K];
    uint32_t bid[VPL][K];
    for (int j = 0; j < VPL; ++j)
        for (int k = 0; k < K; ++k) { bsq[j][k] = 1e10f; bsg[j][k] = 10.0f; bid[j][k] = NO_TRI; }
    // start at the brick centre's position along the Morton curve
    const uint32_t mc = morton_of((bmin + bmax) * 0.5f, a);
    uint32_t lo = 0, hi = a.nclusters;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (a.ccode[mid] < mc) lo = mid + 1; else hi = mid;
    }
    const int32_t c0 = int32_t(min(lo, a.nclusters - 1));
    float bound = 3.0e38f;
    int32_t up = c0, down = c0 - 1;
    for (uint32_t step = 0; up < int32_t(a.nclusters) || down >= 0; ++step) {
        int32_t c;
        if ((step & 1u) == 0u) { if (up < int32_t(a.nclusters)) c = up++; else c = down--; }
        else { if (down >= 0) c = down--; else c = up++; }
        if ((step & 3u) == 0u) {
            float worst = -1.0f;
            for (int j = 0; j < VPL; ++j)
                if (ok[j]) worst = fmaxf(worst, bsq[j][K - 1]);
            bound = wave_max_f(worst);
        }
        const float4 cl = a.clo[c], ch = a.chi[c];
        // gap between the boxes, shortened by 1e-5 m and 1e-4 relative so that rounding of the
        // exact distances can never make a skipped triangle the better candidate
        const float gx = fmaxf(0.0f, fmaxf(cl.x - bmax.x, bmin.x - ch.x) - 1e-5f);
        const float gy = fmaxf(0.0f, fmaxf(cl.y - bmax.y, bmin.y - ch.y) - 1e-5f);
        const float gz = fmaxf(0.0f, fmaxf(cl.z - bmax.z, bmin.z - ch.z) - 1e-5f);
        const float lb = (gx * gx + gy * gy + gz * gz) * 0.9999f;
        if (lb > bound) continue;
        const uint32_t t1 = min(a.ntris, uint32_t(c + 1) * CLUSTER);
        for (uint32_t t = uint32_t(c) * CLUSTER; t < t1; ++t) {
            const BakeTri tr = a.tris[t];
            const uint32_t idx = a.index[t];
            const f3 A = mk3(tr.p0x, tr.p0y, tr.p0z), B = mk3(tr.p1x, tr.p1y, tr.p1z), C = mk3(tr.p2x, tr.p2y, tr.p2z);
            const f3 N = mk3(tr.nx, tr.ny, tr.nz);
            for (int j = 0; j < VPL; ++j) {
                const f3 q = closest_on_tri(p[j], A, B, C);
                const f3 diff = p[j] - q;
                const float dsq = dot(diff, diff);
                if (!better(dsq, idx, bsq[j][K - 1], bid[j][K - 1])) continue;
                const float sg = (dot(diff, N) >= 0.0f ? 1.0f : -1.0f) * sqrtf(dsq);
                if (K == 1) {
                    bsq[j][0] = dsq; bsg[j][0] = sg; bid[j][0] = idx;
                } else {
                    float is = dsq, ig = sg;
                    uint32_t ii = idx;
                    for (int k = 0; k < K; ++k)
                        if (better(is, ii, bsq[j][k], bid[j][k])) {
                            const float ts = bsq[j][k], tg = bsg[j][k];
                            const uint32_t ti = bid[j][k];
                            bsq[j][k] = is; bsg[j][k] = ig; bid[j][k] = ii;
                            is = ts; ig = tg; ii = ti;
                        }
                }
            }
        }
    }
    for (int j = 0; j < VPL; ++j) {
        if (!ok[j]) continue;
        const size_t flat = (size_t(vz0 + j) * a.h + vy) * a.w + vx;
        a.out[flat] = pack_dist(bsg[j][0]);
        if constexpr (K > 1) {
            if (!a.nearest) continue;
            float rsq[8], rsg[8];
            uint32_t rid[8];
            for (int k = 0; k < 8; ++k) { rsq[k] = bsq[j][k]; rsg[k] = bsg[j][k]; rid[k] = bid[j][k]; }
            bool tie = false;
            for (int k = 0; k < 8; ++k) tie = tie || bsq[j][k] == bsq[j][k + 1];
            if (tie) {
                const float d8 = bsq[j][7];
                for (int k = 0; k < 8; ++k) { rsq[k] = 1e10f; rsg[k] = 10.0f; rid[k] = 0u; }
                if (bsq[j][K - 1] > d8) {
                    uint32_t used = 0u;
                    for (int n = 0; n < K; ++n) {  // entries at or below D8 in increasing triangle index
                        uint32_t pick = NO_TRI;
                        float pd = 0.0f, ps = 0.0f;
                        int pk = -1;
                        for (int k = 0; k < K; ++k)
                            if (!((used >> k) & 1u) && bsq[j][k] <= d8 && bid[j][k] < pick) {
                                pick = bid[j][k]; pd = bsq[j][k]; ps = bsg[j][k]; pk = k;
                            }
                        if (pk < 0) break;
                        used |= 1u << pk;
                        ref_insert(rsq, rsg, rid, pd, ps, pick);
                    }
                } else {
                    for (uint32_t i = 0; i < a.ntris; ++i) {
                        const BakeTri tr = a.otris[i];
                        const f3 A = mk3(tr.p0x, tr.p0y, tr.p0z), B = mk3(tr.p1x, tr.p1y, tr.p1z);
                        const f3 C = mk3(tr.p2x, tr.p2y, tr.p2z);
                        const f3 q = closest_on_tri(p[j], A, B, C);
                        const f3 diff = p[j] - q;
                        const float dsq = dot(diff, diff);
                        if (!(dsq < rsq[7])) continue;
                        ref_insert(rsq, rsg, rid, dsq, (dot(diff, mk3(tr.nx, tr.ny, tr.nz)) >= 0.0f ? 1.0f : -1.0f) * sqrtf(dsq),
                                   i);
                    }
                }
            }
            for (int k = 0; k < 8; ++k) a.nearest[flat * 8 + k] = rid[k] == NO_TRI ? 0u : rid[k];
        }
    }
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
    if (ambiguous) {
        best_sq = 1e10f;
        best_sg = 10.0f;
        for (uint32_t i = 0; i < a.ntris; ++i) {
            const BakeTri tr = a.tris[i];
            const f3 A = mk3(tr.p0x, tr.p0y, tr.p0z), B = mk3(tr.p1x, tr.p1y, tr.p1z), C = mk3(tr.p2x, tr.p2y, tr.p2z);
            const f3 q = closest_on_tri(p, A, B, C);
            const f3 diff = p - q;
            const float dsq = dot(diff, diff);
            if (dsq < best_sq) {
                best_sq = dsq;
                best_sg = sqrtf(dsq) * (dot(diff, mk3(tr.nx, tr.ny, tr.nz)) >= 0.0f ? 1.0f : -1.0f);
            }
        }
    }
    a.out[flat] = pack_dist(best_sg);
}

hipError_t launch_bake_dense(hipStream_t st, const DenseArgs &a) {
    if (a.topk == 1) {
        const size_t nb = size_t((a.w + 7) / 8) * ((a.h + 7) / 8) * ((a.d + 7) / 8);
        hipLaunchKernelGGL((k_bake_walk<8, 1>), dim3(uint32_t((nb + 3) / 4)), dim3(256), 0, st, a);
    } else {
        const size_t nb = size_t((a.w + 3) / 4) * ((a.h + 3) / 4) * ((a.d + 3) / 4);
        hipLaunchKernelGGL((k_bake_walk<1, 16>), dim3(uint32_t((nb + 3) / 4)), dim3(256), 0, st, a);
    }
    return hipGetLastError();
}

hipError_t launch_bake_fine(hipStream_t st, const FineArgs &a) {
    hipLaunchKernelGGL(k_bake_fine, dim3((a.w + 7) / 8, (a.h + 7) / 8, (a.d + 3) / 4), dim3(256), 0, st, a);
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
    std::vector<uint32_t> texels, nearest;
};

namespace {

using tc::sdf::BakeTri;

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

// Runs the walk bake (topk 1: dense, 8: coarse) of `tris` over domain `dm`; fills r.
int run_walk(const std::vector<BakeTri> &tris, const tc::sdf::Domain &dm, uint32_t topk, tc_sdf_result &r) {
    const size_t nvox = size_t(dm.w) * dm.h * dm.d;
    size_t free_b = 0, total_b = 0;
    TC_HIP(hipMemGetInfo(&free_b, &total_b));
    const size_t need = nvox * 4 * (topk == 8 ? 9 : 1) + tris.size() * 80;
    if (need > free_b / 10 * 9)
        return fail(TC_ERR_RUNTIME, "SDF resolution " + std::to_string(dm.w) + "x" + std::to_string(dm.h) + "x" +
                                        std::to_string(dm.d) + " needs " + std::to_string(need >> 20) +
                                        " MiB of device memory; use a larger voxel size");
    const tc::sdf::Clusters cl = tc::sdf::build_clusters(tris, dm);
    DevAlloc dev;
    BakeTri *dt = nullptr;
    BakeTri *dot_ = nullptr;
    uint32_t *didx = nullptr, *dcode = nullptr, *dout = nullptr, *dnear = nullptr;
    float4 *dlo = nullptr, *dhi = nullptr;
    TC_HIP(dev.get(&dt, cl.tris.size()));
    TC_HIP(dev.get(&didx, cl.index.size()));
    TC_HIP(dev.get(&dcode, cl.code.size()));
    TC_HIP(dev.get(&dlo, cl.lo.size()));
    TC_HIP(dev.get(&dhi, cl.hi.size()));
    TC_HIP(dev.get(&dout, nvox));
    if (topk == 8) {
        TC_HIP(dev.get(&dnear, nvox * 8));
        TC_HIP(dev.get(&dot_, tris.size()));
        TC_HIP(hipMemcpy(dot_, tris.data(), sizeof(BakeTri) * tris.size(), hipMemcpyHostToDevice));
    }
    TC_HIP(hipMemcpy(dt, cl.tris.data(), sizeof(BakeTri) * cl.tris.size(), hipMemcpyHostToDevice));
    TC_HIP(hipMemcpy(didx, cl.index.data(), 4 * cl.index.size(), hipMemcpyHostToDevice));
    TC_HIP(hipMemcpy(dcode, cl.code.data(), 4 * cl.code.size(), hipMemcpyHostToDevice));
    TC_HIP(hipMemcpy(dlo, cl.lo.data(), 16 * cl.lo.size(), hipMemcpyHostToDevice));
    TC_HIP(hipMemcpy(dhi, cl.hi.data(), 16 * cl.hi.size(), hipMemcpyHostToDevice));
    tc::sdf::DenseArgs a;
    a.tris = dt; a.otris = dot_; a.index = didx; a.clo = dlo; a.chi = dhi; a.ccode = dcode;
    a.ntris = uint32_t(cl.tris.size()); a.nclusters = uint32_t(cl.lo.size());
    for (int k = 0; k < 3; ++k) { a.mo[k] = cl.origin[k]; a.lmin[k] = dm.min[k]; a.lmax[k] = dm.max[k]; }
    a.ms = cl.scale;
    a.w = dm.w; a.h = dm.h; a.d = dm.d;
    a.topk = topk;
    a.out = dout;
    a.nearest = dnear;
    TC_HIP(tc::sdf::launch_bake_dense(nullptr, a));
    TC_HIP(hipDeviceSynchronize());
    r.w = dm.w; r.h = dm.h; r.d = dm.d;
    r.texels.resize(nvox);
    TC_HIP(hipMemcpy(r.texels.data(), dout, 4 * nvox, hipMemcpyDeviceToHost));
    if (topk == 8) {
        r.nearest.resize(nvox * 8);
        TC_HIP(hipMemcpy(r.nearest.data(), dnear, 4 * nvox * 8, hipMemcpyDeviceToHost));
    }
    return TC_OK;
}

}  // namespace

int tc_bake_mesh_sdf(const float *verts, uint32_t nv, const int32_t *tris, uint32_t nt, float voxel_size,
                     float margin, float friction, float thickness, float restitution, int close_open_boundaries,
                     int method, tc_sdf_result **out) {
    if (!out) return fail(TC_ERR_VALUE, "null output pointer");
    *out = nullptr;
    if (method < 0 || method > 2) return fail(TC_ERR_VALUE, "method must be 0 (dense), 1 (hierarchical) or 2 (coarse)");
    if (nv == 0 || !verts) return fail(TC_ERR_RUNTIME, "the mesh has no vertices");
    if (nt && !tris) return fail(TC_ERR_VALUE, "null triangle array");
    std::vector<float> v(verts, verts + 3 * size_t(nv));
    std::vector<int32_t> t(tris, tris + 3 * size_t(nt));
    for (float x : v)
        if (!std::isfinite(x)) return fail(TC_ERR_VALUE, "vertex coordinates must be finite");
    if (close_open_boundaries) tc::sdf::close_open_boundary_loops(v, t);
    tc_sdf_result *r = new tc_sdf_result();
    const std::vector<BakeTri> bt = tc::sdf::collect_bake_triangles(v, t);
    int rc = TC_OK;
    if (method == 0) {
        const tc::sdf::Domain dm = tc::sdf::mesh_sdf_domain(v, voxel_size, margin, thickness);
        tc::sdf::mesh_sdf_bone_info(dm, friction, thickness, restitution, r->bone_info);
        rc = run_walk(bt, dm, 1, *r);
    } else if (method == 2) {
        const float cv = std::fmax(voxel_size * 4.0f, 0.001f);
        const tc::sdf::Domain dm = tc::sdf::mesh_sdf_domain(v, cv, margin, thickness);
        tc::sdf::mesh_sdf_bone_info(dm, friction, thickness, restitution, r->bone_info);
        r->coarse_voxel = cv;
        rc = run_walk(bt, dm, 8, *r);
    } else {
        // bake_mesh_sdf_hierarchical_ext (sdf_baker.rs:1625-1882)
        const tc::sdf::Domain dm = tc::sdf::mesh_sdf_domain(v, voxel_size, margin, thickness);
        tc::sdf::mesh_sdf_bone_info(dm, friction, thickness, restitution, r->bone_info);
        const float cv = std::fmax(voxel_size * 4.0f, 0.001f);
        const tc::sdf::Domain cdm = tc::sdf::mesh_sdf_domain(v, cv, margin, thickness);
        tc_sdf_result coarse;
        rc = run_walk(bt, cdm, 8, coarse);
        if (rc == TC_OK) {
            const float h_fine = std::fmax(std::fmax(dm.size[0] / float(dm.w), dm.size[1] / float(dm.h)),
                                           dm.size[2] / float(dm.d));
            const float band = 4.0f * h_fine;
            const float h_coarse_min = std::fmin(std::fmin(cdm.size[0] / float(cdm.w), cdm.size[1] / float(cdm.h)),
                                                 cdm.size[2] / float(cdm.d));
            if (band > 2.0f * h_coarse_min) {
                rc = fail(TC_ERR_RUNTIME, "hierarchical bake: band " + std::to_string(band) +
                                              " m is too wide for coarse cells of " + std::to_string(h_coarse_min) + " m");
            } else {
                const size_t nvox = size_t(dm.w) * dm.h * dm.d;
                DevAlloc dev;
                BakeTri *dt = nullptr;
                uint32_t *dcd = nullptr, *dct = nullptr, *dout = nullptr;
                hipError_t e = hipSuccess;
                if ((e = dev.get(&dt, bt.size())) == hipSuccess && (e = dev.get(&dcd, coarse.texels.size())) == hipSuccess &&
                    (e = dev.get(&dct, coarse.nearest.size())) == hipSuccess && (e = dev.get(&dout, nvox)) == hipSuccess &&
                    (e = hipMemcpy(dt, bt.data(), sizeof(BakeTri) * bt.size(), hipMemcpyHostToDevice)) == hipSuccess &&
                    (e = hipMemcpy(dcd, coarse.texels.data(), 4 * coarse.texels.size(), hipMemcpyHostToDevice)) == hipSuccess &&
                    (e = hipMemcpy(dct, coarse.nearest.data(), 4 * coarse.nearest.size(), hipMemcpyHostToDevice)) == hipSuccess) {
                    tc::sdf::FineArgs fa;
                    fa.tris = dt; fa.ntris = uint32_t(bt.size());
                    fa.coarse_dist = dcd; fa.coarse_tri = dct;
                    fa.cw = cdm.w; fa.ch = cdm.h; fa.cd = cdm.d;
                    fa.band = band;
                    for (int k = 0; k < 3; ++k) { fa.lmin[k] = dm.min[k]; fa.lmax[k] = dm.max[k]; }
                    fa.w = dm.w; fa.h = dm.h; fa.d = dm.d;
                    fa.out = dout;
                    if ((e = tc::sdf::launch_bake_fine(nullptr, fa)) == hipSuccess && (e = hipDeviceSynchronize()) == hipSuccess) {
                        r->w = dm.w; r->h = dm.h; r->d = dm.d;
                        r->texels.resize(nvox);
                        e = hipMemcpy(r->texels.data(), dout, 4 * nvox, hipMemcpyDeviceToHost);
                    }
                }
                if (e != hipSuccess) rc = fail(TC_ERR_RUNTIME, std::string("hierarchical bake: ") + hipGetErrorString(e));
            }
        }
    }
    if (rc != TC_OK) {
        delete r;
        return rc;
    }
    *out = r;
    return TC_OK;
}

int tc_sdf_result_info(const tc_sdf_result *r, uint32_t *dims, float *bone_info, float *coarse_voxel) {
    if (!r) return fail(TC_ERR_VALUE, "null result");
    if (dims) { dims[0] = r->w; dims[1] = r->h; dims[2] = r->d; }
    if (bone_info) std::memcpy(bone_info, r->bone_info, sizeof(r->bone_info));
    if (coarse_voxel) *coarse_voxel = r->coarse_voxel;
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
