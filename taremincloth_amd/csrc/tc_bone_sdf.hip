// Bone-SDF atlas bake: per bone, the signed distance to the bone's skinned triangles in the bone's
// bind (local) frame on a res^3 tile of a 3D atlas, with the bone weight interpolated at the nearest
// point, packed as pack2x16float(distance, weight).
//
// Reference semantics (paths relative to the reference repository root):
//   host preparation  crates/cloth_core/src/sdf_baker.rs:137-163 (compute_3d_atlas_layout),
//                     203-240 (open boundary edges), 242-452 (loop capping with cap weights),
//                     494-716 (active bones, local boxes, tiles, 20-float infos, bake triangles)
//   kernel            crates/cloth_core/src/shaders/bake_sdf.wgsl:34-196 (nearest triangle with the
//                     near-tie rule of Baerentzen & Aanaes 2005: within 1e-7 max(1, d^2) the larger
//                     |normal projection| wins; open-boundary sign fix)
//   entry point       crates/cloth_py/src/lib.rs:1193-1355 (bake_bone_sdf_gpu)
//
// One wavefront per 4x4x4 brick of a bone's tile: the bone's triangles pass through LDS 64 at a time
// and every lane tests each one against its voxel in the reference's order, so every voxel keeps
// the reference's sequence of comparisons.
#include "tc_sim_impl.h"
#include "tc_sdf.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <set>
#include <string>

namespace tc {
namespace bone {

__device__ __forceinline__ f3 v3(const float *p) { return mk3(p[0], p[1], p[2]); }
__device__ __forceinline__ f3 normalize3(f3 v) { return v / sqrtf(dot(v, v)); }

__global__ void __launch_bounds__(64) k_bone_bake(const Params *__restrict__ params, const BakeTri *__restrict__ tris,
                                                  uint32_t *__restrict__ out, uint32_t bricks_per_axis) {
    __shared__ BakeTri s_tri[64];
    const uint32_t per_bone = bricks_per_axis * bricks_per_axis * bricks_per_axis;
    const uint32_t bone = blockIdx.x / per_bone, brick = blockIdx.x % per_bone;
    const Params P = params[bone];
    const uint32_t lane = threadIdx.x;
    const uint32_t bx = brick % bricks_per_axis, by = (brick / bricks_per_axis) % bricks_per_axis,
                   bz = brick / (bricks_per_axis * bricks_per_axis);
    const uint32_t vx = bx * 4u + (lane & 3u), vy = by * 4u + ((lane >> 2) & 3u), vz = bz * 4u + (lane >> 4);
    const bool live = vx < P.res && vy < P.res && vz < P.res;
    const float fr = float(P.res);
    const f3 uvw = mk3((float(vx) + 0.5f) / fr, (float(vy) + 0.5f) / fr, (float(vz) + 0.5f) / fr);
    const f3 lmin = v3(P.lmin), lmax = v3(P.lmax);
    const f3 lp = lmax - lmin;
    const f3 p = mk3(lmin.x + uvw.x * lp.x, lmin.y + uvw.y * lp.y, lmin.z + uvw.z * lp.z);
    float min_dsq = 1e10f, best_abs = -1.0f, best_sd = 10.0f, best_alpha = 0.0f;
    for (uint32_t t0 = 0; t0 < P.tri_count; t0 += 64u) {
        __syncthreads();
        if (t0 + lane < P.tri_count) s_tri[lane] = tris[P.tri_start + t0 + lane];
        __syncthreads();
        const uint32_t nt = min(64u, P.tri_count - t0);
        for (uint32_t k = 0; k < nt; ++k) {
            const BakeTri &T = s_tri[k];
            const f3 a = v3(T.p0), b = v3(T.p1), c = v3(T.p2);
            const f3 ab = b - a, ac = c - a, ap = p - a;
            const float d1 = dot(ab, ap), d2 = dot(ac, ap);
            f3 cp = a, bary = mk3(1.0f, 0.0f, 0.0f);
            if (!(d1 <= 0.0f && d2 <= 0.0f)) {
                const f3 bp = p - b;
                const float d3 = dot(ab, bp), d4 = dot(ac, bp);
                if (d3 >= 0.0f && d4 <= d3) {
                    cp = b; bary = mk3(0.0f, 1.0f, 0.0f);
                } else {
                    const float vc = d1 * d4 - d3 * d2;
                    if (vc <= 0.0f && d1 >= 0.0f && d3 <= 0.0f) {
                        const float v = d1 / (d1 - d3);
                        cp = a + ab * v; bary = mk3(1.0f - v, v, 0.0f);
                    } else {
                        const f3 cq = p - c;
                        const float d5 = dot(ab, cq), d6 = dot(ac, cq);
                        if (d6 >= 0.0f && d5 <= d6) {
                            cp = c; bary = mk3(0.0f, 0.0f, 1.0f);
                        } else {
                            const float vb = d5 * d2 - d1 * d6;
                            if (vb <= 0.0f && d2 >= 0.0f && d6 <= 0.0f) {
                                const float w = d2 / (d2 - d6);
                                cp = a + ac * w; bary = mk3(1.0f - w, 0.0f, w);
                            } else {
                                const float va = d3 * d6 - d5 * d4;
                                if (va <= 0.0f && (d4 - d3) >= 0.0f && (d5 - d6) >= 0.0f) {
                                    const float w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
                                    cp = b + (c - b) * w; bary = mk3(0.0f, 1.0f - w, w);
                                } else {
                                    const float denom = 1.0f / (va + vb + vc);
                                    const float v = vb * denom, w = vc * denom;
                                    cp = a + ab * v + ac * w; bary = mk3(1.0f - v - w, v, w);
                                }
                            }
                        }
                    }
                }
            }
            const f3 diff = p - cp;
            const float dsq = dot(diff, diff);
            const f3 tn = v3(T.n);
            const float proj = dot(diff, tn), aproj = fabsf(proj);
            const float eps_sq = 1e-7f * fmaxf(1.0f, min_dsq);
            const bool closer = dsq + eps_sq < min_dsq;
            const bool tie_better = fabsf(dsq - min_dsq) <= eps_sq && aproj > best_abs;
            if (!(closer || tie_better)) continue;
            if (dsq < min_dsq) min_dsq = dsq;
            best_abs = aproj;
            const float d = sqrtf(dsq);
            float sign = proj >= 0.0f ? 1.0f : -1.0f;
            const uint32_t bmask = uint32_t(T.bmask + 0.5f);
            if (bmask != 0u && sign < 0.0f) {  // outside an open rim the face sign would say inside
                float out_proj = -1.0f;
                if ((bmask & 1u) && bary.z <= 1e-5f) out_proj = fmaxf(out_proj, dot(diff, normalize3(cross(b - a, tn))));
                if ((bmask & 2u) && bary.x <= 1e-5f) out_proj = fmaxf(out_proj, dot(diff, normalize3(cross(c - b, tn))));
                if ((bmask & 4u) && bary.y <= 1e-5f) out_proj = fmaxf(out_proj, dot(diff, normalize3(cross(a - c, tn))));
                if (out_proj > 0.0f && out_proj >= aproj * 0.25f) sign = 1.0f;
            }
            best_sd = d * sign;
            best_alpha = bary.x * T.w0 + bary.y * T.w1 + bary.z * T.w2;
        }
    }
    if (!live) return;
    const uint32_t gx = P.col * P.res + vx, gy = P.row * P.res + vy, gz = P.layer * P.res + vz;
    const uint32_t pitch = P.pitch > 0u ? P.pitch : P.tw;
    const size_t flat = (size_t(gz) * P.th + gy) * pitch + gx;
    const _Float16 hd = (_Float16)best_sd, ha = (_Float16)best_alpha;
    uint16_t bd, ba;
    __builtin_memcpy(&bd, &hd, 2);
    __builtin_memcpy(&ba, &ha, 2);
    out[flat] = uint32_t(bd) | (uint32_t(ba) << 16);
}

hipError_t launch_bone_bake(hipStream_t st, const Params *params, uint32_t nparams, uint32_t res, const BakeTri *tris,
                            uint32_t *out) {
    if (nparams == 0 || res == 0) return hipSuccess;
    const uint32_t bpa = (res + 3u) / 4u;
    hipLaunchKernelGGL(k_bone_bake, dim3(nparams * bpa * bpa * bpa), dim3(64), 0, st, params, tris, out, bpa);
    return hipGetLastError();
}

// compute_3d_atlas_layout (sdf_baker.rs:137-163): (cols, rows, layers, res, width, height, depth)
void atlas_layout(size_t n, size_t res, size_t max_dim, size_t o[7]) {
    if (n == 0) { o[0] = o[1] = o[2] = 0; o[3] = res; o[4] = o[5] = o[6] = 0; return; }
    const size_t cols = std::max<size_t>(size_t(std::ceil(std::pow(double(n), 1.0 / 3.0))), 1);
    const size_t rows = std::max<size_t>(size_t(std::ceil(std::sqrt(double((n + cols - 1) / cols)))), 1);
    const size_t layers = std::max<size_t>((n + cols * rows - 1) / (cols * rows), 1);
    const size_t mt = std::max(cols, std::max(rows, layers));
    const size_t r = mt * res > max_dim ? std::max<size_t>(max_dim / mt, 16) : res;
    o[0] = cols; o[1] = rows; o[2] = layers; o[3] = r; o[4] = cols * r; o[5] = rows * r; o[6] = layers * r;
}

// transform_point (sdf_baker.rs:167-178): row-major 4x4, divided by w unless w is 0 or 1
inline void xform(const float *m, const float *p, float *o) {
    const float x = m[0] * p[0] + m[1] * p[1] + m[2] * p[2] + m[3];
    const float y = m[4] * p[0] + m[5] * p[1] + m[6] * p[2] + m[7];
    const float z = m[8] * p[0] + m[9] * p[1] + m[10] * p[2] + m[11];
    const float w = m[12] * p[0] + m[13] * p[1] + m[14] * p[2] + m[15];
    if (std::fabs(w) > 1e-7f && std::fabs(w - 1.0f) > 1e-5f) {
        const float iw = 1.0f / w;
        o[0] = x * iw; o[1] = y * iw; o[2] = z * iw;
    } else {
        o[0] = x; o[1] = y; o[2] = z;
    }
}

}  // namespace bone
}  // namespace tc

struct tc_bone_sdf_result {
    uint32_t w = 0, h = 0, d = 0;
    std::vector<uint32_t> texels;
    std::vector<float> infos;        // 20 per active bone
    std::vector<uint32_t> active;    // input index of each active bone
    std::vector<float> binds;        // 16 per active bone
    float kernel_ms = 0.0f;
};

extern "C" {

int tc_bake_bone_sdf(const float *verts, uint32_t nv, const int32_t *tris, uint32_t nt, const float *weights,
                     const float *binds, uint32_t nbones, uint32_t resolution, float margin, float weight_threshold,
                     float blend_k, float friction, float thickness, float restitution, int close_open_boundaries,
                     tc_bone_sdf_result **out) {
    using namespace tc::bone;
    if (!out) return fail(TC_ERR_VALUE, "out pointer is null");
    *out = nullptr;
    if (resolution == 0) return fail(TC_ERR_VALUE, "resolution must be 1 or more");
    if ((nv && !verts) || (nt && !tris) || (nbones && (!weights || !binds)))
        return fail(TC_ERR_VALUE, "null input array");
    std::vector<float> V(verts, verts + 3 * size_t(nv));
    std::vector<int32_t> T(tris, tris + 3 * size_t(nt));
    std::vector<std::vector<float>> W(nbones);
    for (uint32_t b = 0; b < nbones; ++b) W[b].assign(weights + size_t(b) * nv, weights + size_t(b + 1) * nv);
    if (close_open_boundaries) tc::sdf::close_open_boundary_loops(V, T, &W);
    const size_t n = V.size() / 3, m = T.size() / 3;
    // open boundary edges (sdf_baker.rs:203-240)
    std::map<std::pair<int32_t, int32_t>, uint32_t> ecount;
    for (size_t t = 0; t < m; ++t) {
        const int32_t i0 = T[3 * t], i1 = T[3 * t + 1], i2 = T[3 * t + 2];
        if (i0 < 0 || i1 < 0 || i2 < 0 || size_t(i0) >= n || size_t(i1) >= n || size_t(i2) >= n || i0 == i1 ||
            i1 == i2 || i2 == i0)
            continue;
        const int32_t e[3][2] = {{i0, i1}, {i1, i2}, {i2, i0}};
        for (const auto &uv : e) ++ecount[{std::min(uv[0], uv[1]), std::max(uv[0], uv[1])}];
    }
    auto open_edge = [&](int32_t a, int32_t b) {
        auto it = ecount.find({std::min(a, b), std::max(a, b)});
        return it != ecount.end() && it->second == 1u;
    };
    // active bones: largest weight above the threshold
    std::vector<uint32_t> active;
    for (uint32_t b = 0; b < nbones; ++b) {
        float mx = 0.0f;
        for (float w : W[b]) mx = std::fmax(mx, w);  // f32::max ignores NaN
        if (mx > weight_threshold) active.push_back(b);
    }
    if (active.empty()) return fail(TC_ERR_RUNTIME, "no bone has a skin weight above the threshold");
    size_t L[7];
    atlas_layout(active.size(), resolution, 2048, L);
    const size_t cols = L[0], rows = L[1], layers = L[2], res = L[3];
    std::vector<Params> params;
    std::vector<BakeTri> all;
    tc_bone_sdf_result *r = new tc_bone_sdf_result();
    r->w = uint32_t(L[4]); r->h = uint32_t(L[5]); r->d = uint32_t(L[6]);
    r->active = active;
    for (size_t ob = 0; ob < active.size(); ++ob) {
        const uint32_t sb = active[ob];
        const float *mat = binds + 16 * size_t(sb);
        r->binds.insert(r->binds.end(), mat, mat + 16);
        float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
        size_t valid = 0;
        for (size_t v = 0; v < W[sb].size(); ++v)
            if (W[sb][v] > weight_threshold && v < n) {
                float pl[3];
                xform(mat, &V[3 * v], pl);
                for (int k = 0; k < 3; ++k) { mn[k] = std::fmin(mn[k], pl[k]); mx[k] = std::fmax(mx[k], pl[k]); }
                ++valid;
            }
        if (!valid) for (int k = 0; k < 3; ++k) { mn[k] = -0.1f; mx[k] = 0.1f; }
        float size[3];
        for (int k = 0; k < 3; ++k) size[k] = std::fmax(mx[k] - mn[k], 0.05f);
        for (int k = 0; k < 3; ++k) {
            const float pad = size[k] * margin + std::fmax(thickness, 0.0f);
            mn[k] -= pad; mx[k] += pad;
            size[k] = mx[k] - mn[k];
        }
        const size_t bpl = cols * rows;
        const uint32_t l = uint32_t(ob / bpl), rem = uint32_t(ob % bpl), row = uint32_t(rem / cols),
                       col = uint32_t(rem % cols);
        float info[20];
        for (int k = 0; k < 3; ++k) {
            const float scale = 1.0f / size[k], off = -mn[k] / size[k];
            const float nn = float(k == 0 ? cols : k == 1 ? rows : layers);
            const float idx = float(k == 0 ? col : k == 1 ? row : l);
            info[k] = mn[k];
            info[4 + k] = mx[k];
            info[8 + k] = scale / nn;
            info[12 + k] = (off + idx) / nn;
        }
        info[3] = float(ob); info[7] = float(res); info[11] = blend_k; info[15] = 0.0f;
        info[16] = friction; info[17] = thickness; info[18] = restitution; info[19] = weight_threshold;
        r->infos.insert(r->infos.end(), info, info + 20);
        const uint32_t tri_start = uint32_t(all.size());
        for (size_t t = 0; t < m; ++t) {
            const int32_t i0 = T[3 * t], i1 = T[3 * t + 1], i2 = T[3 * t + 2];
            if (size_t(uint32_t(i0)) >= n || size_t(uint32_t(i1)) >= n || size_t(uint32_t(i2)) >= n) continue;
            const float w0 = size_t(i0) < W[sb].size() ? W[sb][i0] : 0.0f;
            const float w1 = size_t(i1) < W[sb].size() ? W[sb][i1] : 0.0f;
            const float w2 = size_t(i2) < W[sb].size() ? W[sb][i2] : 0.0f;
            if (!(w0 > 0.0f || w1 > 0.0f || w2 > 0.0f)) continue;
            float a[3], b[3], c[3];
            xform(mat, &V[3 * size_t(i0)], a);
            xform(mat, &V[3 * size_t(i1)], b);
            xform(mat, &V[3 * size_t(i2)], c);
            bool overlap = true;
            for (int k = 0; k < 3; ++k) {
                const float tmin = std::fmin(std::fmin(a[k], b[k]), c[k]), tmax = std::fmax(std::fmax(a[k], b[k]), c[k]);
                overlap = overlap && tmin <= mx[k] && tmax >= mn[k];
            }
            if (!overlap) continue;
            const float e01[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]}, e02[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
            const float cr[3] = {e01[1] * e02[2] - e01[2] * e02[1], e01[2] * e02[0] - e01[0] * e02[2],
                                 e01[0] * e02[1] - e01[1] * e02[0]};
            const float ls = cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2];
            float nn[3] = {0.0f, 1.0f, 0.0f};
            if (ls > 1e-12f) { const float il = 1.0f / std::sqrt(ls); nn[0] = cr[0] * il; nn[1] = cr[1] * il; nn[2] = cr[2] * il; }
            uint32_t bm = 0;
            if (open_edge(i0, i1)) bm |= 1u;
            if (open_edge(i1, i2)) bm |= 2u;
            if (open_edge(i2, i0)) bm |= 4u;
            all.push_back({{a[0], a[1], a[2]}, w0, {b[0], b[1], b[2]}, w1, {c[0], c[1], c[2]}, w2, {nn[0], nn[1], nn[2]},
                           float(bm)});
        }
        params.push_back({{mn[0], mn[1], mn[2]}, tri_start, {mx[0], mx[1], mx[2]}, uint32_t(all.size()) - tri_start,
                          col, row, l, uint32_t(res), r->w, r->h, r->d, 0u});
    }
    if (all.empty()) all.push_back({{0, 0, 0}, 0, {0, 0, 0}, 0, {0, 0, 0}, 0, {0.0f, 1.0f, 0.0f}, 0.0f});
    const size_t nvox = size_t(r->w) * r->h * r->d;
    if (r->w > 2048u || r->h > 2048u || r->d > 2048u) {  // the 3D texture limit the atlas is sized for
        const std::string msg = "bone SDF atlas " + std::to_string(r->w) + "x" + std::to_string(r->h) + "x" +
                                std::to_string(r->d) + " exceeds 2048 texels per axis";
        delete r;
        return fail(TC_ERR_RUNTIME, msg);
    }
    r->texels.assign(nvox, 0u);
    auto bail = [&](hipError_t e) { delete r; return fail(TC_ERR_RUNTIME, std::string("bone SDF bake: ") + hipGetErrorString(e)); };
    Params *dp = nullptr;
    BakeTri *dt = nullptr;
    uint32_t *dout = nullptr;
    hipError_t e = hipMalloc(reinterpret_cast<void **>(&dp), sizeof(Params) * params.size());
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void **>(&dt), sizeof(BakeTri) * all.size());
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void **>(&dout), 4 * nvox);
    if (e == hipSuccess) e = hipMemcpy(dp, params.data(), sizeof(Params) * params.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dt, all.data(), sizeof(BakeTri) * all.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(dout, 0, 4 * nvox);
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    if (e == hipSuccess) e = hipEventCreate(&ev0);
    if (e == hipSuccess) e = hipEventCreate(&ev1);
    if (e == hipSuccess) {
        e = hipEventRecord(ev0, nullptr);
        if (e == hipSuccess) e = launch_bone_bake(nullptr, dp, uint32_t(params.size()), uint32_t(res), dt, dout);
        if (e == hipSuccess) e = hipEventRecord(ev1, nullptr);
        if (e == hipSuccess) e = hipEventSynchronize(ev1);
        if (e == hipSuccess) (void)hipEventElapsedTime(&r->kernel_ms, ev0, ev1);
    }
    if (e == hipSuccess) e = hipMemcpy(r->texels.data(), dout, 4 * nvox, hipMemcpyDeviceToHost);
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (dp) (void)hipFree(dp);
    if (dt) (void)hipFree(dt);
    if (dout) (void)hipFree(dout);
    if (e != hipSuccess) return bail(e);
    *out = r;
    return TC_OK;
}

int tc_bone_sdf_result_info(const tc_bone_sdf_result *r, uint32_t *dims, uint32_t *n_active, float *kernel_ms) {
    if (!r) return fail(TC_ERR_VALUE, "null result");
    if (dims) { dims[0] = r->w; dims[1] = r->h; dims[2] = r->d; }
    if (n_active) *n_active = uint32_t(r->active.size());
    if (kernel_ms) *kernel_ms = r->kernel_ms;
    return TC_OK;
}

int tc_bone_sdf_result_arrays(const tc_bone_sdf_result *r, uint32_t *texels, float *infos, uint32_t *active,
                              float *binds) {
    if (!r) return fail(TC_ERR_VALUE, "null result");
    if (texels && !r->texels.empty()) std::memcpy(texels, r->texels.data(), 4 * r->texels.size());
    if (infos && !r->infos.empty()) std::memcpy(infos, r->infos.data(), 4 * r->infos.size());
    if (active && !r->active.empty()) std::memcpy(active, r->active.data(), 4 * r->active.size());
    if (binds && !r->binds.empty()) std::memcpy(binds, r->binds.data(), 4 * r->binds.size());
    return TC_OK;
}

void tc_bone_sdf_result_free(tc_bone_sdf_result *r) { delete r; }

}  // extern "C"
