// Dynamic bone SDF: the bone-SDF atlas of a skinned collider mesh re-baked on the device at the start
// of a frame, straight into the simulator's SDF texture, so that the collider follows the pose.
//
// Reference semantics (paths relative to the reference repository root):
//   setup        crates/cloth_py/src/lib.rs:2362-2588 (setup_dynamic_bone_sdf: open-edge masks of the
//                triangle sources, per-bone triangle ranges, atlas tiles of the bake parameters),
//                crates/cloth_core/src/simulation/mod.rs:1998-2312 (texture, infos, identity transforms)
//   per frame    crates/cloth_core/src/simulation/mod.rs:2320-2457 (dispatch_dynamic_bone_sdf_update:
//                update interval, skin matrices S = W B, dirty bones), called by step, step_async and
//                step_buffered before the frame (simulation/dispatch.rs:624, 760, 1312)
//   kernels      crates/cloth_core/src/shaders/skinning.wgsl:17-43 (linear blend skinning),
//                prep_bone_triangles.wgsl:31-67 (triangles into their bones' frames), bake_sdf.wgsl
//                (the bake of tc_bone_sdf.hip)
//
// Three launches per update: one lane per vertex (skinning), one lane per triangle (bone frame and
// normal), then the atlas bake of the bones to refresh (one wavefront per 4x4x4 brick).  The per-frame
// matrices and the bake parameters reach the device through one pinned staging block; the host waits
// for the previous update's copy before it rewrites the block.
#include "tc_sim_impl.h"
#include "tc_sdf.h"

#include <array>
#include <cmath>
#include <cstring>
#include <map>
#include <set>
#include <string>

namespace tc {
namespace dyn {

// m holds a 4x4 matrix as the reference uploads it for WGSL: column c at m[4c .. 4c+3], so
// (m * (v, 1)).r = m[r] v.x + m[4+r] v.y + m[8+r] v.z + m[12+r], summed in that order.
__device__ __forceinline__ float3 apply(const float *__restrict__ m, float x, float y, float z) {
    return make_float3(m[0] * x + m[4] * y + m[8] * z + m[12], m[1] * x + m[5] * y + m[9] * z + m[13],
                       m[2] * x + m[6] * y + m[10] * z + m[14]);
}

// skinning.wgsl:17-43: p = sum over the four influences with w > 0 of w * (S_b * (rest, 1)); a vertex
// whose weights sum to 1e-6 or less keeps its rest position.
__global__ void __launch_bounds__(256) k_skin(const float4 *__restrict__ rest, const float *__restrict__ skin,
                                              float4 *__restrict__ out, uint32_t nv) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nv) return;
    const float4 pos = rest[4 * size_t(i)];
    const float4 idx = rest[4 * size_t(i) + 2];
    const float4 wt = rest[4 * size_t(i) + 3];
    const float ws[4] = {wt.x, wt.y, wt.z, wt.w};
    const uint32_t bs[4] = {__float_as_uint(idx.x), __float_as_uint(idx.y), __float_as_uint(idx.z),
                            __float_as_uint(idx.w)};
    float px = 0.0f, py = 0.0f, pz = 0.0f, wsum = 0.0f;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const float w = ws[k];
        if (w > 0.0f) {
            const float3 t = apply(skin + 16 * size_t(bs[k]), pos.x, pos.y, pos.z);
            px = px + w * t.x;
            py = py + w * t.y;
            pz = pz + w * t.z;
            wsum = wsum + w;
        }
    }
    if (wsum <= 1e-6f) { px = pos.x; py = pos.y; pz = pos.z; }
    out[i] = make_float4(px, py, pz, 1.0f);
}

// prep_bone_triangles.wgsl:31-67: the skinned corners through the bone's inverse world matrix, the
// unit face normal (0, 1, 0 for a degenerate triangle), the corner weights and the open-edge mask.
__global__ void __launch_bounds__(256) k_prep(const float4 *__restrict__ skinned, const float *__restrict__ inv,
                                              const uint4 *__restrict__ src, tc::bone::BakeTri *__restrict__ out,
                                              uint32_t nt) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nt) return;
    const uint4 s0 = src[2 * size_t(t)], s1 = src[2 * size_t(t) + 1];
    const float *m = inv + 16 * size_t(s0.w);
    const float4 a = skinned[s0.x], b = skinned[s0.y], c = skinned[s0.z];
    const float3 p0 = apply(m, a.x, a.y, a.z), p1 = apply(m, b.x, b.y, b.z), p2 = apply(m, c.x, c.y, c.z);
    const float e1x = p1.x - p0.x, e1y = p1.y - p0.y, e1z = p1.z - p0.z;
    const float e2x = p2.x - p0.x, e2y = p2.y - p0.y, e2z = p2.z - p0.z;
    const float cx = e1y * e2z - e1z * e2y, cy = e1z * e2x - e1x * e2z, cz = e1x * e2y - e1y * e2x;
    const float ls = cx * cx + cy * cy + cz * cz;
    float nx = 0.0f, ny = 1.0f, nz = 0.0f;
    if (ls > 1e-12f) {
        const float il = 1.0f / sqrtf(ls);
        nx = cx * il; ny = cy * il; nz = cz * il;
    }
    tc::bone::BakeTri o;
    o.p0[0] = p0.x; o.p0[1] = p0.y; o.p0[2] = p0.z; o.w0 = __uint_as_float(s1.x);
    o.p1[0] = p1.x; o.p1[1] = p1.y; o.p1[2] = p1.z; o.w1 = __uint_as_float(s1.y);
    o.p2[0] = p2.x; o.p2[1] = p2.y; o.p2[2] = p2.z; o.w2 = __uint_as_float(s1.z);
    o.n[0] = nx; o.n[1] = ny; o.n[2] = nz; o.bmask = __uint_as_float(s1.w);
    out[t] = o;
}

// Host 4x4 helpers in the reference's f32 arithmetic (simulation/mod.rs:3192-3214).
void mat4_mul(const float *a, const float *b, float *o) {  // row-major o = a b
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j)
            o[4 * i + j] = a[4 * i] * b[j] + a[4 * i + 1] * b[4 + j] + a[4 * i + 2] * b[8 + j] + a[4 * i + 3] * b[12 + j];
}
void mat4_transpose(const float *m, float *o) {
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) o[4 * i + j] = m[4 * j + i];
}

}  // namespace dyn
}  // namespace tc

int dyn_sdf_frame(tc_sim *s) {
    using namespace tc::dyn;
    tc_sim::DynSdf &D = s->dyn;
    if (!D.enabled) return TC_OK;
    D.counter += 1u;  // wrapping, as the reference's frame counter
    if ((D.counter - 1u) % D.interval != 0u) return TC_OK;
    const uint32_t nb = D.nbones;
    if (s->sdf_xf_h.size() < size_t(nb) * 32) return TC_OK;  // fewer transforms than bones
    // the staging block of the previous update may still be in flight
    TC_HIP(hipEventSynchronize(D.mats_ev));
    float *skin = D.mats_host, *inv = D.mats_host + 16 * size_t(nb);
    for (uint32_t b = 0; b < nb; ++b) {
        float w_row[16], s_row[16];
        mat4_transpose(&s->sdf_xf_h[32 * size_t(b)], w_row);
        mat4_mul(w_row, &D.bind_inv[16 * size_t(b)], s_row);
        mat4_transpose(s_row, skin + 16 * size_t(b));
        std::memcpy(inv + 16 * size_t(b), &s->sdf_xf_h[32 * size_t(b) + 16], 64);
    }
    TC_HIP(hipMemcpyAsync(D.mats.p, D.mats_host, 4 * 32 * size_t(nb), hipMemcpyHostToDevice, s->stream));
    hipLaunchKernelGGL(k_skin, dim3((D.nv + 255) / 256), dim3(256), 0, s->stream, D.rest.p, D.mats.p, D.skinned.p,
                       D.nv);
    TC_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_prep, dim3((D.nt + 255) / 256), dim3(256), 0, s->stream, D.skinned.p,
                       D.mats.p + 16 * size_t(nb), D.src.p, D.tris.p, D.nt);
    TC_HIP(hipGetLastError());
    // which bones to bake: every bone until the first bake, then the dirty list when one is set
    // (an empty list skips the bake, after the skinning as in the reference), else every bone
    uint32_t count = nb;
    bool dirty_mode = false;
    if (D.baked && D.dirty_set) {
        if (D.dirty.empty()) {
            ++D.skips;
            TC_HIP(hipEventRecord(D.mats_ev, s->stream));
            return TC_OK;
        }
        count = uint32_t(D.dirty.size());
        dirty_mode = true;
    }
    if (dirty_mode) {  // the valid entries go to the front; the rest of the buffer keeps what it held
        size_t j = 0;
        for (uint32_t b : D.dirty)
            if (b < D.params.size() && j < D.params_dev.size()) D.params_dev[j++] = D.params[b];
    } else {
        D.params_dev = D.params;
    }
    std::memcpy(D.mats_host + 32 * size_t(nb), D.params_dev.data(), sizeof(tc::bone::Params) * nb);
    TC_HIP(hipMemcpyAsync(D.dparams.p, D.mats_host + 32 * size_t(nb), sizeof(tc::bone::Params) * nb,
                          hipMemcpyHostToDevice, s->stream));
    TC_HIP(hipEventRecord(D.mats_ev, s->stream));
    // the bake reads parameter entries [0, count) of a buffer of nb entries (bake_sdf.wgsl:35-48)
    const uint32_t nbake = std::min(count, nb);
    TC_HIP(tc::bone::launch_bone_bake(s->stream, D.dparams.p, nbake, D.res, D.tris.p, s->sdf_tex.p));
    D.bakes += nbake;
    D.baked = true;
    return TC_OK;
}

extern "C" {

int tc_sim_setup_dynamic_bone_sdf(tc_sim *s, uint32_t width, uint32_t height, uint32_t depth, uint32_t res,
                                  const float *infos, uint32_t num_bones, const float *rest_positions,
                                  uint32_t num_verts, const uint32_t *bone_indices, const float *bone_weights,
                                  const uint32_t *tri_sources, const float *tri_weights, uint32_t num_tris,
                                  const float *bind_inv, uint32_t num_bind, uint32_t update_interval) {
    if (!s) return fail(TC_ERR_VALUE, "null simulator");
    if (res == 0 || width == 0 || height == 0 || depth == 0)
        return fail(TC_ERR_VALUE, "width, height, depth and res must all be 1 or more");
    if ((num_bones && !infos) || (num_verts && (!rest_positions || !bone_indices || !bone_weights)) ||
        (num_tris && (!tri_sources || !tri_weights)) || (num_bind && !bind_inv))
        return fail(TC_ERR_VALUE, "null input array");
    if (num_bind < num_bones)
        return fail(TC_ERR_VALUE, "bone_bind_inv_matrices must hold one 4x4 matrix per bone (shape (N_bones, 4, 4))");
    // indices the kernels would read out of bounds are rejected here
    for (size_t i = 0; i < size_t(num_verts) * 4; ++i)
        if (bone_weights[i] > 0.0f && bone_indices[i] >= num_bones)
            return fail(TC_ERR_VALUE, "vertex " + std::to_string(i / 4) + " has a weighted bone index " +
                                          std::to_string(bone_indices[i]) + " out of range");
    for (uint32_t t = 0; t < num_tris; ++t) {
        const uint32_t *q = tri_sources + 4 * size_t(t);
        if (q[0] >= num_verts || q[1] >= num_verts || q[2] >= num_verts || q[3] >= num_bones)
            return fail(TC_ERR_VALUE, "triangle source " + std::to_string(t) + " has a vertex or bone index out of range");
    }
    // simulation/mod.rs:1998-2009: an empty setup leaves the simulator as it is
    if (num_verts == 0 || num_tris == 0 || num_bones == 0) {
        s->dyn.interval = std::max(update_interval, 1u);
        return TC_OK;
    }
    // atlas tiles (lib.rs:2532-2567): bone b at column b % cols, row (b / cols) % rows, layer b / (cols rows)
    const uint32_t n_cols = std::max(width / res, 1u), n_rows = std::max(height / res, 1u);
    const uint32_t bpl = n_cols * n_rows;
    for (uint32_t b = 0; b < num_bones; ++b) {
        const uint32_t l = b / bpl, rem = b % bpl, r = rem / n_cols, c = rem % n_cols;
        if (uint64_t(c + 1) * res > width || uint64_t(r + 1) * res > height || uint64_t(l + 1) * res > depth)
            return fail(TC_ERR_VALUE, "bone " + std::to_string(b) + "'s " + std::to_string(res) + "^3 tile lies outside the " +
                                          std::to_string(width) + "x" + std::to_string(height) + "x" +
                                          std::to_string(depth) + " atlas");
    }
    const size_t nvox = size_t(width) * height * depth;
    if (width > 2048u || height > 2048u || depth > 2048u)
        return fail(TC_ERR_VALUE, "the atlas exceeds 2048 texels per axis");

    // open-boundary masks (lib.rs:2440-2490): undirected edges used once by the distinct triangles
    std::set<std::array<uint32_t, 3>> seen;
    std::map<std::pair<int32_t, int32_t>, uint32_t> ecount;
    for (uint32_t t = 0; t < num_tris; ++t) {
        const uint32_t *q = tri_sources + 4 * size_t(t);
        std::array<uint32_t, 3> key = {q[0], q[1], q[2]};
        std::sort(key.begin(), key.end());
        if (!seen.insert(key).second) continue;
        if (q[0] > uint32_t(INT32_MAX) || q[1] > uint32_t(INT32_MAX) || q[2] > uint32_t(INT32_MAX)) continue;
        const int32_t i0 = int32_t(q[0]), i1 = int32_t(q[1]), i2 = int32_t(q[2]);
        if (i0 == i1 || i1 == i2 || i2 == i0) continue;  // indices are < num_verts (checked above)
        const int32_t e[3][2] = {{i0, i1}, {i1, i2}, {i2, i0}};
        for (const auto &uv : e) ++ecount[{std::min(uv[0], uv[1]), std::max(uv[0], uv[1])}];
    }
    auto open_edge = [&](int32_t a, int32_t b) {
        auto it = ecount.find({std::min(a, b), std::max(a, b)});
        return it != ecount.end() && it->second == 1u;
    };
    std::vector<uint4> src(2 * size_t(num_tris));
    std::vector<uint32_t> tri_start(num_bones, 0u), tri_count(num_bones, 0u);
    for (uint32_t t = 0; t < num_tris; ++t) {
        const uint32_t *q = tri_sources + 4 * size_t(t);
        const float *w = tri_weights + 3 * size_t(t);
        uint32_t bm = 0;
        if (q[0] <= uint32_t(INT32_MAX) && q[1] <= uint32_t(INT32_MAX) && q[2] <= uint32_t(INT32_MAX)) {
            const int32_t i0 = int32_t(q[0]), i1 = int32_t(q[1]), i2 = int32_t(q[2]);
            if (i0 != i1 && i1 != i2 && i2 != i0) {
                if (open_edge(i0, i1)) bm |= 1u;
                if (open_edge(i1, i2)) bm |= 2u;
                if (open_edge(i2, i0)) bm |= 4u;
            }
        }
        const float fbm = float(bm);
        uint32_t wb[3], mb;
        std::memcpy(wb, w, 12);
        std::memcpy(&mb, &fbm, 4);
        src[2 * size_t(t)] = make_uint4(q[0], q[1], q[2], q[3]);
        src[2 * size_t(t) + 1] = make_uint4(wb[0], wb[1], wb[2], mb);
        const uint32_t b = q[3];
        if (tri_count[b] == 0) tri_start[b] = t;
        ++tri_count[b];
    }
    std::vector<tc::bone::Params> params(num_bones);
    for (uint32_t b = 0; b < num_bones; ++b) {
        const float *I = infos + 20 * size_t(b);
        tc::bone::Params &P = params[b];
        const uint32_t l = b / bpl, rem = b % bpl;
        P.lmin[0] = I[0]; P.lmin[1] = I[1]; P.lmin[2] = I[2];
        P.lmax[0] = I[4]; P.lmax[1] = I[5]; P.lmax[2] = I[6];
        P.tri_start = tri_start[b];
        P.tri_count = tri_count[b];
        P.col = rem % n_cols; P.row = rem / n_cols; P.layer = l;
        P.res = res; P.tw = width; P.th = height; P.td = depth;
        P.pitch = 0;  // the texture is one linear array of width-long rows
    }
    std::vector<float4> rest(4 * size_t(num_verts));
    for (uint32_t i = 0; i < num_verts; ++i) {
        const float *p = rest_positions + 3 * size_t(i);
        const uint32_t *bi = bone_indices + 4 * size_t(i);
        const float *bw = bone_weights + 4 * size_t(i);
        rest[4 * size_t(i)] = make_float4(p[0], p[1], p[2], 0.0f);
        rest[4 * size_t(i) + 1] = make_float4(0.0f, 1.0f, 0.0f, 0.0f);  // the rest normal: carried, never read
        float4 idx;
        std::memcpy(&idx, bi, 16);  // the indices' bits, read back as u32 by the kernel
        rest[4 * size_t(i) + 2] = idx;
        rest[4 * size_t(i) + 3] = make_float4(bw[0], bw[1], bw[2], bw[3]);
    }

    TC_HIP(hipStreamSynchronize(s->stream));
    tc_sim::DynSdf &D = s->dyn;
    // the texture (zero until the first bake), the infos and the transforms (existing ones kept,
    // identities appended up to one per bone; simulation/mod.rs:2036-2087)
    TC_HIP(s->sdf_tex.alloc(nvox));
    TC_HIP(hipMemset(s->sdf_tex.p, 0, 4 * nvox));
    s->sdf_info_h.assign(infos, infos + 20 * size_t(num_bones));
    TC_HIP(s->sdf_info.alloc(s->sdf_info_h.size()));
    TC_HIP(hipMemcpy(s->sdf_info.p, s->sdf_info_h.data(), 4 * s->sdf_info_h.size(), hipMemcpyHostToDevice));
    const size_t have = s->sdf_xf_h.size() / 32;
    s->sdf_xf_h.resize(32 * size_t(num_bones), 0.0f);
    for (size_t b = have; b < num_bones; ++b)
        for (int m = 0; m < 2; ++m)
            for (int k = 0; k < 4; ++k) s->sdf_xf_h[b * 32 + m * 16 + k * 5] = 1.0f;
    TC_HIP(s->sdf_xf.alloc(s->sdf_xf_h.size()));
    TC_HIP(hipMemcpy(s->sdf_xf.p, s->sdf_xf_h.data(), 4 * s->sdf_xf_h.size(), hipMemcpyHostToDevice));
    s->sdf_w = width; s->sdf_h = height; s->sdf_d = depth;
    s->sdf_nb = num_bones;

    D.nbones = num_bones; D.nv = num_verts; D.nt = num_tris; D.res = res;
    D.bind_inv.assign(bind_inv, bind_inv + 16 * size_t(num_bones));
    D.params = params;
    D.params_dev = params;
    TC_HIP(D.rest.alloc(rest.size()));
    TC_HIP(hipMemcpy(D.rest.p, rest.data(), sizeof(float4) * rest.size(), hipMemcpyHostToDevice));
    TC_HIP(D.skinned.alloc(num_verts));
    TC_HIP(D.src.alloc(src.size()));
    TC_HIP(hipMemcpy(D.src.p, src.data(), sizeof(uint4) * src.size(), hipMemcpyHostToDevice));
    TC_HIP(D.tris.alloc(num_tris));
    TC_HIP(D.dparams.alloc(num_bones));
    TC_HIP(hipMemcpy(D.dparams.p, params.data(), sizeof(tc::bone::Params) * num_bones, hipMemcpyHostToDevice));
    TC_HIP(D.mats.alloc(32 * size_t(num_bones)));
    if (D.mats_host) { (void)hipHostFree(D.mats_host); D.mats_host = nullptr; }
    TC_HIP(hipHostMalloc(reinterpret_cast<void **>(&D.mats_host), 4 * 48 * size_t(num_bones), hipHostMallocDefault));
    if (!D.mats_ev) TC_HIP(hipEventCreateWithFlags(&D.mats_ev, hipEventDisableTiming));
    TC_HIP(hipEventRecord(D.mats_ev, s->stream));
    D.ready = true;
    D.enabled = true;
    D.baked = false;
    D.dirty_set = false;
    D.dirty.clear();
    D.interval = std::max(update_interval, 1u);
    s->touch();
    return TC_OK;
}

int tc_sim_set_dynamic_bone_sdf_dirty_bones(tc_sim *s, const uint32_t *bones, uint32_t n, int has_list) {
    if (!s) return fail(TC_ERR_VALUE, "null simulator");
    if (has_list && n && !bones) return fail(TC_ERR_VALUE, "null bone list");
    s->dyn.dirty_set = has_list != 0;
    s->dyn.dirty.assign(bones, bones + (has_list ? n : 0));
    return TC_OK;
}

int tc_sim_set_dynamic_bone_sdf_enabled(tc_sim *s, int enabled) {
    if (!s) return fail(TC_ERR_VALUE, "null simulator");
    s->dyn.enabled = enabled != 0 && s->dyn.ready;
    return TC_OK;
}

int tc_sim_set_dynamic_bone_sdf_update_interval(tc_sim *s, uint32_t interval) {
    if (!s) return fail(TC_ERR_VALUE, "null simulator");
    s->dyn.interval = std::max(interval, 1u);
    return TC_OK;
}

int tc_sim_dynamic_bone_sdf_stats(const tc_sim *s, uint32_t *out) {
    if (!s || !out) return fail(TC_ERR_VALUE, "null argument");
    out[0] = s->dyn.enabled ? 1u : 0u;
    out[1] = s->dyn.baked ? 1u : 0u;
    out[2] = s->dyn.bakes;
    out[3] = s->dyn.counter;
    out[4] = s->dyn.skips;
    return TC_OK;
}

}  // extern "C"
