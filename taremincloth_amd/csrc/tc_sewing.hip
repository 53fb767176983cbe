// Sewing springs: the per-substep shrink of their rest lengths and their XPBD projection, one
// launch per colour.
//
// Reference semantics (paths relative to the reference repository root):
//   shrink        crates/cloth_core/src/shaders/sewing_shrink.wgsl:40-49: current rest length moves
//                 toward the target (0) at shrink_speed, once per substep (dispatch.rs:121-134)
//   projection    crates/cloth_core/src/shaders/sewing.wgsl:56-110: XPBD with the simulator's sewing
//                 compliance (0 for stiffness >= 5000, else 1 / max(stiffness, 1), dispatch.rs:29-33);
//                 a pair closer than max(thickness sum, lock distance) with the lock on projects at
//                 zero compliance to the target length and keeps that as its current rest length
//   placement     in every iteration after the distance colours (dispatch.rs:183-191) and once more
//                 before the velocity update (dispatch.rs:318-339)
//   records       GpuSewingConstraint (mesh.rs:51-60), built and coloured in tc_mesh.cpp
//
// A record is 16 bytes {v0, v1, current rest bits, target rest bits}; the lambdas are zeroed by the
// shrink launch at the start of each substep (the reference clears them there, dispatch.rs:117-119).
#include "tc_kernels.h"
#include "tc_sim_impl.h"

#include <cstring>

namespace tc {

__global__ void __launch_bounds__(256) k_sew_shrink(uint4 *__restrict__ rec, float *__restrict__ lam, uint32_t n,
                                                    SewArgs a) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    uint4 r = rec[k];
    const float dt = fmaxf(a.dt, 1e-8f);
    const float cur = fmaxf(__uint_as_float(r.w), __uint_as_float(r.z) - a.speed * dt);
    r.z = __float_as_uint(cur);
    rec[k] = r;
    lam[k] = 0.0f;
}

__global__ void __launch_bounds__(256) k_sewing(float4 *__restrict__ P, const float4 *__restrict__ X,
                                                uint4 *__restrict__ rec, float *__restrict__ lam, uint32_t off,
                                                uint32_t cnt, SewArgs a) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= cnt) return;
    const uint32_t k = off + t;
    const uint4 r = rec[k];
    const float4 q0 = P[r.x], q1 = P[r.y];
    const float w0 = q0.w, w1 = q1.w, wsum = w0 + w1;
    if (wsum <= 1e-7f) return;
    const f3 delta = xyz(q0) - xyz(q1);
    const float dist = length(delta);
    if (dist <= 1e-7f) return;
    const float dt = fmaxf(a.dt, 1e-8f);
    const f3 dir = delta / dist;
    float target = __uint_as_float(r.z), comp = a.compliance;
    const float lock_thresh = fmaxf(X[r.x].w + X[r.y].w, a.lock_dist);  // X.w: vertex thickness
    if (a.lock > 0.5f && dist <= lock_thresh) {
        comp = 0.0f;
        target = __uint_as_float(r.w);
        rec[k].z = r.w;
    }
    const float alpha = comp / (dt * dt);
    const float c = dist - target;
    const float l = lam[k];
    const float dl = (-c - alpha * l) / (wsum + alpha);
    lam[k] = l + dl;
    const f3 c0 = dir * (w0 * dl), c1 = dir * (w1 * dl);
    P[r.x] = make_float4(q0.x + c0.x, q0.y + c0.y, q0.z + c0.z, w0);
    P[r.y] = make_float4(q1.x - c1.x, q1.y - c1.y, q1.z - c1.z, w1);
}

hipError_t launch_sew_shrink(hipStream_t s, uint4 *rec, float *lam, uint32_t n, const SewArgs &a) {
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(k_sew_shrink, dim3((n + 255) / 256), dim3(256), 0, s, rec, lam, n, a);
    return hipGetLastError();
}

hipError_t launch_sewing(hipStream_t s, float4 *P, const float4 *X, uint4 *rec, float *lam, uint32_t off, uint32_t cnt,
                         const SewArgs &a, uint32_t block) {
    if (!cnt) return hipSuccess;
    hipLaunchKernelGGL(k_sewing, dim3((cnt + block - 1) / block), dim3(block), 0, s, P, X, rec, lam, off, cnt, a);
    return hipGetLastError();
}

}  // namespace tc

namespace {
tc::SewArgs sew_args(const tc_sim *s, float dt_sub) {
    tc::SewArgs a;
    a.dt = dt_sub;
    a.compliance = s->sew_stiffness >= 5000.0f ? 0.0f : 1.0f / std::max(s->sew_stiffness, 1.0f);
    a.lock = s->sew_lock ? 1.0f : 0.0f;
    a.lock_dist = s->sew_lock_dist;
    a.speed = s->sew_speed;
    return a;
}
}  // namespace

hipError_t enqueue_sew_shrink(tc_sim *s, float dt_sub) {
    return tc::launch_sew_shrink(s->stream, s->srec.p, s->slam.p, uint32_t(s->mesh.s_v0.size()), sew_args(s, dt_sub));
}

hipError_t enqueue_sewing(tc_sim *s, float4 *P, float dt_sub, float *lam) {
    const tc::Mesh &m = s->mesh;
    const tc::SewArgs a = sew_args(s, dt_sub);
    for (size_t c = 0; c < m.s_col.count.size(); ++c) {
        const hipError_t e = tc::launch_sewing(s->stream, P, s->X.p, s->srec.p, lam, m.s_col.offset[c],
                                               m.s_col.count[c], a, s->block);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// Device records of the sewing constraints and, for meshes with seams, the bending gate table.
int upload_sewing(tc_sim *s) {
    const tc::Mesh &m = s->mesh;
    const size_t n = m.s_v0.size();
    if (n) {
        std::vector<uint4> rec(n);
        for (size_t k = 0; k < n; ++k) {
            rec[k].x = m.s_v0[k];
            rec[k].y = m.s_v1[k];
            std::memcpy(&rec[k].z, &m.s_rest[k], 4);
            std::memcpy(&rec[k].w, &m.s_target[k], 4);
        }
        TC_HIP(s->srec.upload(rec.data(), n, s->stream));
        TC_HIP(s->slam.alloc(n));
        TC_HIP(s->slam2.alloc(n));
        TC_HIP(hipMemsetAsync(s->slam.p, 0, 4 * n, s->stream));
        TC_HIP(hipMemsetAsync(s->slam2.p, 0, 4 * n, s->stream));
    }
    if (m.has_seams) {
        const size_t nb = m.b_v0.size();
        std::vector<uint4> g(nb);
        for (size_t k = 0; k < nb; ++k) {
            const bool seam = m.b_pad0[k] > 0.5f && m.b_pad0[k] < 1.5f;
            g[k].x = seam ? m.b_v0[k] : 0u;
            g[k].y = seam ? m.b_v1[k] : 0u;
            std::memcpy(&g[k].z, &m.b_pad1[k], 4);
            g[k].w = seam ? 1u : 0u;
        }
        TC_HIP(s->bgate.upload(g.data(), nb, s->stream));
    }
    s->sew_speed = m.sew_shrink_speed;
    return TC_OK;
}

extern "C" {

int tc_sim_set_sewing_stiffness(tc_sim *s, float stiffness) {  // mod.rs:1211-1213
    if (!s) return fail(TC_ERR_VALUE, "null simulator");
    s->sew_stiffness = std::max(stiffness, 1.0f);
    s->touch();
    return TC_OK;
}

int tc_sim_set_sewing_shrink_speed(tc_sim *s, float speed) {  // mod.rs:1216-1239: the current rest lengths stay
    if (!s) return fail(TC_ERR_VALUE, "null simulator");
    s->sew_speed = std::isfinite(speed) ? std::max(speed, 0.0f) : 1.0f;
    s->touch();
    return TC_OK;
}

int tc_sim_set_enable_sewing_lock(tc_sim *s, int enable) {  // mod.rs:1248-1250
    if (!s) return fail(TC_ERR_VALUE, "null simulator");
    s->sew_lock = enable != 0;
    s->touch();
    return TC_OK;
}

int tc_sim_set_sewing_lock_distance(tc_sim *s, float distance) {  // mod.rs:1253-1255
    if (!s) return fail(TC_ERR_VALUE, "null simulator");
    s->sew_lock_dist = std::max(distance, 0.0f);
    s->touch();
    return TC_OK;
}

int tc_sim_get_sewing_state(const tc_sim *s, float *out) {
    if (!s || !out) return fail(TC_ERR_VALUE, "null argument");
    out[0] = s->sew_stiffness;
    out[1] = s->sew_speed;
    out[2] = s->sew_lock ? 1.0f : 0.0f;
    out[3] = s->sew_lock_dist;
    out[4] = float(s->mesh.s_v0.size());
    return TC_OK;
}

int tc_sim_get_sewing_rest_lengths(tc_sim *s, float *out, uint32_t n) {  // dispatch.rs:912-926
    if (!s) return fail(TC_ERR_VALUE, "null simulator");
    const size_t ns = s->mesh.s_v0.size();
    if (n != ns) return fail(TC_ERR_VALUE, "expected " + std::to_string(ns) + " rest lengths");
    if (!ns) return TC_OK;
    if (!out) return fail(TC_ERR_VALUE, "null output");
    std::vector<uint4> rec(ns);
    TC_HIP(hipStreamSynchronize(s->stream));
    TC_HIP(hipMemcpy(rec.data(), s->srec.p, 16 * ns, hipMemcpyDeviceToHost));
    for (size_t k = 0; k < ns; ++k) std::memcpy(&out[k], &rec[k].z, 4);
    return TC_OK;
}

int tc_sim_set_sewing_rest_lengths(tc_sim *s, const float *values, uint32_t n) {  // dispatch.rs:928-950
    if (!s) return fail(TC_ERR_VALUE, "null simulator");
    const size_t ns = s->mesh.s_v0.size();
    if (n != ns || ns == 0) return TC_OK;  // the reference ignores a list of the wrong length
    if (!values) return fail(TC_ERR_VALUE, "null values");
    std::vector<uint4> rec(ns);
    TC_HIP(hipStreamSynchronize(s->stream));
    TC_HIP(hipMemcpy(rec.data(), s->srec.p, 16 * ns, hipMemcpyDeviceToHost));
    for (size_t k = 0; k < ns; ++k) std::memcpy(&rec[k].z, &values[k], 4);
    TC_HIP(hipMemcpy(s->srec.p, rec.data(), 16 * ns, hipMemcpyHostToDevice));
    return TC_OK;
}

// mod.rs:1381-1420: the share of sewing pairs closer than merge_dist; pairs with an index outside
// the positions are skipped, and with none counted the share is 1
int tc_sewing_closure_ratio(const float *positions, uint32_t num_vertices, const uint32_t *pairs, uint32_t num_pairs,
                            float merge_dist, float *ratio) {
    if (!ratio || (num_vertices && !positions) || (num_pairs && !pairs)) return fail(TC_ERR_VALUE, "null argument");
    const float m2 = merge_dist * merge_dist;
    uint32_t closed = 0, total = 0;
    for (uint32_t k = 0; k < num_pairs; ++k) {
        const uint32_t a = pairs[2 * k], b = pairs[2 * k + 1];
        if (a >= num_vertices || b >= num_vertices) continue;
        ++total;
        const float dx = positions[3 * a] - positions[3 * b];
        const float dy = positions[3 * a + 1] - positions[3 * b + 1];
        const float dz = positions[3 * a + 2] - positions[3 * b + 2];
        if (dx * dx + dy * dy + dz * dz <= m2) ++closed;
    }
    *ratio = total ? float(closed) / float(total) : 1.0f;
    return TC_OK;
}

int tc_sim_set_sewing_priority_options(tc_sim *s, int enabled, float threshold, float merge_dist,
                                       uint32_t ramp_frames, uint32_t max_frames) {  // mod.rs:1262-1292
    if (!s) return fail(TC_ERR_VALUE, "null simulator");
    const bool on = enabled != 0;
    const float th = std::min(std::max(threshold, 0.0f), 1.0f);
    const float md = std::min(std::max(merge_dist, 0.0005f), 0.1f);
    const uint32_t rf = std::min(ramp_frames, 120u), mf = std::min(max_frames, 100000u);
    if (s->prio_on == on && s->prio_threshold == th && s->prio_merge == md && s->prio_ramp == rf && s->prio_max == mf)
        return TC_OK;  // an unchanged setting keeps the latch
    s->prio_on = on;
    s->prio_threshold = th;
    s->prio_merge = md;
    s->prio_ramp = rf;
    s->prio_max = mf;
    s->prio_latched = false;
    s->prio_frame = 0;
    s->prio_ramp_t = 0.0f;
    s->prio_ratio = 0.0f;
    s->touch();
    return TC_OK;
}

int tc_sim_set_sewing_priority_state(tc_sim *s, int latched, uint32_t frame, float ramp_t,
                                     float ratio) {  // mod.rs:1307-1318
    if (!s) return fail(TC_ERR_VALUE, "null simulator");
    s->prio_latched = latched != 0;
    s->prio_frame = frame;
    s->prio_ramp_t = std::min(std::max(ramp_t, 0.0f), 1.0f);
    s->prio_ratio = std::min(std::max(ratio, 0.0f), 1.0f);
    s->touch();
    return TC_OK;
}

int tc_sim_get_sewing_priority_state(const tc_sim *s, double *out) {
    if (!s || !out) return fail(TC_ERR_VALUE, "null argument");
    const bool live = s->prio_on && !s->mesh.s_v0.empty();
    out[0] = s->prio_on ? 1.0 : 0.0;
    out[1] = s->prio_threshold;
    out[2] = s->prio_merge;
    out[3] = s->prio_ramp;
    out[4] = s->prio_max;
    out[5] = s->prio_latched ? 1.0 : 0.0;
    out[6] = s->prio_frame;
    out[7] = s->prio_ramp_t;
    out[8] = s->prio_ratio;
    out[9] = s->priority_scale();
    out[10] = live ? s->prio_ratio : 1.0;                  // mod.rs:1379-1384
    out[11] = live && !s->prio_latched ? 1.0 : 0.0;       // mod.rs:1387-1391
    return TC_OK;
}

int tc_sim_update_sewing_priority(tc_sim *s, const float *positions, uint32_t num_vertices,
                                  float *out) {  // mod.rs:1323-1361
    if (!s || !out) return fail(TC_ERR_VALUE, "null argument");
    if (!s->prio_on || s->mesh.s_v0.empty()) {
        out[0] = 1.0f; out[1] = 1.0f; out[2] = 1.0f;
        return TC_OK;
    }
    const tc::Mesh &m = s->mesh;
    std::vector<uint32_t> pairs(2 * m.s_v0.size());
    for (size_t k = 0; k < m.s_v0.size(); ++k) { pairs[2 * k] = m.s_v0[k]; pairs[2 * k + 1] = m.s_v1[k]; }
    float ratio = 1.0f;
    if (int rc = tc_sewing_closure_ratio(positions, num_vertices, pairs.data(), uint32_t(m.s_v0.size()),
                                         s->prio_merge, &ratio))
        return rc;
    const float before = s->priority_scale();
    s->prio_ratio = ratio;
    if (s->prio_frame != UINT32_MAX) ++s->prio_frame;
    if (!s->prio_latched && (ratio >= s->prio_threshold || (s->prio_max > 0 && s->prio_frame >= s->prio_max)))
        s->prio_latched = true;
    if (s->prio_latched && s->prio_ramp > 0)
        s->prio_ramp_t = std::min(s->prio_ramp_t + 1.0f / float(s->prio_ramp), 1.0f);
    else
        s->prio_ramp_t = s->prio_latched ? 1.0f : 0.0f;
    const float after = s->priority_scale();
    if (after != before) s->touch();
    out[0] = ratio;
    out[1] = after;
    out[2] = s->prio_latched ? 1.0f : 0.0f;
    return TC_OK;
}

}  // extern "C"
