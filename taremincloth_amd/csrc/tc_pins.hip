// Dynamic pins: the compliant pin projection and the host API that keeps the
// pin table and the vertices' inverse masses.
//
// Reference semantics (paths relative to the reference repository root):
//   pin projection   crates/cloth_core/src/shaders/pin.wgsl:55-92, run at the start of every solver
//                    iteration and once more after the decoupled collider
//                    (crates/cloth_core/src/simulation/dispatch.rs:176-181, 238-250)
//   host API         crates/cloth_core/src/simulation/mod.rs:2768-2905 (set_pin_target,
//                    set_pin_targets_batch, release_pin(s_batch), clear_dynamic_pins,
//                    set_vertex_inv_masses, get_vertex_inv_masses), dispatch.rs:952-986
//                    (set_positions_and_velocities keeps them), dispatch.rs:1411-1427 (reset)
//
// A pin is {vertex, target, weight}.  Its weight w (clamped to [0, 1]) scales the vertex's
// inverse mass to orig * (1 - w) and sets the pin's compliance to 1e-3 (1 - w) / w: w = 1 snaps the
// vertex onto the target every iteration, smaller weights pull it with a finite XPBD spring.  The
// pins of one table touch distinct vertices, so one thread per pin needs no atomics and the
// result does not depend on the table's order.  Inverse masses live in P.w (both ping-pong
// buffers), where predict and every projection read them.
#include "tc_sim_impl.h"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace tc {

__global__ void __launch_bounds__(256) k_pins(float4 *__restrict__ P, const uint32_t *__restrict__ pin_v,
                                              const float4 *__restrict__ pin_t, uint32_t npins, SubstepArgs a) {
    if (a.run_flag && *a.run_flag == 0u) return;
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= npins) return;
    const float4 pt = pin_t[t];
    const float w = fminf(fmaxf(pt.w, 0.0f), 1.0f);  // NaN -> 0
    if (w <= 0.0f) return;
    const uint32_t v = pin_v[t];
    const float4 q = P[v];
    const float inv_m = q.w;
    if (inv_m <= 0.0f) {  // held: snap onto the target
        P[v] = make_float4(pt.x, pt.y, pt.z, q.w);
        return;
    }
    const f3 d = mk3(q.x - pt.x, q.y - pt.y, q.z - pt.z);
    const float dist = length(d);
    if (!(dist > 1e-7f)) return;
    const f3 n = d / dist;
    const float dt = fmaxf(a.dt, 1e-8f);
    const float compliance = 1e-3f * (1.0f - w) / fmaxf(w, 1e-6f);
    const float alpha = compliance / (dt * dt);
    const float corr = -dist / (inv_m + alpha);
    const float s = inv_m * corr;
    P[v] = make_float4(q.x + s * n.x, q.y + s * n.y, q.z + s * n.z, q.w);
}

hipError_t launch_pins(hipStream_t st, float4 *P, const uint32_t *pin_v, const float4 *pin_t, uint32_t npins,
                       const SubstepArgs &a) {
    if (!npins) return hipSuccess;
    hipLaunchKernelGGL(k_pins, dim3((npins + 255) / 256), dim3(256), 0, st, P, pin_v, pin_t, npins, a);
    return hipGetLastError();
}

// P[idx[k]].w = Palt[idx[k]].w = val[k]
__global__ void __launch_bounds__(256) k_scatter_inv_mass(float4 *__restrict__ P, float4 *__restrict__ Palt,
                                                          const uint32_t *__restrict__ idx,
                                                          const float *__restrict__ val, uint32_t n) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const uint32_t v = idx[t];
    P[v].w = val[t];
    Palt[v].w = val[t];
}

}  // namespace tc

namespace {

float clamped_weight(float w) { return std::isfinite(w) ? std::min(std::max(w, 0.0f), 1.0f) : 0.0f; }

// Writes the inverse masses of `idx` (mod.rs:2785-2789: queue writes, ordered before the next step).
int scatter_inv_masses(tc_sim *s, const std::vector<uint32_t> &idx, const std::vector<float> &val) {
    if (idx.empty()) return TC_OK;
    for (size_t k = 0; k < idx.size(); ++k) s->inv_eff[idx[k]] = val[k];
    TC_HIP(hipStreamSynchronize(s->stream));
    if (s->scat_idx.n < idx.size()) {
        TC_HIP(s->scat_idx.alloc(idx.size()));
        TC_HIP(s->scat_val.alloc(idx.size()));
    }
    TC_HIP(hipMemcpy(s->scat_idx.p, idx.data(), 4 * idx.size(), hipMemcpyHostToDevice));
    TC_HIP(hipMemcpy(s->scat_val.p, val.data(), 4 * val.size(), hipMemcpyHostToDevice));
    const uint32_t n = uint32_t(idx.size());
    hipLaunchKernelGGL(tc::k_scatter_inv_mass, dim3((n + 255) / 256), dim3(256), 0, s->stream, s->P.p, s->Palt.p,
                       s->scat_idx.p, s->scat_val.p, n);
    TC_HIP(hipGetLastError());
    TC_HIP(hipStreamSynchronize(s->stream));
    return TC_OK;
}

}  // namespace

// Uploads the pin table (upload_pins, mod.rs:2930-2970).  The capacity grows to a power of two; a
// new buffer or a new count changes what the captured graphs bake in.
int upload_pins(tc_sim *s) {
    const uint32_t n = uint32_t(s->pins.size());
    std::vector<uint32_t> v;
    std::vector<float4> t;
    v.reserve(n);
    t.reserve(n);
    for (const auto &kv : s->pins) {
        v.push_back(kv.first);
        t.push_back(make_float4(kv.second.target[0], kv.second.target[1], kv.second.target[2], kv.second.weight));
    }
    TC_HIP(hipStreamSynchronize(s->stream));
    if (n > s->pin_cap) {
        uint32_t cap = 256;
        while (cap < n) cap <<= 1;
        TC_HIP(s->pin_v.alloc(cap));
        TC_HIP(s->pin_t.alloc(cap));
        s->pin_cap = cap;
        s->touch();
    }
    if (n) {
        TC_HIP(hipMemcpy(s->pin_v.p, v.data(), 4 * n, hipMemcpyHostToDevice));
        TC_HIP(hipMemcpy(s->pin_t.p, t.data(), 16 * n, hipMemcpyHostToDevice));
    }
    if (n != s->npins) {
        s->npins = n;
        s->touch();
    }
    return TC_OK;
}

hipError_t enqueue_pins(tc_sim *s, float4 *P, const tc::SubstepArgs &a) {
    return tc::launch_pins(s->stream, P, s->pin_v.p, s->pin_t.p, s->npins, a);
}

// The device inverse masses a set_positions / reset leaves (dispatch.rs:972-980): the original
// inverse mass, scaled by (1 - w) for a pinned vertex.
float effective_inv_mass(const tc_sim *s, uint32_t i) {
    const auto it = s->pins.find(i);
    return it == s->pins.end() ? s->inv_orig[i] : s->inv_orig[i] * (1.0f - clamped_weight(it->second.weight));
}

extern "C" {

int tc_sim_set_pins(tc_sim *s, const uint32_t *idx, const float *targets, const float *weights, uint32_t n) {
    if (!s) return fail(TC_ERR_VALUE, "null simulator");
    if (n && (!idx || !targets || !weights)) return fail(TC_ERR_VALUE, "indices, targets and weights are required");
    std::vector<uint32_t> si;
    std::vector<float> sv;
    for (uint32_t k = 0; k < n; ++k) {
        const uint32_t v = idx[k];
        if (v >= s->mesh.nv) continue;  // out-of-range vertices are ignored (mod.rs:2805)
        tc_sim::PinRec r;
        std::memcpy(r.target, targets + 3 * size_t(k), 12);
        r.weight = weights[k];
        s->pins[v] = r;
        si.push_back(v);
        sv.push_back(s->inv_orig[v] * (1.0f - clamped_weight(weights[k])));
    }
    if (int rc = upload_pins(s)) return rc;
    return scatter_inv_masses(s, si, sv);
}

int tc_sim_release_pins(tc_sim *s, const uint32_t *idx, uint32_t n) {
    if (!s) return fail(TC_ERR_VALUE, "null simulator");
    if (n && !idx) return fail(TC_ERR_VALUE, "indices are required");
    std::vector<uint32_t> si;
    std::vector<float> sv;
    for (uint32_t k = 0; k < n; ++k) {
        if (s->pins.erase(idx[k]) == 0) continue;
        si.push_back(idx[k]);
        sv.push_back(s->inv_orig[idx[k]]);
    }
    if (si.empty()) return TC_OK;  // nothing removed: no upload (mod.rs:2859)
    if (int rc = upload_pins(s)) return rc;
    return scatter_inv_masses(s, si, sv);
}

int tc_sim_clear_pins(tc_sim *s) {
    if (!s) return fail(TC_ERR_VALUE, "null simulator");
    std::vector<uint32_t> si;
    std::vector<float> sv;
    for (const auto &kv : s->pins) {
        si.push_back(kv.first);
        sv.push_back(s->inv_orig[kv.first]);
    }
    s->pins.clear();
    if (int rc = upload_pins(s)) return rc;
    return scatter_inv_masses(s, si, sv);
}

int tc_sim_pin_count(const tc_sim *s, uint32_t *n) {
    if (!s || !n) return fail(TC_ERR_VALUE, "null argument");
    *n = uint32_t(s->pins.size());
    return TC_OK;
}

int tc_sim_get_pins(const tc_sim *s, uint32_t *idx, float *targets, float *weights, uint32_t cap, uint32_t *count) {
    if (!s || !count) return fail(TC_ERR_VALUE, "null argument");
    *count = uint32_t(s->pins.size());
    uint32_t k = 0;
    for (const auto &kv : s->pins) {
        if (k >= cap) break;
        if (idx) idx[k] = kv.first;
        if (targets) std::memcpy(targets + 3 * size_t(k), kv.second.target, 12);
        if (weights) weights[k] = kv.second.weight;
        ++k;
    }
    return TC_OK;
}

int tc_sim_set_vertex_inv_masses(tc_sim *s, const uint32_t *idx, const float *inv_masses, uint32_t n) {
    if (!s) return fail(TC_ERR_VALUE, "null simulator");
    if (n && (!idx || !inv_masses)) return fail(TC_ERR_VALUE, "indices and inverse masses are required");
    std::vector<uint32_t> si;
    std::vector<float> sv;
    for (uint32_t k = 0; k < n; ++k) {
        const uint32_t v = idx[k];
        if (v >= s->mesh.nv) continue;
        s->inv_orig[v] = inv_masses[k];
        si.push_back(v);
        sv.push_back(effective_inv_mass(s, v));
    }
    return scatter_inv_masses(s, si, sv);
}

int tc_sim_get_vertex_inv_masses(const tc_sim *s, float *out, size_t cap) {
    if (!s) return fail(TC_ERR_VALUE, "null simulator");
    if (cap < s->inv_orig.size() || (!out && !s->inv_orig.empty()))
        return fail(TC_ERR_VALUE, "output buffer must hold num_vertices floats");
    if (!s->inv_orig.empty()) std::memcpy(out, s->inv_orig.data(), 4 * s->inv_orig.size());
    return TC_OK;
}

}  // extern "C"
