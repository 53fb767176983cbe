// Frame-level host policies of the simulator: adaptive substeps and iterations, asynchronous and
// buffered position readback.
//
// Reference semantics (paths relative to the reference repository root):
//   compute_effective_substeps / compute_effective_iterations
//                       crates/cloth_core/src/simulation/dispatch.rs:476-607
//   feedback from the positions read back (max displacement, P95 strain sample)
//                       dispatch.rs:355-435
//   step / step_async / fetch_positions_flat
//                       dispatch.rs:610-671, 738-863
//   step_buffered / fetch_buffered_positions / clear_frame_buffer
//                       dispatch.rs:1293-1403
//   mesh constants     crates/cloth_core/src/simulation/mod.rs:380-420
//   setters            simulation/mod.rs:723-790
//
// The frame itself is the simulator's graph-replayed frame (tc_sim.hip); the positions go through
// the extract kernel into the simulator's device staging array and from there, asynchronously,
// into pinned host memory, with a HIP event marking when a copy has landed.  The policy reads the
// feedback of the last copy that has landed when the next step starts, without waiting for one
// still in flight, as the reference polls its mapped buffers.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "tc_sim_impl.h"

namespace {

// simulation/mod.rs:380-420: characteristic length for the CFL margin, mean edge length and the
// mesh diameter in edge hops, from the rest lengths of the distance constraints.
void mesh_constants(tc_sim *s) {
    const tc::Mesh &m = s->mesh;
    float min_sq = 3.4e38f;
    double sum = 0.0;
    for (float L : m.d_rest) {
        const float d2 = L * L;
        if (d2 > 1e-8f && d2 < min_sq) min_sq = d2;
        sum += L;
    }
    const float min_edge = min_sq < 3.4e38f ? std::sqrt(min_sq) : 0.02f;
    s->mesh_char_len = std::max(std::min(s->cloth_thickness * 2.0f, min_edge), 0.001f);
    float lo[3] = {3.4e38f, 3.4e38f, 3.4e38f}, hi[3] = {-3.4e38f, -3.4e38f, -3.4e38f};
    for (uint32_t i = 0; i < m.nv; ++i)
        for (int a = 0; a < 3; ++a) {
            const float v = m.pos0[3 * size_t(i) + a];
            lo[a] = std::min(lo[a], v);
            hi[a] = std::max(hi[a], v);
        }
    float diag = 0.0f;
    if (m.nv)
        diag = std::sqrt((hi[0] - lo[0]) * (hi[0] - lo[0]) + (hi[1] - lo[1]) * (hi[1] - lo[1]) +
                         (hi[2] - lo[2]) * (hi[2] - lo[2]));
    s->mesh_avg_edge_len = m.d_rest.empty() ? 0.02f : std::max(float(sum / double(m.d_rest.size())), 1e-4f);
    const double hops = std::ceil(double(diag) / double(s->mesh_avg_edge_len));
    s->mesh_diameter_hops = uint32_t(std::min(std::max(hops, 10.0), 60.0));
}

// dispatch.rs:355-435: the largest per-axis displacement since the last feedback, and the 95th
// percentile of the positive strains over at most about 2048 distance constraints taken with a
// fixed stride; the positions become the next reference.
void update_feedback(tc_sim *s, const float *pos) {
    const tc::Mesh &m = s->mesh;
    const size_t n = m.nv;
    float max_d = 0.0f;
    if (s->prev_step_pos.size() == 3 * n) {
        for (size_t i = 0; i < 3 * n; i += 3) {
            const float d = std::max(std::fabs(pos[i] - s->prev_step_pos[i]),
                                     std::max(std::fabs(pos[i + 1] - s->prev_step_pos[i + 1]),
                                              std::fabs(pos[i + 2] - s->prev_step_pos[i + 2])));
            max_d = std::max(max_d, d);
        }
    }
    float rep = 0.0f;
    const size_t nd = m.d_v0.size();
    if (nd) {
        const float min_ref = std::max(0.5f * s->mesh_avg_edge_len, 0.002f);
        const size_t stride = std::max<size_t>(nd / 2048, 1);
        std::vector<float> pos_strain;
        pos_strain.reserve(nd / stride + 1);
        for (size_t k = 0; k < nd; k += stride) {
            const float l0 = m.d_rest[k];
            if (l0 <= 1e-6f) continue;
            const size_t a = 3 * size_t(m.d_v0[k]), b = 3 * size_t(m.d_v1[k]);
            const float dx = pos[a] - pos[b], dy = pos[a + 1] - pos[b + 1], dz = pos[a + 2] - pos[b + 2];
            const float d2 = dx * dx + dy * dy + dz * dz;
            if (d2 > l0 * l0) {
                const float strain = (std::sqrt(d2) - l0) / std::max(l0, min_ref);
                if (strain > 0.0f) pos_strain.push_back(strain);
            }
        }
        if (!pos_strain.empty()) {
            size_t i95 = size_t(float(pos_strain.size()) * 0.95f);
            i95 = std::min(i95, pos_strain.size() - 1);
            std::nth_element(pos_strain.begin(), pos_strain.begin() + i95, pos_strain.end());
            rep = pos_strain[i95];
        }
    }
    s->last_max_strain = rep;
    s->prev_step_pos.assign(pos, pos + 3 * n);
    s->last_max_disp = max_d;
}

hipError_t ensure_host(tc_sim::HostReadback &b, size_t floats) {
    if (!b.host) {
        hipError_t e = hipHostMalloc(reinterpret_cast<void **>(&b.host), sizeof(float) * std::max<size_t>(floats, 1),
                                     hipHostMallocDefault);
        if (e != hipSuccess) return e;
    }
    if (!b.ev) return hipEventCreateWithFlags(&b.ev, hipEventDisableTiming);
    return hipSuccess;
}

// Positions of the last frame -> pinned host buffer `b`, asynchronously; its event marks arrival.
hipError_t enqueue_readback(tc_sim *s, tc_sim::HostReadback &b) {
    const uint32_t n = s->mesh.nv;
    hipError_t e = ensure_host(b, size_t(n) * 3);
    if (e != hipSuccess) return e;
    if ((e = tc::launch_extract_positions(s->stream, s->X.p, s->readback.p, n)) != hipSuccess) return e;
    if ((e = hipMemcpyAsync(b.host, s->readback.p, sizeof(float) * 3 * n, hipMemcpyDeviceToHost, s->stream)) !=
        hipSuccess)
        return e;
    if ((e = hipEventRecord(b.ev, s->stream)) != hipSuccess) return e;
    b.pending = true;
    return hipSuccess;
}

// Consumes the feedback copies that have landed (oldest first), without waiting.
void poll_feedback(tc_sim *s) {
    for (uint32_t k = 0; k < 2; ++k) {
        tc_sim::HostReadback &b = s->fb[(s->fb_next + k) & 1u];
        if (!b.pending) continue;
        if (hipEventQuery(b.ev) != hipSuccess) continue;  // still in flight (or failed: tried again later)
        b.pending = false;
        update_feedback(s, b.host);
    }
}

}  // namespace

void frames_init(tc_sim *s) {
    mesh_constants(s);
    const tc::Mesh &m = s->mesh;
    s->prev_step_pos.assign(m.pos0.begin(), m.pos0.begin() + 3 * size_t(m.nv));
}

namespace {

// dispatch.rs:476-570
uint32_t effective_substeps(tc_sim *s, float dt, uint32_t requested) {
    if (!s->adaptive) {
        s->eff_substeps = requested > 0 ? requested : s->base_substeps;
        return s->eff_substeps;
    }
    poll_feedback(s);
    const uint32_t min_steps = std::max(s->min_substeps, 1u);
    const uint32_t max_steps = std::max(s->max_substeps, min_steps);
    if (dt <= 0.0f || s->mesh.nv == 0) {
        s->eff_substeps = min_steps;
        return min_steps;
    }
    const float external = s->pending_external;
    const float max_disp = std::max(s->last_max_disp, external);
    s->pending_external = 0.0f;
    s->last_eff_disp = max_disp;
    const float cfl_margin = std::max(std::min(0.5f * s->mesh_char_len, s->cloth_thickness), 0.0025f);
    const uint32_t cfl_steps =
        (cfl_margin > 1e-6f && max_disp > 1e-6f) ? uint32_t(std::ceil(max_disp / cfl_margin)) : min_steps;
    const float g = std::sqrt(s->gravity[0] * s->gravity[0] + s->gravity[1] * s->gravity[1] +
                              s->gravity[2] * s->gravity[2]);
    const float freefall = 0.5f * g * dt * dt;
    const uint32_t grav_steps =
        (cfl_margin > 1e-6f && freefall > 1e-6f) ? uint32_t(std::ceil(std::sqrt(freefall / cfl_margin))) : 1u;
    uint32_t strain_steps = min_steps;
    if (s->strain_adaptive && s->strain_tolerance > 1e-5f) {
        const float ratio = s->last_max_strain / s->strain_tolerance;
        if (ratio > 1.25f) strain_steps = uint32_t(std::ceil(float(min_steps) * std::min(ratio, 2.0f)));
    }
    const uint32_t target = std::min(std::max(std::max(cfl_steps, std::max(grav_steps, strain_steps)), min_steps),
                                     max_steps);
    uint32_t cur;
    if (s->eff_substeps == 0) {
        cur = target;
    } else {
        const uint32_t prev = std::min(std::max(s->eff_substeps, min_steps), max_steps);
        const int diff = int(target) - int(prev);
        if (std::abs(diff) <= 1) {
            cur = prev;  // deadband
        } else if (diff > 0) {
            const bool emergency = external > 1e-4f || max_disp > cfl_margin * 4.0f ||
                                   (s->strain_adaptive && s->last_max_strain > s->strain_tolerance * 3.0f);
            cur = emergency ? target : std::min(target, prev + 2);
        } else {
            cur = std::max(target, prev > 2 ? prev - 2 : 0u);
        }
    }
    cur = std::min(std::max(cur, min_steps), max_steps);
    s->eff_substeps = cur;
    return cur;
}

// dispatch.rs:573-607
uint32_t effective_iterations(tc_sim *s, uint32_t substeps) {
    if (!s->adaptive || !s->auto_compensate) {
        s->eff_iters = 0;
        return s->iterations;
    }
    const uint32_t budget = std::max(s->base_substeps, 1u) * std::max(s->iterations, 1u);
    const uint32_t floor_hops = std::min(std::max(s->mesh_diameter_hops, 15u), 30u);
    const uint32_t base_target = std::min(budget, std::max(floor_hops, 20u));
    uint32_t target = base_target;
    if (s->strain_adaptive && s->strain_tolerance > 1e-5f) {
        const float ratio = s->last_max_strain / s->strain_tolerance;
        if (ratio > 1.0f) target = uint32_t(std::ceil(float(base_target) * std::min(ratio, 2.0f)));
    }
    const uint32_t comp = (target + substeps - 1) / std::max(substeps, 1u);
    const uint32_t max_iters = std::max(s->iterations * 4, 8u);
    const uint32_t eff = std::min(std::max(comp, std::max(s->iterations, 1u)), max_iters);
    s->eff_iters = eff;
    return eff;
}

// The frame of step / step_async / step_buffered: the policy picks substeps and iterations.
int policy_frame(tc_sim *s, float dt, uint32_t requested, uint32_t *substeps_out) {
    const uint32_t S = effective_substeps(s, dt, requested);
    effective_iterations(s, S);
    if (substeps_out) *substeps_out = S;
    const int r = dyn_sdf_frame(s);
    if (r != TC_OK) return r;
    return tc_run_frame(s, dt, S);
}

}  // namespace

int tc_sim_step(tc_sim *s, float dt, uint32_t substeps, uint32_t iterations, uint32_t *eff) {
    if (!s) return fail(TC_ERR_VALUE, "null simulator");
    if (iterations > 0) s->iterations = iterations;
    if (s->mesh.nv == 0) {  // dispatch.rs:611-613
        if (eff) *eff = substeps > 0 ? substeps : s->base_substeps;
        return TC_OK;
    }
    uint32_t S = 0;
    int r = policy_frame(s, dt, substeps, &S);
    if (eff) *eff = S;
    if (r != TC_OK) return r;
    // dispatch.rs:628-650: with the policy on, the frame's positions go to a free feedback buffer
    if (s->adaptive && s->compact_readback) {
        tc_sim::HostReadback &b = s->fb[s->fb_next];
        if (!b.pending) {
            TC_HIP(enqueue_readback(s, b));
            s->fb_next ^= 1u;
        }
    }
    return TC_OK;
}

int tc_sim_step_async(tc_sim *s, float dt, uint32_t substeps, uint32_t iterations) {
    if (!s) return fail(TC_ERR_VALUE, "null simulator");
    if (iterations > 0) s->iterations = iterations;
    if (s->mesh.nv == 0) return TC_OK;
    if (s->async_pending) {  // dispatch.rs:747-750: the previous frame's positions are collected first
        std::vector<float> dummy(3 * size_t(s->mesh.nv));
        int ok = 0;
        int r = tc_sim_fetch_positions(s, dummy.data(), dummy.size(), &ok);
        if (r != TC_OK) return r;
    }
    int r = policy_frame(s, dt, substeps, nullptr);
    if (r != TC_OK) return r;
    tc_sim::HostReadback &b = s->as[s->as_next];
    TC_HIP(enqueue_readback(s, b));
    s->as_ready = s->as_next;
    s->as_next ^= 1u;
    s->async_pending = true;
    return TC_OK;
}

int tc_sim_fetch_positions(tc_sim *s, float *out, size_t n_floats, int *ok) {
    if (!s) return fail(TC_ERR_VALUE, "null simulator");
    if (ok) *ok = 0;
    if (!s->async_pending || s->mesh.nv == 0) return TC_OK;
    if (!out) return fail(TC_ERR_VALUE, "output buffer is null");
    tc_sim::HostReadback &b = s->as[s->as_ready];
    TC_HIP(hipEventSynchronize(b.ev));
    b.pending = false;
    const size_t n3 = 3 * size_t(s->mesh.nv);
    std::memcpy(out, b.host, sizeof(float) * std::min(n_floats, n3));
    if (n_floats >= n3) update_feedback(s, b.host);
    s->async_pending = false;
    if (ok) *ok = 1;
    return TC_OK;
}

int tc_sim_step_buffered(tc_sim *s, float dt, uint32_t substeps, uint32_t iterations, uint32_t *count) {
    if (!s) return fail(TC_ERR_VALUE, "null simulator");
    if (iterations > 0) s->iterations = iterations;
    if (count) *count = 0;
    if (s->mesh.nv == 0) return TC_OK;
    const size_t n3 = 3 * size_t(s->mesh.nv);
    if (!s->ring.p) TC_HIP(s->ring.alloc(n3 * s->ring_frames));
    const uint32_t slot = std::min(s->ring_count, s->ring_frames - 1);
    int r = policy_frame(s, dt, substeps, nullptr);
    if (r != TC_OK) return r;
    TC_HIP(tc::launch_extract_positions(s->stream, s->X.p, s->ring.p + size_t(slot) * n3, s->mesh.nv));
    if (s->ring_count < s->ring_frames) ++s->ring_count;
    if (count) *count = s->ring_count;
    return TC_OK;
}

int tc_sim_fetch_buffered_positions(tc_sim *s, float *out, size_t n_floats, uint32_t *count) {
    if (!s) return fail(TC_ERR_VALUE, "null simulator");
    if (count) *count = 0;
    const uint32_t c = s->ring_count;
    if (c == 0 || s->mesh.nv == 0) return TC_OK;
    if (!out) return fail(TC_ERR_VALUE, "output buffer is null");
    const size_t n3 = 3 * size_t(s->mesh.nv), total = size_t(c) * n3;
    if (!s->ring_host)
        TC_HIP(hipHostMalloc(reinterpret_cast<void **>(&s->ring_host), sizeof(float) * n3 * s->ring_frames,
                             hipHostMallocDefault));
    TC_HIP(hipMemcpyAsync(s->ring_host, s->ring.p, sizeof(float) * total, hipMemcpyDeviceToHost, s->stream));
    TC_HIP(hipStreamSynchronize(s->stream));
    const size_t copy = std::min(n_floats, total);
    std::memcpy(out, s->ring_host, sizeof(float) * copy);
    if (copy >= total) update_feedback(s, s->ring_host + (size_t(c) - 1) * n3);
    s->ring_count = 0;
    if (count) *count = c;
    return TC_OK;
}

int tc_sim_get_buffered_frame_count(const tc_sim *s, uint32_t *count) {
    if (!s || !count) return fail(TC_ERR_VALUE, "null argument");
    *count = s->ring_count;
    return TC_OK;
}

int tc_sim_clear_frame_buffer(tc_sim *s) {
    if (!s) return fail(TC_ERR_VALUE, "null simulator");
    s->ring_count = 0;
    return TC_OK;
}

int tc_sim_set_adaptive_substep_options(tc_sim *s, int enabled, uint32_t min_substeps, uint32_t max_substeps) {
    if (!s) return fail(TC_ERR_VALUE, "null simulator");
    const uint32_t lo = std::max(min_substeps, 1u), hi = std::max(max_substeps, lo);
    const bool was = s->adaptive;
    s->adaptive = enabled != 0;
    s->min_substeps = lo;
    s->max_substeps = hi;
    // simulation/mod.rs:733-743: the history survives re-applying the options while enabled
    if (!s->adaptive || !was) {
        s->eff_substeps = 0;
        s->pending_external = 0.0f;
    } else if (s->eff_substeps > 0) {
        s->eff_substeps = std::min(std::max(s->eff_substeps, lo), hi);
    }
    if (!s->adaptive) s->eff_iters = 0;
    return TC_OK;
}

int tc_sim_set_auto_compensate_iterations(tc_sim *s, int enabled) {
    if (!s) return fail(TC_ERR_VALUE, "null simulator");
    s->auto_compensate = enabled != 0;
    return TC_OK;
}

int tc_sim_set_strain_adaptive_options(tc_sim *s, int enabled, float tolerance) {
    if (!s) return fail(TC_ERR_VALUE, "null simulator");
    s->strain_adaptive = enabled != 0;
    s->strain_tolerance = std::min(std::max(tolerance, 0.0001f), 0.1f);
    return TC_OK;
}

int tc_sim_set_enable_compact_readback(tc_sim *s, int enabled) {
    if (!s) return fail(TC_ERR_VALUE, "null simulator");
    s->compact_readback = enabled != 0;
    return TC_OK;
}

int tc_sim_note_external_displacement(tc_sim *s, float displacement) {
    if (!s) return fail(TC_ERR_VALUE, "null simulator");
    if (displacement > s->pending_external) s->pending_external = displacement;
    return TC_OK;
}

int tc_sim_frame_policy_state(const tc_sim *s, double *out) {
    if (!s || !out) return fail(TC_ERR_VALUE, "null argument");
    out[0] = s->eff_substeps;
    out[1] = s->eff_iters ? s->eff_iters : s->iterations;
    out[2] = s->last_max_strain;
    out[3] = s->last_max_disp;
    out[4] = s->last_eff_disp;
    out[5] = s->mesh_char_len;
    out[6] = s->mesh_diameter_hops;
    out[7] = s->mesh_avg_edge_len;
    return TC_OK;
}
