// Simulator object and C ABI (include/tc_sim.h) of the MI355X XPBD substep library.
//
// The frame loop restates GpuClothSimulator::step / encode_simulation_steps
// (crates/cloth_core/src/simulation/dispatch.rs:7-354, 610-671) with adaptive
// substeps off, decoupled colliders (the path plain step() + get_positions()
// always takes, dispatch.rs:94-102) and self-collision mode 0 / DIRECT.
// One substep enqueues, on the simulator's own HIP stream:
//   predict -> iterations x (distance colours, bending colours)
//   -> collider resolve -> [vertex hash, edge hash, V-V/V-T, E-E, apply]
//   -> update_vel
// The whole frame is captured once into a hipGraph per (dt, substeps,
// iterations, parameter epoch) and replayed; the lambda clear of every substep
// is folded into the first sweep (lambda is read as zero there).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "tc_layout.h"
#include "../../include/tc_sim.h"
#include "tc_kernels.h"
#include "tc_mesh.h"
#include "tc_tiles.h"

#include "tc_sim_impl.h"

namespace tc {
thread_local std::string g_err;
}
using tc::g_err;

namespace {

hipError_t enqueue_hashes(tc_sim *s, const float4 *P) {
    return tc::launch_hashes(s->stream, P, s->mesh.nv, s->edges.p, uint32_t(s->mesh.e_v0.size()), s->cell_size,
                             s->table_size, s->vhash(), s->ehash());
}

// One substep up to (optionally) the velocity update.  Predict writes P buffer 0;
// the per-iteration path ping-pongs P and lambda between two buffers, and the
// rest of the substep works on the buffer holding the latest positions.
}  // namespace

// `count` projection iterations.  The per-iteration path reads P / lambda from buffer
// pc / lc and leaves them in the other one after each launch; the first iteration
// (first == true) treats lambda as zero.
hipError_t enqueue_iterations(tc_sim *s, const tc::SubstepArgs &a, uint32_t count, bool first, uint32_t &pc,
                              uint32_t &lc) {
    const tc::Mesh &m = s->mesh;
    const uint32_t B = s->block;
    hipError_t e = hipSuccess;
    if (s->fused_active()) {
        const tc::FusedArgs fa = s->fused_args();
        const tc::FusedPlanDev pd = s->plan_dev();
        for (uint32_t it = 0; it < count; ++it) {
            const int q = s->prof_begin("solver_iter");
            if (s->npins && (e = enqueue_pins(s, s->Pbuf(pc), a)) != hipSuccess) return e;
            e = tc::launch_tile_iteration(s->stream, s->Pbuf(pc), s->Pbuf(pc ^ 1u), s->X.p, pd, s->dlam_buf(lc),
                                          s->dlam_buf(lc ^ 1u), s->blam_buf(lc), s->blam_buf(lc ^ 1u),
                                          s->slam_buf(lc), s->slam_buf(lc ^ 1u), fa, a, first && it == 0,
                                          s->fused_block);
            s->prof_end(q);
            if (e != hipSuccess) return e;
            pc ^= 1u;
            lc ^= 1u;
        }
        return e;
    }
    for (uint32_t it = 0; it < count; ++it) {
        const bool f = first && it == 0;
        const int q = s->prof_begin("solver_iter");
        if (s->npins && (e = enqueue_pins(s, s->P.p, a)) != hipSuccess) return e;  // dispatch.rs:176-181
        if (s->solver_mode == 1) {  // dispatch.rs:1436-1446
            e = tc::launch_distance_jacobi(s->stream, s->P.p, s->X.p, s->drec.p, s->dcc.p, s->dlam.p,
                                           uint32_t(m.d_v0.size()), s->jacobi_acc.p, m.nv, a, f, s->split_compliance);
            if (e != hipSuccess) return e;
        } else {
            for (size_t c = 0; c < m.d_col.count.size(); ++c) {
                e = tc::launch_distance(s->stream, s->P.p, s->X.p, s->drec.p, s->dcc.p, s->dlam.p, m.d_col.offset[c],
                                        m.d_col.count[c], a, f, s->split_compliance, B);
                if (e != hipSuccess) return e;
            }
        }
        if (!m.s_v0.empty() && (e = enqueue_sewing(s, s->P.p, a.dt, s->slam.p)) != hipSuccess) return e;  // dispatch.rs:183-191
        for (size_t c = 0; c < m.b_col.count.size(); ++c) {
            e = tc::launch_bending(s->stream, s->P.p, s->X.p, s->brec.p, s->blam.p, m.b_col.offset[c],
                                   m.b_col.count[c], a, f, B, s->bgate.p);
            if (e != hipSuccess) return e;
        }
        s->prof_end(q);
    }
    return e;
}

// The check after the budget's last chunk: it runs only when that chunk ran (the gate still open)
// and counts the substeps that end above the threshold; nothing reads its flag.
hipError_t enqueue_final_check(tc_sim *s, const tc::SubstepArgs &a, uint32_t pc, uint32_t lc) {
    const tc::Mesh &m = s->mesh;
    const int q = s->prof_begin("residual_gate");
    hipError_t e = tc::launch_residual(s->stream, s->Pbuf(pc), s->X.p, s->drec.p, s->dcc.p, s->dlam_buf(lc),
                                       uint32_t(m.d_v0.size()), s->brec.p, s->blam_buf(lc), uint32_t(m.b_v0.size()),
                                       a, s->split_compliance, s->resid_part.p, s->ws_res.p, s->ws_gate.p,
                                       s->ws_threshold, s->ws_gate.p, nullptr, true);
    s->prof_end(q);
    return e;
}

// Residual-gated fallback: chunks of the warm-start iteration count (rounded up to even, so that
// the P / lambda buffer parity is the same whether a chunk runs or not), each preceded by a
// residual check; the first check that passes closes the gate for the rest of the substep.
hipError_t enqueue_gated_fallback(tc_sim *s, const tc::SubstepArgs &a, float dt_sub, uint32_t &pc, uint32_t &lc) {
    const tc::Mesh &m = s->mesh;
    const uint32_t chunk = std::max<uint32_t>(2u, s->ws_iters + (s->ws_iters & 1u));
    tc::SubstepArgs fa = a;
    fa.run_flag = s->ws_gate.p;
    hipError_t e = hipSuccess;
    for (uint32_t done = 0; done < s->ws_fallback; done += chunk) {
        const int q = s->prof_begin("residual_gate");
        e = tc::launch_residual(s->stream, s->Pbuf(pc), s->X.p, s->drec.p, s->dcc.p, s->dlam_buf(lc), uint32_t(m.d_v0.size()),
                                s->brec.p, s->blam_buf(lc), uint32_t(m.b_v0.size()), a, s->split_compliance,
                                s->resid_part.p, s->ws_res.p, s->ws_gate.p, s->ws_threshold,
                                done == 0 ? nullptr : s->ws_gate.p);
        s->prof_end(q);
        if (e != hipSuccess) return e;
        if ((e = enqueue_iterations(s, fa, std::min(chunk, s->ws_fallback - done), false, pc, lc)) != hipSuccess)
            return e;
    }
    return enqueue_final_check(s, a, pc, lc);
}

namespace {

// The substep up to the end of the warm iterations.  With a warm start and a fallback, either the
// whole gated fallback follows on the device (gate_to_host false) or only the first residual
// check, whose kernel also writes the gate's fallback flag to s->gate_host for the host to read.
hipError_t enqueue_substep_head(tc_sim *s, float dt_sub, bool gate_to_host, uint32_t &pc, uint32_t &lc) {
    const tc::SubstepArgs a = s->args(dt_sub);
    const tc::Mesh &m = s->mesh;
    const uint32_t n = m.nv, B = s->block;
    int q;
    hipError_t e;
    if (!m.s_v0.empty()) {  // the rest lengths shrink once per substep; the sewing lambdas clear (dispatch.rs:117-134)
        q = s->prof_begin("sew_shrink");
        e = enqueue_sew_shrink(s, dt_sub);
        s->prof_end(q);
        if (e != hipSuccess) return e;
    }
    q = s->prof_begin("predict");
    e = tc::launch_predict(s->stream, s->P.p, s->X.p, s->V.p, n, a, B);
    s->prof_end(q);
    if (e != hipSuccess) return e;
    if (s->wrinkle_on && s->wrinkle_tex.p && s->priority_scale() > 0.0f) {  // dispatch.rs:151-165
        q = s->prof_begin("wrinkle_field");
        e = enqueue_wrinkle(s, dt_sub);
        s->prof_end(q);
        if (e != hipSuccess) return e;
    }
    if (s->ws_tracking()) {
        q = s->prof_begin("warm_start");
        e = ws_pre(s, dt_sub);
        s->prof_end(q);
        if (e != hipSuccess) return e;
    }
    const uint32_t iters = s->iters_effective();
    pc = 0;
    lc = 0;
    if (iters == 0) {  // lambdas are still cleared at the substep start (dispatch.rs:117-119)
        if ((e = tc::launch_fill_u32(s->stream, reinterpret_cast<uint32_t *>(s->dlam.p), 0u, m.d_v0.size())) != hipSuccess) return e;
        if ((e = tc::launch_fill_u32(s->stream, reinterpret_cast<uint32_t *>(s->blam.p), 0u, m.b_v0.size())) != hipSuccess) return e;
        return e;
    }
    if ((e = enqueue_iterations(s, a, iters, true, pc, lc)) != hipSuccess) return e;
    if (!s->ws_active() || s->ws_fallback == 0) return e;
    if (!gate_to_host) return enqueue_gated_fallback(s, a, dt_sub, pc, lc);
    q = s->prof_begin("residual_gate");
    e = tc::launch_residual(s->stream, s->Pbuf(pc), s->X.p, s->drec.p, s->dcc.p, s->dlam_buf(lc), uint32_t(m.d_v0.size()),
                            s->brec.p, s->blam_buf(lc), uint32_t(m.b_v0.size()), a, s->split_compliance,
                            s->resid_part.p, s->ws_res.p, s->ws_gate.p, s->ws_threshold, nullptr, s->gate_host);
    s->prof_end(q);
    return e;
}

// One fallback chunk of `count` iterations (gated on the device flag as well), then, unless it
// is the last chunk of the budget, the next residual check and the flag's copy to the host.  The
// same kernels in the same order as enqueue_gated_fallback, minus those the gate would skip.
hipError_t enqueue_fallback_chunk(tc_sim *s, float dt_sub, uint32_t count, bool last, uint32_t &pc, uint32_t &lc) {
    const tc::Mesh &m = s->mesh;
    tc::SubstepArgs fa = s->args(dt_sub);
    fa.run_flag = s->ws_gate.p;
    hipError_t e = enqueue_iterations(s, fa, count, false, pc, lc);
    if (e != hipSuccess) return e;
    if (last) return enqueue_final_check(s, fa, pc, lc);
    e = tc::launch_residual(s->stream, s->Pbuf(pc), s->X.p, s->drec.p, s->dcc.p, s->dlam_buf(lc), uint32_t(m.d_v0.size()),
                            s->brec.p, s->blam_buf(lc), uint32_t(m.b_v0.size()), fa, s->split_compliance,
                            s->resid_part.p, s->ws_res.p, s->ws_gate.p, s->ws_threshold, s->ws_gate.p,
                            s->gate_host);
    return e;
}

// The substep after the projection loop: the warm start's bookkeeping, colliders, self-collision
// and the velocity update, on the buffers (pc, lc) that hold the latest positions and lambdas.
hipError_t enqueue_substep_tail(tc_sim *s, float dt_sub, uint32_t pc, uint32_t lc, bool collide, bool self_collide,
                                bool update) {
    const tc::Mesh &m = s->mesh;
    const uint32_t n = m.nv, B = s->block;
    float4 *Pc = s->Pbuf(pc);
    s->last_p = pc;
    s->last_lam = lc;
    hipError_t e = hipSuccess;
    int q;
    if (s->ws_tracking()) {
        q = s->prof_begin("warm_start_update");
        e = ws_post(s, dt_sub, Pc);
        s->prof_end(q);
        if (e != hipSuccess) return e;
    }
    if (collide && (!s->cols_h.empty() || !s->tris_h.empty() || s->sdf_nb > 0)) {
        q = s->prof_begin("collider_decoupled");
        e = tc::launch_collision(s->stream, Pc, s->X.p, n, s->col_order.p, s->cols.p, uint32_t(s->cols_h.size()), s->tris.p,
                                 s->tri_bounds.p, uint32_t(s->tris_h.size()), s->tri_cbox.p, s->single_sided_recovery, dt_sub, B, s->sdf_args());
        s->prof_end(q);
        if (e != hipSuccess) return e;
    }
    if (s->npins) {  // the final pin after the iterations and the decoupled collider (dispatch.rs:238-250)
        q = s->prof_begin("final_pin");
        e = enqueue_pins(s, Pc, s->args(dt_sub));
        s->prof_end(q);
        if (e != hipSuccess) return e;
    }
    // collider-edge collision and its apply (dispatch.rs:252-278)
    if (collide && s->edge_collision && !s->mesh.e_v0.empty() &&
        (!s->cols_h.empty() || !s->tris_h.empty() || s->sdf_nb > 0)) {
        q = s->prof_begin("edge_collision");
        e = enqueue_edge_collision(s, Pc);
        s->prof_end(q);
        if (e != hipSuccess) return e;
    }
    if (self_collide && s->self_collision && s->pair_cache && !(s->pair_final_fallback && s->sub_last)) {
        // dispatch.rs:1554-1645: collect on the frame's first substep, then solve the cached pairs
        const int scope = s->prof_begin("self_collision_Outer");
        const tc::SelfColArgs sa = s->sc_args();
        const tc::SelfColTopo t = s->topo();
        if (sa.untangle) {
            q = s->prof_begin("sc_normals");
            e = tc::launch_compute_normals(s->stream, s->X.p, t, s->normals.p, n);
            s->prof_end(q);
            if (e != hipSuccess) return e;
        }
        if (s->sub_first) {
            q = s->prof_begin("hash_build");
            e = enqueue_hashes(s, Pc);
            s->prof_end(q);
            if (e != hipSuccess) return e;
            if ((e = tc::launch_fill_u32(s->stream, s->pair_counters.p, 0u, 4)) != hipSuccess) return e;
            q = s->prof_begin("pair_collect");
            e = tc::launch_pair_collect(s->stream, Pc, s->X.p, s->V.p, t, {s->adj_off.p, s->adj_idx.p, s->island.p},
                                        s->vhash(), s->pair_vt.p, s->pair_ee.p, s->pair_counters.p, s->pair_args());
            s->prof_end(q);
            if (e != hipSuccess) return e;
        }
        const tc::PairArgs pa = s->pair_args();
        q = s->prof_begin("pair_vt");
        e = tc::launch_pair_solve_vt(s->stream, Pc, s->X.p, s->pair_vt.p, n * pa.quota_vt, s->accum.p, n, t,
                                     sa.untangle);
        s->prof_end(q);
        if (e != hipSuccess) return e;
        q = s->prof_begin("pair_ee");
        e = tc::launch_pair_solve_ee(s->stream, Pc, s->X.p, s->pair_ee.p, n * pa.quota_ee, s->accum.p, n, t,
                                     sa.untangle);
        s->prof_end(q);
        if (e != hipSuccess) return e;
        if (sa.untangle) {
            q = s->prof_begin("sc_untangle");
            e = tc::launch_sc_untangle(s->stream, Pc, s->X.p, t, s->accum.p, sa);
            s->prof_end(q);
            if (e != hipSuccess) return e;
        }
        q = s->prof_begin("pair_apply");
        e = tc::launch_sc_apply(s->stream, Pc, s->local_len.p, s->accum.p, sa, B);
        s->prof_end(q);
        s->prof_end(scope);
        if (e != hipSuccess) return e;
    } else if (self_collide && s->self_collision) {
        const int scope = s->prof_begin("self_collision_Outer");
        const tc::SelfColArgs sa = s->sc_args();
        const tc::SelfColTopo t = s->topo();
        if (sa.untangle) {
            q = s->prof_begin("sc_normals");
            e = tc::launch_compute_normals(s->stream, s->X.p, t, s->normals.p, n);
            s->prof_end(q);
            if (e != hipSuccess) return e;
        }
        q = s->prof_begin("hash_build");
        e = enqueue_hashes(s, Pc);
        s->prof_end(q);
        if (e != hipSuccess) return e;
        q = s->prof_begin("sc_solve");
        e = tc::launch_sc_solve(s->stream, Pc, s->X.p, t, s->vhash(), s->accum.p, sa);
        s->prof_end(q);
        if (e != hipSuccess) return e;
        if (sa.untangle) {
            q = s->prof_begin("sc_untangle");
            e = tc::launch_sc_untangle(s->stream, Pc, s->X.p, t, s->accum.p, sa);
            s->prof_end(q);
            if (e != hipSuccess) return e;
        }
        q = s->prof_begin("sc_solve_ee");
        e = tc::launch_sc_solve_ee(s->stream, Pc, s->X.p, t, s->ehash(), s->accum.p, sa);
        s->prof_end(q);
        if (e != hipSuccess) return e;
        q = s->prof_begin("sc_apply");
        e = tc::launch_sc_apply(s->stream, Pc, s->local_len.p, s->accum.p, sa, B);
        s->prof_end(q);
        s->prof_end(scope);
        if (e != hipSuccess) return e;
    }
    if (!m.s_v0.empty()) {  // the final sewing pass before the velocity update (dispatch.rs:318-339)
        q = s->prof_begin("final_sewing");
        e = enqueue_sewing(s, Pc, dt_sub, s->slam_buf(lc));  // the lambdas the iterations left
        s->prof_end(q);
        if (e != hipSuccess) return e;
    }
    if (update) {
        q = s->prof_begin("update_vel");
        e = tc::launch_update_vel(s->stream, Pc, s->X.p, s->V.p, n, dt_sub, B);
        s->prof_end(q);
        return e;
    }
    return hipSuccess;
}

hipError_t enqueue_substep(tc_sim *s, float dt_sub, bool collide, bool self_collide, bool update) {
    uint32_t pc = 0, lc = 0;
    hipError_t e = enqueue_substep_head(s, dt_sub, false, pc, lc);
    if (e != hipSuccess) return e;
    return enqueue_substep_tail(s, dt_sub, pc, lc, collide, self_collide, update);
}

// Picks the training-sample slot the next substep writes (recording only).
void next_record_slot(tc_sim *s) {
    s->rec_slot = -1;
    if (s->rec_cap == 0 || s->rec_count >= s->rec_cap) return;
    if (s->rec_tick++ % s->rec_stride == 0) s->rec_slot = int(s->rec_count++);
}

hipError_t enqueue_frame(tc_sim *s, float dt, uint32_t substeps) {
    const float dt_sub = dt / float(substeps);
    s->pair_dt_frame = dt;
    for (uint32_t k = 0; k < substeps; ++k) {
        s->sub_first = k == 0;
        s->sub_last = k + 1 == substeps;
        next_record_slot(s);
        hipError_t e = enqueue_substep(s, dt_sub, true, true, true);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// Replays `pg` (captured from `enqueue` when its key changes).
template <class Enqueue>
int replay_part(tc_sim *s, tc_sim::PartGraph &pg, const uint64_t key[4], Enqueue enqueue) {
    if (!pg.x || std::memcmp(key, pg.key, sizeof(pg.key)) != 0) {
        pg.clear();
        TC_HIP(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
        hipError_t e = enqueue();
        hipGraph_t g = nullptr;
        hipError_t e2 = hipStreamEndCapture(s->stream, &g);
        if (e != hipSuccess) { if (g) (void)hipGraphDestroy(g); return fail(TC_ERR_RUNTIME, std::string("capture: ") + hipGetErrorString(e)); }
        TC_HIP(e2);
        pg.g = g;
        TC_HIP(hipGraphInstantiate(&pg.x, pg.g, nullptr, nullptr, 0));
        std::memcpy(pg.key, key, sizeof(pg.key));
    }
    TC_HIP(hipGraphLaunch(pg.x, s->stream));
    return TC_OK;
}

// A frame of warm-started substeps with a fallback, from graphs of the substeps' parts.  Captured
// whole, a substep carries every fallback chunk, and a chunk whose gate has passed still costs its
// launches (about 2.7 us each for the per-iteration kernel, 56 of them); here the host launches
// chunks only while the gate stays open, one chunk ahead of the check it waits for: the next chunk
// is queued before the host waits for the previous check, so the device never idles between
// chunks, and a queued chunk whose gate has closed returns at once (its kernels are gated on the
// device flag).  The chunks that run, and their order, are those of the captured frame.  The host
// flag is written by every check, a closed gate stays closed, and everything queued before a
// substep's head has finished when the host reads the head's flag, so any value read after the
// awaited check reflects that check or a later one of the same substep.
int run_frame_host_gated(tc_sim *s, float dt, uint32_t substeps) {
    if (!s->gate_host) TC_HIP(hipHostMalloc(reinterpret_cast<void **>(&s->gate_host), 16, hipHostMallocDefault));
    for (hipEvent_t &ev : s->gate_ev)
        if (!ev) TC_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    const float dt_sub = dt / float(substeps);
    uint32_t dtb;
    std::memcpy(&dtb, &dt_sub, 4);
    const uint64_t fused = s->fused_active() ? 1u : 0u;
    const uint32_t chunk = std::max<uint32_t>(2u, s->ws_iters + (s->ws_iters & 1u));
    uint32_t p0 = 0, l0 = 0;
    s->pair_dt_frame = dt;
    uint32_t dtf;
    std::memcpy(&dtf, &dt, 4);
    for (uint32_t k = 0; k < substeps; ++k) {
        s->sub_first = k == 0;
        s->sub_last = k + 1 == substeps;
        // the tail of the frame's first and last substeps differs with the pair cache
        const uint64_t kind = s->pair_cache ? (uint64_t(s->sub_first) | (uint64_t(s->sub_last) << 1) | (uint64_t(dtf) << 2)) : 0u;
        const uint64_t kh[4] = {dtb, s->ws_iters, fused, s->epoch};
        int r = replay_part(s, s->pg_head, kh, [&] { return enqueue_substep_head(s, dt_sub, true, p0, l0); });
        if (r != TC_OK) return r;
        // the buffers holding P and lambda after the warm iterations (chunks have an even size)
        p0 = l0 = (fused && s->ws_iters) ? (s->ws_iters & 1u) : 0u;
        TC_HIP(hipEventRecord(s->gate_ev[0], s->stream));
        uint32_t launched = 0, ev_next = 1;
        bool last_queued = s->ws_fallback == 0;
        // queues the next chunk and records its completion event; returns the event of the chunk
        auto queue_chunk = [&](hipEvent_t &out) -> int {
            const uint32_t cnt = std::min(chunk, s->ws_fallback - launched);
            const bool last = launched + cnt >= s->ws_fallback;
            const uint64_t kc[4] = {dtb, cnt, fused | (uint64_t(p0) << 1), s->epoch};
            uint32_t pc = p0, lc = l0;
            const int rr = replay_part(s, last ? s->pg_last : s->pg_chunk, kc,
                                       [&] { return enqueue_fallback_chunk(s, dt_sub, cnt, last, pc, lc); });
            if (rr != TC_OK) return rr;
            launched += cnt;
            last_queued = last;
            out = s->gate_ev[ev_next];
            ev_next = ev_next == 1u ? 2u : 1u;
            TC_HIP(hipEventRecord(out, s->stream));
            return TC_OK;
        };
        hipEvent_t awaited = s->gate_ev[0], queued = nullptr;
        if (!last_queued && (r = queue_chunk(queued)) != TC_OK) return r;
        for (;;) {
            TC_HIP(hipEventSynchronize(awaited));
            if (*s->gate_host == 0u || last_queued) break;  // the gate closed, or the budget is queued
            hipEvent_t next = nullptr;
            if ((r = queue_chunk(next)) != TC_OK) return r;
            awaited = queued;
            queued = next;
        }
        const uint64_t kt[4] = {dtb, uint64_t(p0) | (uint64_t(l0) << 1) | (kind << 2), fused, s->epoch};
        r = replay_part(s, s->pg_tail, kt, [&] { return enqueue_substep_tail(s, dt_sub, p0, l0, true, true, true); });
        if (r != TC_OK) return r;
    }
    return TC_OK;
}

int run_frame(tc_sim *s, float dt, uint32_t substeps) {
    const uint32_t iters = s->iters_effective();
    if (s->fused_active() && iters > 0) {  // host bookkeeping of the buffer parity the graph bakes in
        s->last_p = iters & 1u;
        s->last_lam = iters & 1u;
    } else {
        s->last_p = 0;
        s->last_lam = 0;
    }
    if (s->prof.enabled) {  // pass events are recorded between the launches, kernel by kernel
        s->prof.frame_begin();
        s->prof.active = true;
        const hipError_t e = enqueue_frame(s, dt, substeps);
        s->prof.active = false;
        TC_HIP(e);
        return TC_OK;
    }
    if (!s->use_graph || s->rec_cap > 0) {  // recording picks its sample slot per substep on the host
        TC_HIP(enqueue_frame(s, dt, substeps));
        return TC_OK;
    }
    if (s->ws_active() && s->ws_fallback > 0 && iters > 0) return run_frame_host_gated(s, dt, substeps);
    uint32_t dtb;
    std::memcpy(&dtb, &dt, 4);
    const uint64_t key[4] = {dtb, substeps, uint64_t(iters) | (uint64_t(s->fused_active()) << 32), s->epoch};
    if (!s->graph_exec || std::memcmp(key, s->graph_key, sizeof(key)) != 0) {
        if (s->graph_exec) { (void)hipGraphExecDestroy(s->graph_exec); s->graph_exec = nullptr; }
        if (s->graph) { (void)hipGraphDestroy(s->graph); s->graph = nullptr; }
        TC_HIP(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
        hipError_t e = enqueue_frame(s, dt, substeps);
        hipGraph_t g = nullptr;
        hipError_t e2 = hipStreamEndCapture(s->stream, &g);
        if (e != hipSuccess) { if (g) (void)hipGraphDestroy(g); return fail(TC_ERR_RUNTIME, std::string("capture: ") + hipGetErrorString(e)); }
        TC_HIP(e2);
        s->graph = g;
        TC_HIP(hipGraphInstantiate(&s->graph_exec, s->graph, nullptr, nullptr, 0));
        std::memcpy(s->graph_key, key, sizeof(key));
    }
    TC_HIP(hipGraphLaunch(s->graph_exec, s->stream));
    return TC_OK;
}

// (Re)builds the one-launch-per-iteration plan; on failure the per-colour path is used.
void build_plan(tc_sim *s) {
    s->fused_ready = false;
    s->fused_reason.clear();
    int dev = 0;
    int max_lds = 65536;
    if (hipGetDevice(&dev) == hipSuccess) {
        int v = 0;
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeSharedMemPerBlockOptin, dev) == hipSuccess && v > 0) max_lds = v;
    }
    const uint32_t max_vertices = uint32_t(max_lds / int(2 * sizeof(float4)));  // room for the damped variant
    std::string why = tc::build_tile_plan(s->mesh, s->tile_vertices, max_vertices, s->plan);
    if (!why.empty()) { s->fused_reason = why; return; }
    hipError_t e;
    if ((e = s->pl_ext_off.upload(s->plan.ext_off.data(), s->plan.ext_off.size(), s->stream)) != hipSuccess ||
        (e = s->pl_ext_vid.upload(s->plan.ext_vid.data(), s->plan.ext_vid.size(), s->stream)) != hipSuccess ||
        (e = s->pl_n_own.upload(s->plan.n_own.data(), s->plan.n_own.size(), s->stream)) != hipSuccess ||
        (e = s->pl_ph_off.upload(s->plan.ph_off.data(), s->plan.ph_off.size(), s->stream)) != hipSuccess ||
        (e = s->pl_items.upload(s->plan.items.data(), s->plan.items.size(), s->stream)) != hipSuccess ||
        (e = tc::set_tile_lds_limit(size_t(s->plan.max_ext) * 2 * sizeof(float4))) != hipSuccess ||
        (e = hipStreamSynchronize(s->stream)) != hipSuccess) {
        s->fused_reason = std::string("plan upload: ") + hipGetErrorString(e);
        return;
    }
    s->fused_ready = true;
}

void build_records(tc_sim *s) {
    const tc::Mesh &m = s->mesh;
    const size_t nd = m.d_v0.size(), nb = m.b_v0.size();
    s->drec_h.resize(nd);
    s->dcc_h.resize(nd);
    s->split_compliance = false;
    for (size_t k = 0; k < nd; ++k) {
        uint4 r;
        r.x = m.d_v0[k];
        r.y = m.d_v1[k] | (m.d_type[k] == 1u ? 0x80000000u : 0u);
        std::memcpy(&r.z, &m.d_rest[k], 4);
        std::memcpy(&r.w, &m.d_tc[k], 4);
        s->drec_h[k] = r;
        s->dcc_h[k] = m.d_cc[k];
        if (m.d_cc[k] != m.d_tc[k]) s->split_compliance = true;
    }
    s->brec_h.resize(nb);
    for (size_t k = 0; k < nb; ++k) {
        uint4 r;
        r.x = m.b_v2[k];
        r.y = m.b_v3[k];
        std::memcpy(&r.z, &m.b_rest[k], 4);
        std::memcpy(&r.w, &m.b_comp[k], 4);
        s->brec_h[k] = r;
    }
}

int upload_records(tc_sim *s) {
    TC_HIP(s->drec.upload(s->drec_h.data(), s->drec_h.size(), s->stream));
    TC_HIP(s->dcc.upload(s->dcc_h.data(), s->dcc_h.size(), s->stream));
    TC_HIP(s->brec.upload(s->brec_h.data(), s->brec_h.size(), s->stream));
    TC_HIP(hipStreamSynchronize(s->stream));
    build_plan(s);  // compliance classes are baked into the plan
    s->touch();
    return TC_OK;
}

int upload_colliders(tc_sim *s) {
    TC_HIP(s->cols.upload(s->cols_h.data(), s->cols_h.size(), s->stream));
    TC_HIP(hipStreamSynchronize(s->stream));
    s->touch();
    return TC_OK;
}

bool finite3(const float *p) { return std::isfinite(p[0]) && std::isfinite(p[1]) && std::isfinite(p[2]); }

}  // namespace

int tc_run_frame(tc_sim *s, float dt, uint32_t substeps) { return run_frame(s, dt, substeps); }

extern "C" {

void tc_sim_params_default(tc_sim_params *p) {
    if (!p) return;
    p->thickness = 0.005f;
    p->layer_id = 0;
    p->stiffness = 1000.0f;
    p->bending_stiffness = 10.0f;
    p->compression_stiffness = -1.0f;
    p->shear_stiffness = -1.0f;
    p->workgroup_size = 0;
    p->solver_mode = 0;
}

static int finish_create(tc_sim *s, const tc_sim_params &pp, tc_sim **out);

static tc::MeshInput mesh_input(const float *positions, uint32_t nv, const uint32_t *edges, uint32_t ne,
                                const uint32_t *faces, uint32_t nf, const float *inv_masses, const float *thicknesses,
                                const uint32_t *layer_ids, const tc_sim_params &pp) {
    tc::MeshInput in;
    in.positions = positions; in.nv = nv;
    in.edges = edges; in.ne = ne;
    in.faces = (faces && nf) ? faces : nullptr; in.nf = (faces && nf) ? nf : 0;
    in.inv_masses = inv_masses; in.thicknesses = thicknesses; in.layer_ids = layer_ids;
    in.layer_id = pp.layer_id; in.thickness = pp.thickness;
    in.tension_k = pp.stiffness;
    in.compression_k = pp.compression_stiffness >= 0.0f ? pp.compression_stiffness : pp.stiffness;
    in.shear_k = pp.shear_stiffness >= 0.0f ? pp.shear_stiffness : pp.stiffness * 0.5f;
    in.bending_k = pp.bending_stiffness;
    return in;
}

static int check_topology(const float *positions, uint32_t nv, const uint32_t *edges, uint32_t ne,
                          const uint32_t *faces, uint32_t nf) {
    if (nv > 0 && !positions) return fail(TC_ERR_VALUE, "positions is null");
    if (ne > 0 && !edges) return fail(TC_ERR_VALUE, "edges is null");
    if (nv >= 0x80000000u) return fail(TC_ERR_VALUE, "too many vertices (limit 2^31 - 1)");
    for (uint32_t e = 0; e < 2 * ne; ++e)
        if (edges[e] >= nv)
            return fail(TC_ERR_VALUE, "edges contain an index >= number of vertices (" + std::to_string(nv) + "): " +
                                          std::to_string(edges[e]));
    if (faces)
        for (uint32_t f = 0; f < 3 * nf; ++f)
            if (faces[f] >= nv)
                return fail(TC_ERR_VALUE, "faces contain an index >= number of vertices (" + std::to_string(nv) +
                                              "): " + std::to_string(faces[f]));
    return TC_OK;
}

void tc_sewing_params_default(tc_sewing_params *p) {
    if (!p) return;
    p->shrink_speed = 1.0f;
    p->stiffness = -1.0f;
    p->lock = -1;
    p->lock_distance = -1.0f;
}

// Sewing inputs of the mesh build (lib.rs:1643-1665 validation, mesh.rs:909-947 defaults).
static int sewing_input(uint32_t nv, const uint32_t *sew, uint32_t nsew, const tc_sewing_params *sp,
                        tc::MeshInput &in) {
    if (nsew && !sew) return fail(TC_ERR_VALUE, "sewing springs are null");
    for (size_t k = 0; k < size_t(nsew) * 2; k += 2)
        if (sew[k] >= nv || sew[k + 1] >= nv)
            return fail(TC_ERR_VALUE, "sewing_springs contain an index >= number of vertices (" + std::to_string(nv) +
                                          "): [" + std::to_string(sew[k]) + ", " + std::to_string(sew[k + 1]) + "]");
    tc_sewing_params def;
    tc_sewing_params_default(&def);
    const tc_sewing_params &q = sp ? *sp : def;
    in.sew = nsew ? sew : nullptr;
    in.nsew = nsew;
    in.sew_shrink_speed = q.shrink_speed;
    in.sew_stiffness = q.stiffness >= 0.0f ? q.stiffness : 10000.0f;
    in.sew_lock = q.lock != 0;
    return TC_OK;
}

int tc_sim_create_sewn(const float *positions, uint32_t nv, const uint32_t *edges, uint32_t ne, const uint32_t *faces,
                       uint32_t nf, const float *inv_masses, const float *thicknesses, const uint32_t *layer_ids,
                       const uint32_t *sewing, uint32_t num_sewing, const tc_sewing_params *sewing_params,
                       const tc_sim_params *params, tc_sim **out) {
    g_err.clear();
    if (!out) return fail(TC_ERR_VALUE, "out handle pointer is null");
    *out = nullptr;
    tc_sim_params def;
    tc_sim_params_default(&def);
    const tc_sim_params &pp = params ? *params : def;
    if (int rc = check_topology(positions, nv, edges, ne, faces, nf)) return rc;
    tc::MeshInput in = mesh_input(positions, nv, edges, ne, faces, nf, inv_masses, thicknesses, layer_ids, pp);
    if (int rc = sewing_input(nv, sewing, num_sewing, sewing_params, in)) return rc;

    tc_sim *s = new tc_sim();
    std::string err = tc::build_mesh(in, s->mesh);
    if (!err.empty()) { delete s; return fail(TC_ERR_VALUE, err); }
    if (sewing_params && sewing_params->lock_distance >= 0.0f)  // lib.rs:1749-1751
        s->sew_lock_dist = sewing_params->lock_distance;
    return finish_create(s, pp, out);
}

int tc_sim_create(const float *positions, uint32_t nv, const uint32_t *edges, uint32_t ne, const uint32_t *faces,
                  uint32_t nf, const float *inv_masses, const float *thicknesses, const uint32_t *layer_ids,
                  const tc_sim_params *params, tc_sim **out) {
    return tc_sim_create_sewn(positions, nv, edges, ne, faces, nf, inv_masses, thicknesses, layer_ids, nullptr, 0,
                              nullptr, params, out);
}

// Device state of a simulator whose mesh is built; takes ownership of s.
// Bucket count of the device hash tables.  The reference caps its tables at
// 65,536 buckets (crates/cloth_core/src/spatial_hash.rs:48); above that many
// elements most buckets hold several cells, and every neighbour-cell walk reads
// entries of cells that it then rejects.  Buckets list their elements in index
// order, so the candidates of a cell, and their order, do not depend on the
// bucket count: a larger table changes no result, only the number of rejected
// reads.  Four buckets per element, at most 2^22 (32 MiB for both tables).
static uint32_t device_table_size(const tc::Mesh &m) {
    const uint64_t n = std::max<uint64_t>(uint64_t(m.nv) + m.n_virtual, m.e_v0.size());
    uint32_t ts = std::max<uint32_t>(m.table_size, 1024u);
    while (ts < 4 * n && ts < (1u << 22)) ts <<= 1;
    return ts;
}

// Vertices in the Morton order of their rest positions (10 bits per axis over the bounding
// cube, ties by index).  The collider kernel takes its vertices in this order, so that the
// lanes of a wavefront lie close together on the cloth and need the same triangle clusters.
// It changes no result: each vertex's collision pass reads and writes only that vertex.
static std::vector<uint32_t> vertex_morton_order(const std::vector<float> &pos, uint32_t nv) {
    std::vector<uint32_t> idx(nv);
    for (uint32_t i = 0; i < nv; ++i) idx[i] = i;
    if (nv == 0) return idx;
    float lo[3] = {3.4e38f, 3.4e38f, 3.4e38f}, hi[3] = {-3.4e38f, -3.4e38f, -3.4e38f};
    for (uint32_t i = 0; i < nv; ++i)
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::min(lo[a], pos[3 * size_t(i) + a]);
            hi[a] = std::max(hi[a], pos[3 * size_t(i) + a]);
        }
    auto spread = [](uint32_t v) {
        v = (v * 0x00010001u) & 0xFF0000FFu;
        v = (v * 0x00000101u) & 0x0F00F00Fu;
        v = (v * 0x00000011u) & 0xC30C30C3u;
        v = (v * 0x00000005u) & 0x49249249u;
        return v;
    };
    // one scale for all axes: a flat cloth's jitter along its thin axis must not fill that
    // axis's 10 bits and scatter neighbours far apart in the order
    const float ext = std::max(std::max(hi[0] - lo[0], hi[1] - lo[1]), std::max(hi[2] - lo[2], 1e-9f));
    std::vector<uint32_t> code(nv);
    for (uint32_t i = 0; i < nv; ++i) {
        uint32_t c[3];
        for (int a = 0; a < 3; ++a) {
            const float u = std::min(std::max((pos[3 * size_t(i) + a] - lo[a]) / ext, 0.0f), 1.0f);
            c[a] = uint32_t(u * 1023.0f);
        }
        code[i] = (spread(c[0]) << 2) | (spread(c[1]) << 1) | spread(c[2]);
    }
    std::stable_sort(idx.begin(), idx.end(), [&](uint32_t a, uint32_t b) { return code[a] < code[b]; });
    return idx;
}

static int finish_create(tc_sim *s, const tc_sim_params &pp, tc_sim **out) {
    const uint32_t nv = s->mesh.nv;
    if (pp.workgroup_size == 64 || pp.workgroup_size == 128 || pp.workgroup_size == 256) s->block = pp.workgroup_size;
    if (const char *g = std::getenv("TCLOTH_USE_GRAPH")) s->use_graph = std::strcmp(g, "0") != 0;
    // Default tile size: small enough for one tile per CU or more.  An iteration's 14 colour
    // phases cost about 13-15 us of barriers and LDS round trips however little a tile holds, so
    // small meshes want many small tiles (measured at 128^2 and 256^2: 256-vertex tiles 14.3 and
    // 15.0 us per iteration, 1024-vertex tiles 19.2 and 19.6 us); from 512^2 on, 1024-vertex
    // tiles keep the cone halos small.  tc_sim_autotune times the alternatives per mesh.
    s->tile_vertices = 256;
    while (s->tile_vertices < 1024u && size_t(s->tile_vertices) * 2u * 256u <= nv) s->tile_vertices *= 2u;
    if (const char *tv = std::getenv("TCLOTH_TILE_VERTICES")) {
        const long v = std::strtol(tv, nullptr, 10);
        if (v >= 64 && v <= 16384) s->tile_vertices = uint32_t(v);
    }
    if (const char *fb = std::getenv("TCLOTH_TILE_BLOCK")) {
        const long v = std::strtol(fb, nullptr, 10);
        if (v == 256 || v == 512 || v == 1024) s->fused_block = uint32_t(v);
    }
    s->solver_mode = pp.solver_mode;
    s->pair_max_vt = s->pair_max_ee = uint32_t(std::min<uint64_t>(std::max<uint64_t>(uint64_t(nv) * 8u, 8192u), 262144u));
    s->pair_quota_vt = s->pair_quota_ee = nv ? std::min<uint32_t>(std::max<uint32_t>(s->pair_max_vt / nv, 1u), 8u) : 1u;
    s->cell_size = s->mesh.cell_size;
    s->table_size = device_table_size(s->mesh);

    auto bail = [&](int code) { delete s; return code; };
    hipError_t he = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
    if (he != hipSuccess) return bail(fail(TC_ERR_RUNTIME, std::string("hipStreamCreate: ") + hipGetErrorString(he)));

    const tc::Mesh &m = s->mesh;
    s->X0h.resize(nv);
    s->P0h.resize(nv);
    for (uint32_t i = 0; i < nv; ++i) {
        s->X0h[i] = make_float4(m.pos0[3 * i], m.pos0[3 * i + 1], m.pos0[3 * i + 2], m.thick[i]);
        s->P0h[i] = make_float4(m.pos0[3 * i], m.pos0[3 * i + 1], m.pos0[3 * i + 2], m.inv_mass[i]);
    }
    s->inv_orig.assign(m.inv_mass.begin(), m.inv_mass.end());
    s->inv_eff = s->inv_orig;
    std::vector<float4> zeros(nv, make_float4(0.0f, 0.0f, 0.0f, 0.0f));
    build_records(s);
    std::vector<uint2> star(m.star_v0.size()), edg(m.e_v0.size());
    for (size_t k = 0; k < star.size(); ++k) star[k] = make_uint2(m.star_v0[k], m.star_v1[k]);
    for (size_t k = 0; k < edg.size(); ++k) edg[k] = make_uint2(m.e_v0[k], m.e_v1[k]);
    const std::vector<uint32_t> col_order = vertex_morton_order(m.pos0, nv);
    const size_t T = s->table_size, NE = m.e_v0.size();
#define UP(buf, src, cnt)                                                                        \
    if ((he = (buf).upload((src), (cnt), s->stream)) != hipSuccess)                              \
        return bail(fail(TC_ERR_RUNTIME, std::string("device upload: ") + hipGetErrorString(he)));
#define AL(buf, cnt)                                                                             \
    if ((he = (buf).alloc(cnt)) != hipSuccess)                                                   \
        return bail(fail(TC_ERR_RUNTIME, std::string("hipMalloc: ") + hipGetErrorString(he)));
    UP(s->X, s->X0h.data(), nv);
    UP(s->P, s->P0h.data(), nv);
    UP(s->Palt, s->P0h.data(), nv);
    UP(s->V, zeros.data(), nv);
    UP(s->drec, s->drec_h.data(), s->drec_h.size());
    UP(s->dcc, s->dcc_h.data(), s->dcc_h.size());
    UP(s->brec, s->brec_h.data(), s->brec_h.size());
    AL(s->dlam, m.d_v0.size());
    AL(s->blam, m.b_v0.size());
    AL(s->dlam2, m.d_v0.size());
    AL(s->blam2, m.b_v0.size());
    UP(s->local_len, m.local_len.data(), m.local_len.size());
    UP(s->th_off, m.th_off.data(), m.th_off.size());
    UP(s->th_idx, m.th_idx.data(), m.th_idx.size());
    UP(s->th_rest, m.th_rest.data(), m.th_rest.size());
    UP(s->star_off, m.star_off.data(), m.star_off.size());
    UP(s->star, star.data(), star.size());
    UP(s->edges, edg.data(), edg.size());
    UP(s->edge_bbits, m.e_bbits.data(), m.e_bbits.size());
    if (m.has_multiple_layers) {  // the untangling pass's inputs
        UP(s->layer_d, m.layer.data(), m.layer.size());
        AL(s->normals, nv);
        UP(s->adj_off, m.adj_off.data(), m.adj_off.size());
        UP(s->adj_idx, m.adj_idx.data(), m.adj_idx.size());
    }
    UP(s->col_order, col_order.data(), col_order.size());
    AL(s->vkeys, nv); AL(s->vsorted, nv); AL(s->vcounts, T + (T + 1023) / 1024); AL(s->vstarts, T + 1);
    AL(s->ekeys, NE); AL(s->esorted, NE); AL(s->ecounts, T + (T + 1023) / 1024); AL(s->estarts, T + 1);
    AL(s->accum, nv);
    AL(s->jacobi_acc, nv);
    AL(s->readback, size_t(nv) * 3);
    AL(s->resid, 4);
    AL(s->resid_part, 3 * 1024);
#undef UP
#undef AL
    if ((he = hipMemsetAsync(s->accum.p, 0, sizeof(tc::Accum) * std::max<size_t>(nv, 1), s->stream)) != hipSuccess ||
        (he = hipMemsetAsync(s->jacobi_acc.p, 0, sizeof(int4) * std::max<size_t>(nv, 1), s->stream)) != hipSuccess ||
        (he = hipMemsetAsync(s->dlam.p, 0, sizeof(float) * std::max<size_t>(m.d_v0.size(), 1), s->stream)) != hipSuccess ||
        (he = hipMemsetAsync(s->blam.p, 0, sizeof(float) * std::max<size_t>(m.b_v0.size(), 1), s->stream)) != hipSuccess ||
        (he = hipMemsetAsync(s->dlam2.p, 0, sizeof(float) * std::max<size_t>(m.d_v0.size(), 1), s->stream)) != hipSuccess ||
        (he = hipMemsetAsync(s->blam2.p, 0, sizeof(float) * std::max<size_t>(m.b_v0.size(), 1), s->stream)) != hipSuccess ||
        (he = hipStreamSynchronize(s->stream)) != hipSuccess)
        return bail(fail(TC_ERR_RUNTIME, std::string("device init: ") + hipGetErrorString(he)));
    if (int rc = upload_sewing(s)) return bail(rc);
    if ((he = hipStreamSynchronize(s->stream)) != hipSuccess)
        return bail(fail(TC_ERR_RUNTIME, std::string("device init: ") + hipGetErrorString(he)));
    build_plan(s);
    frames_init(s);
    *out = s;
    return TC_OK;
}

int tc_sim_create_batch(const tc_instance *instances, uint32_t n, const tc_sim_params *params, tc_sim **out) {
    g_err.clear();
    if (!out) return fail(TC_ERR_VALUE, "out handle pointer is null");
    *out = nullptr;
    if (n > 0 && !instances) return fail(TC_ERR_VALUE, "instances is null");
    tc_sim_params def;
    tc_sim_params_default(&def);
    const tc_sim_params &pp = params ? *params : def;
    std::vector<tc::Mesh> parts(n);
    uint64_t total = 0;
    for (uint32_t k = 0; k < n; ++k) {
        const tc_instance &I = instances[k];
        if (int rc = check_topology(I.positions, I.num_vertices, I.edges, I.num_edges, I.faces, I.num_faces)) {
            const std::string why = g_err;
            return fail(rc, "instance " + std::to_string(k) + ": " + why);
        }
        total += I.num_vertices;
        if (total >= 0x80000000ull) return fail(TC_ERR_VALUE, "too many vertices in the batch (limit 2^31 - 1)");
        const tc::MeshInput in = mesh_input(I.positions, I.num_vertices, I.edges, I.num_edges, I.faces, I.num_faces,
                                            I.inv_masses, I.thicknesses, nullptr, pp);
        const std::string err = tc::build_mesh(in, parts[k]);
        if (!err.empty()) return fail(TC_ERR_VALUE, "instance " + std::to_string(k) + ": " + err);
    }
    tc_sim *s = new tc_sim();
    tc::merge_meshes(parts, s->mesh);
    s->self_collision = false;  // instances are independent; overlapping ones would collide
    return finish_create(s, pp, out);
}

int tc_sim_instance_count(const tc_sim *s, uint32_t *n) {
    if (!s || !n) return fail(TC_ERR_VALUE, "null argument");
    *n = s->mesh.part_off.empty() ? 1u : uint32_t(s->mesh.part_off.size() - 1);
    return TC_OK;
}

int tc_sim_instance_range(const tc_sim *s, uint32_t i, uint32_t *first_vertex, uint32_t *num_vertices) {
    if (!s) return fail(TC_ERR_VALUE, "null simulator");
    const auto &po = s->mesh.part_off;
    const uint32_t count = po.empty() ? 1u : uint32_t(po.size() - 1);
    if (i >= count) return fail(TC_ERR_VALUE, "instance " + std::to_string(i) + " out of range (" + std::to_string(count) + ")");
    const uint32_t a = po.empty() ? 0u : po[i], b = po.empty() ? s->mesh.nv : po[i + 1];
    if (first_vertex) *first_vertex = a;
    if (num_vertices) *num_vertices = b - a;
    return TC_OK;
}

void tc_sim_destroy(tc_sim *sim) {
    if (!sim) return;
    if (sim->stream) (void)hipStreamSynchronize(sim->stream);
    delete sim;
}

int tc_sim_step_single_substep(tc_sim *s, float dt_sub) {
    if (!s) return fail(TC_ERR_VALUE, "null simulator");
    if (s->mesh.nv == 0) return TC_OK;
    int r = run_frame(s, dt_sub, 1);
    if (r != TC_OK) return r;
    TC_HIP(hipStreamSynchronize(s->stream));
    return TC_OK;
}

int tc_sim_substep_parts(tc_sim *s, float dt_sub, int collide, int self_collide, int update) {
    if (!s) return fail(TC_ERR_VALUE, "null simulator");
    if (s->mesh.nv == 0) return TC_OK;
    s->rec_slot = -1;
    s->sub_first = s->sub_last = true;  // one substep: a frame of its own
    s->pair_dt_frame = dt_sub;
    TC_HIP(enqueue_substep(s, dt_sub, collide != 0, self_collide != 0, update != 0));
    return TC_OK;
}

// Times `its` solver iterations of the current schedule from a freshly predicted state.
static int time_iterations(tc_sim *s, float dt_sub, uint32_t its, bool fused, double *ms_per_iteration) {
    const tc::SubstepArgs a = s->args(dt_sub);
    const tc::Mesh &m = s->mesh;
    hipEvent_t e0, e1;
    TC_HIP(hipEventCreate(&e0));
    TC_HIP(hipEventCreate(&e1));
    hipError_t e = tc::launch_predict(s->stream, s->P.p, s->X.p, s->V.p, m.nv, a, s->block);
    uint32_t pc = 0;
    auto one = [&](uint32_t it) -> hipError_t {
        if (fused) {
            hipError_t r = tc::launch_tile_iteration(s->stream, s->Pbuf(pc), s->Pbuf(pc ^ 1u), s->X.p, s->plan_dev(),
                                                     s->dlam_buf(pc), s->dlam_buf(pc ^ 1u), s->blam_buf(pc),
                                                     s->blam_buf(pc ^ 1u), s->slam_buf(pc), s->slam_buf(pc ^ 1u),
                                                     s->fused_args(), a, it == 0, s->fused_block);
            pc ^= 1u;
            return r;
        }
        for (size_t c = 0; c < m.d_col.count.size(); ++c) {
            hipError_t r = tc::launch_distance(s->stream, s->P.p, s->X.p, s->drec.p, s->dcc.p, s->dlam.p,
                                               m.d_col.offset[c], m.d_col.count[c], a, it == 0, s->split_compliance,
                                               s->block);
            if (r != hipSuccess) return r;
        }
        if (!m.s_v0.empty()) {
            hipError_t r = enqueue_sewing(s, s->P.p, dt_sub, s->slam.p);
            if (r != hipSuccess) return r;
        }
        for (size_t c = 0; c < m.b_col.count.size(); ++c) {
            hipError_t r = tc::launch_bending(s->stream, s->P.p, s->X.p, s->brec.p, s->blam.p, m.b_col.offset[c],
                                              m.b_col.count[c], a, it == 0, s->block, s->bgate.p);
            if (r != hipSuccess) return r;
        }
        return hipSuccess;
    };
    // the sewing lock rewrites rest lengths: time on a copy of the records and put them back after
    DevBuf<uint4> srec_keep;
    const size_t ns = m.s_v0.size();
    if (ns && e == hipSuccess) e = srec_keep.alloc(ns);
    if (ns && e == hipSuccess) e = hipMemcpyAsync(srec_keep.p, s->srec.p, 16 * ns, hipMemcpyDeviceToDevice, s->stream);
    if (e == hipSuccess) e = one(0);  // warm-up
    if (e == hipSuccess) e = hipEventRecord(e0, s->stream);
    for (uint32_t it = 1; e == hipSuccess && it <= its; ++it) e = one(it);
    if (e == hipSuccess) e = hipEventRecord(e1, s->stream);
    if (ns && e == hipSuccess) e = hipMemcpyAsync(s->srec.p, srec_keep.p, 16 * ns, hipMemcpyDeviceToDevice, s->stream);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    float ms = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    TC_HIP(e);
    *ms_per_iteration = double(ms) / double(its);
    s->last_p = 0;
    s->last_lam = 0;
    return TC_OK;
}

static uint32_t pc_final(const tc_sim *s, bool fused) { return fused ? (s->iterations & 1u) : 0u; }

int tc_sim_time_kernels(tc_sim *s, float dt_sub, uint32_t reps, double *out) {
    if (!s || !out) return fail(TC_ERR_VALUE, "null argument");
    for (int k = 0; k < 12; ++k) out[k] = 0.0;
    if (s->mesh.nv == 0 || reps == 0) return TC_OK;
    const bool fused = s->fused_active() && s->iterations > 0;
    const tc::Mesh &m = s->mesh;
    const tc::SubstepArgs a = s->args(dt_sub);
    const uint32_t n = m.nv, B = s->block;
    const size_t ndc = m.d_col.count.size(), nbc = m.b_col.count.size();
    const size_t nev = 2 * (4 + size_t(s->iterations) * (ndc + nbc));
    std::vector<hipEvent_t> ev(nev, nullptr);
    for (auto &e : ev) TC_HIP(hipEventCreate(&e));
    struct Rel { size_t a, b; int slot; };
    std::vector<Rel> rel;
    auto run = [&]() -> hipError_t {
        size_t k = 0;
        auto mark = [&](int slot, auto &&launch) -> hipError_t {
            hipError_t e = hipEventRecord(ev[k], s->stream);
            if (e != hipSuccess) return e;
            if ((e = launch()) != hipSuccess) return e;
            if ((e = hipEventRecord(ev[k + 1], s->stream)) != hipSuccess) return e;
            rel.push_back({k, k + 1, slot});
            k += 2;
            return hipSuccess;
        };
        hipError_t e = mark(0, [&] { return tc::launch_predict(s->stream, s->P.p, s->X.p, s->V.p, n, a, B); });
        if (e != hipSuccess) return e;
        uint32_t pc = 0;
        if (fused) {
            // the iterations run back to back under one event pair: no event packets between them
            const tc::FusedArgs fa = s->fused_args();
            const tc::FusedPlanDev pd = s->plan_dev();
            if ((e = mark(9, [&] {
                     hipError_t r = hipSuccess;
                     for (uint32_t it = 0; it < s->iterations && r == hipSuccess; ++it) {
                         r = tc::launch_tile_iteration(s->stream, s->Pbuf(pc), s->Pbuf(pc ^ 1u), s->X.p, pd,
                                                       s->dlam_buf(pc), s->dlam_buf(pc ^ 1u), s->blam_buf(pc),
                                                       s->blam_buf(pc ^ 1u), s->slam_buf(pc), s->slam_buf(pc ^ 1u),
                                                       fa, a, it == 0, s->fused_block);
                         pc ^= 1u;
                     }
                     return r;
                 })) != hipSuccess) return e;
        }
        float4 *Pc = s->Pbuf(pc);
        for (uint32_t it = 0; !fused && it < s->iterations; ++it) {
            // one event pair around each iteration's distance colours and one around its bending colours
            if (ndc && (e = mark(1, [&] {
                            hipError_t r = hipSuccess;
                            for (size_t c = 0; c < ndc && r == hipSuccess; ++c)
                                r = tc::launch_distance(s->stream, s->P.p, s->X.p, s->drec.p, s->dcc.p, s->dlam.p,
                                                        m.d_col.offset[c], m.d_col.count[c], a, it == 0,
                                                        s->split_compliance, B);
                            return r;
                        })) != hipSuccess) return e;
            if (nbc && (e = mark(3, [&] {
                            hipError_t r = hipSuccess;
                            for (size_t c = 0; c < nbc && r == hipSuccess; ++c)
                                r = tc::launch_bending(s->stream, s->P.p, s->X.p, s->brec.p, s->blam.p,
                                                       m.b_col.offset[c], m.b_col.count[c], a, it == 0, B,
                                                       s->bgate.p);
                            return r;
                        })) != hipSuccess) return e;
        }
        if (!s->cols_h.empty() || !s->tris_h.empty() || s->sdf_nb > 0)
            if ((e = mark(5, [&] { return tc::launch_collision(s->stream, Pc, s->X.p, n, s->col_order.p, s->cols.p, uint32_t(s->cols_h.size()),
                                                               s->tris.p, s->tri_bounds.p, uint32_t(s->tris_h.size()), s->tri_cbox.p,
                                                               s->single_sided_recovery, dt_sub, B, s->sdf_args()); })) != hipSuccess) return e;
        if (s->self_collision)
            if ((e = mark(6, [&] {
                     const tc::SelfColArgs sa = s->sc_args();
                     const tc::SelfColTopo t = s->topo();
                     hipError_t r = sa.untangle ? tc::launch_compute_normals(s->stream, s->X.p, t, s->normals.p, n)
                                                : hipSuccess;
                     if (r != hipSuccess) return r;
                     if ((r = enqueue_hashes(s, Pc)) != hipSuccess) return r;
                     if ((r = tc::launch_sc_solve(s->stream, Pc, s->X.p, t, s->vhash(), s->accum.p, sa)) != hipSuccess) return r;
                     if ((r = tc::launch_sc_untangle(s->stream, Pc, s->X.p, t, s->accum.p, sa)) != hipSuccess) return r;
                     if ((r = tc::launch_sc_solve_ee(s->stream, Pc, s->X.p, t, s->ehash(), s->accum.p, sa)) != hipSuccess) return r;
                     return tc::launch_sc_apply(s->stream, Pc, s->local_len.p, s->accum.p, sa, B);
                 })) != hipSuccess) return e;
        return mark(7, [&] { return tc::launch_update_vel(s->stream, Pc, s->X.p, s->V.p, n, dt_sub, B); });
    };
    double acc[12] = {0};
    for (uint32_t r = 0; r < reps; ++r) {
        rel.clear();
        hipError_t e = run();
        if (e != hipSuccess) {
            for (auto &x : ev) (void)hipEventDestroy(x);
            return fail(TC_ERR_RUNTIME, std::string("timed substep: ") + hipGetErrorString(e));
        }
        TC_HIP(hipStreamSynchronize(s->stream));
        for (const Rel &q : rel) {
            float ms = 0.0f;
            TC_HIP(hipEventElapsedTime(&ms, ev[q.a], ev[q.b]));
            acc[q.slot] += ms;
        }
        float wall = 0.0f;
        TC_HIP(hipEventElapsedTime(&wall, ev[rel.front().a], ev[rel.back().b]));
        acc[8] += wall;
    }
    for (auto &x : ev) TC_HIP(hipEventDestroy(x));
    for (int k = 0; k < 12; ++k) out[k] = acc[k] / double(reps);
    out[2] = fused ? 0.0 : double(ndc) * s->iterations;
    out[4] = fused ? 0.0 : double(nbc) * s->iterations;
    out[10] = fused ? double(s->iterations) : 0.0;
    s->last_p = pc_final(s, fused);
    s->last_lam = s->last_p;
    return TC_OK;
}

int tc_sim_autotune(tc_sim *s, float dt_sub, double *out) {
    if (!s) return fail(TC_ERR_VALUE, "null simulator");
    double best_ms = 0.0;
    // the trial predictions update V in place: keep a copy and put it back afterwards
    DevBuf<float4> vsave;
    TC_HIP(vsave.alloc(s->mesh.nv));
    TC_HIP(hipMemcpyAsync(vsave.p, s->V.p, sizeof(float4) * std::max<size_t>(s->mesh.nv, 1), hipMemcpyDeviceToDevice,
                          s->stream));
    TC_HIP(hipStreamSynchronize(s->stream));
    int r = time_iterations(s, dt_sub, 16, false, &best_ms);
    if (r != TC_OK) return r;
    const double per_colour_ms = best_ms;
    uint32_t best_tile = 0, best_block = 0;
    const uint32_t tiles[] = {256, 512, 1024, 2048};
    const uint32_t blocks[] = {256, 512, 1024};
    const uint32_t keep_tile = s->tile_vertices, keep_block = s->fused_block;
    for (uint32_t tv : tiles) {
        s->tile_vertices = tv;
        build_plan(s);
        if (!s->fused_ready) continue;
        for (uint32_t bk : blocks) {
            s->fused_block = bk;
            double ms = 0.0;
            if ((r = time_iterations(s, dt_sub, 16, true, &ms)) != TC_OK) return r;
            if (ms < best_ms) { best_ms = ms; best_tile = tv; best_block = bk; }
        }
    }
    if (best_tile) {
        s->tile_vertices = best_tile;
        s->fused_block = best_block;
        build_plan(s);
        s->solver_path = 2;
    } else {
        s->tile_vertices = keep_tile;
        s->fused_block = keep_block;
        build_plan(s);
        s->solver_path = 1;
    }
    TC_HIP(hipMemcpyAsync(s->V.p, vsave.p, sizeof(float4) * std::max<size_t>(s->mesh.nv, 1), hipMemcpyDeviceToDevice,
                          s->stream));
    TC_HIP(hipStreamSynchronize(s->stream));
    s->touch();
    if (out) {
        out[0] = best_tile;
        out[1] = best_block;
        out[2] = best_ms;
        out[3] = per_colour_ms;
    }
    return TC_OK;
}

int tc_sim_set_profiling_enabled(tc_sim *s, int enable, int *enabled) {
    if (!s) return fail(TC_ERR_VALUE, "null simulator");
    s->prof.enabled = enable != 0;
    s->prof.frame_begin();
    s->prof.last.clear();
    if (enabled) *enabled = s->prof.enabled ? 1 : 0;
    return TC_OK;
}

int tc_sim_take_profile(tc_sim *s, size_t *count) {
    if (!s) return fail(TC_ERR_VALUE, "null simulator");
    if (s->prof.enabled && !s->prof.cur.empty()) {
        TC_HIP(hipStreamSynchronize(s->stream));
        TcProfiler &P = s->prof;
        std::vector<TcProfiler::Entry> out;
        out.reserve(P.cur.size());
        const hipEvent_t t0 = P.pool[P.cur.front().ev0];
        for (const TcProfiler::Scope &c : P.cur) {
            float ms = 0.0f, start = 0.0f;
            if (c.ev1 != c.ev0) TC_HIP(hipEventElapsedTime(&ms, P.pool[c.ev0], P.pool[c.ev1]));
            TC_HIP(hipEventElapsedTime(&start, t0, P.pool[c.ev0]));
            out.push_back({c.name, c.depth, start, ms});
        }
        P.last.swap(out);
        P.cur.clear();
    }
    if (!s->prof.enabled) s->prof.last.clear();
    if (count) *count = s->prof.last.size();
    return TC_OK;
}

int tc_sim_profile_entry(const tc_sim *s, size_t index, const char **name, float *ms, float *start_ms,
                         uint32_t *depth) {
    if (!s) return fail(TC_ERR_VALUE, "null simulator");
    if (index >= s->prof.last.size()) return fail(TC_ERR_VALUE, "profile entry index out of range");
    const TcProfiler::Entry &e = s->prof.last[index];
    if (name) *name = e.name.c_str();
    if (ms) *ms = e.ms;
    if (start_ms) *start_ms = e.start_ms;
    if (depth) *depth = e.depth;
    return TC_OK;
}

int tc_sim_save_profile_trace(tc_sim *s, const char *path) {
    if (!s) return fail(TC_ERR_VALUE, "null simulator");
    if (!path) return fail(TC_ERR_VALUE, "path is null");
    FILE *f = std::fopen(path, "w");
    if (!f) return fail(TC_ERR_IO, std::string("cannot open ") + path + ": " + std::strerror(errno));
    // Chrome trace event format: complete events ("ph": "X") in microseconds, nested by time.
    std::fputs("{\"traceEvents\": [", f);
    for (size_t i = 0; i < s->prof.last.size(); ++i) {
        const TcProfiler::Entry &e = s->prof.last[i];
        std::fprintf(f, "%s\n{\"name\": \"%s\", \"ph\": \"X\", \"pid\": 0, \"tid\": %u, \"ts\": %.3f, \"dur\": %.3f}",
                     i ? "," : "", e.name.c_str(), e.depth, double(e.start_ms) * 1e3, double(e.ms) * 1e3);
    }
    std::fputs("\n], \"displayTimeUnit\": \"ms\"}\n", f);
    const bool bad = std::ferror(f) != 0;
    if (std::fclose(f) != 0 || bad) return fail(TC_ERR_IO, std::string("cannot write ") + path);
    return TC_OK;
}

int tc_sim_synchronize(tc_sim *s) {
    if (!s) return fail(TC_ERR_VALUE, "null simulator");
    TC_HIP(hipStreamSynchronize(s->stream));
    return TC_OK;
}

int tc_sim_get_positions(tc_sim *s, float *out, size_t n_floats) {
    if (!s) return fail(TC_ERR_VALUE, "null simulator");
    const uint32_t n = s->mesh.nv;
    if (n == 0) return TC_OK;
    if (!out || n_floats < size_t(n) * 3) return fail(TC_ERR_VALUE, "output buffer must hold 3 * num_vertices floats");
    TC_HIP(tc::launch_extract_positions(s->stream, s->X.p, s->readback.p, n));
    TC_HIP(hipMemcpyAsync(out, s->readback.p, sizeof(float) * 3 * n, hipMemcpyDeviceToHost, s->stream));
    TC_HIP(hipStreamSynchronize(s->stream));
    return TC_OK;
}

int tc_sim_set_positions_velocities(tc_sim *s, const float *pos, const float *vel, uint32_t n) {
    if (!s) return fail(TC_ERR_VALUE, "null simulator");
    if (n != s->mesh.nv || !pos) return TC_OK;  // dispatch.rs:958-960: silently ignored
    std::vector<float4> X(n), P(n), V(n);
    for (uint32_t i = 0; i < n; ++i) {
        X[i] = make_float4(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2], s->X0h[i].w);
        P[i] = make_float4(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2], effective_inv_mass(s, i));
        V[i] = vel ? make_float4(vel[3 * i], vel[3 * i + 1], vel[3 * i + 2], 0.0f) : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
    TC_HIP(hipStreamSynchronize(s->stream));
    TC_HIP(hipMemcpy(s->X.p, X.data(), sizeof(float4) * n, hipMemcpyHostToDevice));
    TC_HIP(hipMemcpy(s->P.p, P.data(), sizeof(float4) * n, hipMemcpyHostToDevice));
    TC_HIP(hipMemcpy(s->V.p, V.data(), sizeof(float4) * n, hipMemcpyHostToDevice));
    for (uint32_t i = 0; i < n; ++i) s->inv_eff[i] = P[i].w;
    s->last_p = 0;
    return TC_OK;
}

int tc_sim_reset(tc_sim *s) {
    if (!s) return fail(TC_ERR_VALUE, "null simulator");
    const uint32_t n = s->mesh.nv;
    if (!n) return TC_OK;
    std::vector<float4> V(n, make_float4(0.0f, 0.0f, 0.0f, 0.0f));
    TC_HIP(hipStreamSynchronize(s->stream));
    TC_HIP(hipMemcpy(s->X.p, s->X0h.data(), sizeof(float4) * n, hipMemcpyHostToDevice));
    TC_HIP(hipMemcpy(s->P.p, s->P0h.data(), sizeof(float4) * n, hipMemcpyHostToDevice));
    TC_HIP(hipMemcpy(s->V.p, V.data(), sizeof(float4) * n, hipMemcpyHostToDevice));
    // dispatch.rs:1411-1422: the pins go, the initial inverse masses come back (the base inverse
    // masses of set_vertex_inv_masses are kept for later calls)
    s->pins.clear();
    if (int rc = upload_pins(s)) return rc;
    s->prio_latched = false;
    s->prio_frame = 0;
    s->prio_ramp_t = 0.0f;
    s->prio_ratio = 0.0f;
    s->touch();
    for (uint32_t i = 0; i < n; ++i) s->inv_eff[i] = s->P0h[i].w;
    s->last_p = 0;
    return ws_clear_history(s);
}

int tc_sim_add_sphere(tc_sim *s, const float c[3], float r, float fr, float rest) {
    if (!s || !c) return fail(TC_ERR_VALUE, "null argument");
    tc::Collider k{};
    k.type = 0; k.friction = fr; k.restitution = rest;
    k.ax = c[0]; k.ay = c[1]; k.az = c[2]; k.radius = r;
    s->cols_h.push_back(k);
    return upload_colliders(s);
}

int tc_sim_add_capsule(tc_sim *s, const float a[3], const float b[3], float r, float fr, float rest) {
    if (!s || !a || !b) return fail(TC_ERR_VALUE, "null argument");
    tc::Collider k{};
    k.type = 1; k.friction = fr; k.restitution = rest;
    k.ax = a[0]; k.ay = a[1]; k.az = a[2]; k.radius = r;
    k.bx = b[0]; k.by = b[1]; k.bz = b[2];
    s->cols_h.push_back(k);
    return upload_colliders(s);
}

int tc_sim_add_plane(tc_sim *s, const float p[3], const float n[3], float fr, float rest) {
    if (!s || !p || !n) return fail(TC_ERR_VALUE, "null argument");
    tc::Collider k{};
    k.type = 2; k.friction = fr; k.restitution = rest;
    k.ax = p[0]; k.ay = p[1]; k.az = p[2];
    k.bx = n[0]; k.by = n[1]; k.bz = n[2];
    s->cols_h.push_back(k);
    return upload_colliders(s);
}

int tc_sim_clear_colliders(tc_sim *s) {
    if (!s) return fail(TC_ERR_VALUE, "null simulator");
    s->cols_h.clear();
    s->tris_h.clear();
    s->touch();
    return upload_collider_edges(s);
}

int tc_sim_set_sdf_colliders(tc_sim *s, uint32_t w, uint32_t h, uint32_t d, const void *texels, size_t nbytes,
                             const float *infos, uint32_t nb) {
    if (!s) return fail(TC_ERR_VALUE, "null simulator");
    if (nb > 0 && !infos) return fail(TC_ERR_VALUE, "collider infos pointer is null");
    const size_t nvox = size_t(w) * h * d;
    if (nb > 0 && (w == 0 || h == 0 || d == 0)) return fail(TC_ERR_VALUE, "texture dimensions must be positive");
    if (nb > 0 && (!texels || nbytes < nvox * 4))
        return fail(TC_ERR_VALUE, "texture data holds " + std::to_string(nbytes) + " bytes, " +
                                      std::to_string(nvox * 4) + " needed (Rg16Float, 4 bytes per texel)");
    TC_HIP(hipStreamSynchronize(s->stream));
    s->sdf_nb = 0;
    s->sdf_info_h.clear();
    s->sdf_xf_h.clear();
    s->dyn.clear();  // a static atlas replaces the dynamic one
    if (nb > 0) {
        TC_HIP(s->sdf_tex.alloc(nvox));
        TC_HIP(hipMemcpy(s->sdf_tex.p, texels, nvox * 4, hipMemcpyHostToDevice));
        s->sdf_info_h.assign(infos, infos + size_t(nb) * 20);
        TC_HIP(s->sdf_info.alloc(size_t(nb) * 20));
        TC_HIP(hipMemcpy(s->sdf_info.p, infos, size_t(nb) * 20 * 4, hipMemcpyHostToDevice));
        // identity transforms until update_bone_transforms (simulation/mod.rs:1700-1719)
        std::vector<float> xf(size_t(nb) * 32, 0.0f);
        for (uint32_t b = 0; b < nb; ++b)
            for (int m = 0; m < 2; ++m)
                for (int k = 0; k < 4; ++k) xf[size_t(b) * 32 + m * 16 + k * 5] = 1.0f;
        TC_HIP(s->sdf_xf.alloc(xf.size()));
        TC_HIP(hipMemcpy(s->sdf_xf.p, xf.data(), 4 * xf.size(), hipMemcpyHostToDevice));
        s->sdf_xf_h = std::move(xf);
        s->sdf_w = w; s->sdf_h = h; s->sdf_d = d;
        s->sdf_nb = nb;
    }
    s->touch();
    return TC_OK;
}

int tc_sim_clear_bone_sdf_colliders(tc_sim *s) {
    if (!s) return fail(TC_ERR_VALUE, "null simulator");
    TC_HIP(hipStreamSynchronize(s->stream));
    s->sdf_nb = 0;
    s->sdf_info_h.clear();
    s->sdf_xf_h.clear();
    s->dyn.clear();
    s->touch();
    return TC_OK;
}

int tc_sim_update_bone_transforms(tc_sim *s, const float *world, const float *inv_world, uint32_t n) {
    if (!s) return fail(TC_ERR_VALUE, "null simulator");
    if (!world || !inv_world) return fail(TC_ERR_VALUE, "null matrix pointer");
    if (n != s->sdf_nb)
        return fail(TC_ERR_VALUE, "expected " + std::to_string(s->sdf_nb) + " transforms (one per SDF collider), got " +
                                      std::to_string(n));
    std::vector<float> xf(size_t(n) * 32);
    for (uint32_t b = 0; b < n; ++b) {
        std::memcpy(&xf[size_t(b) * 32], world + size_t(b) * 16, 64);
        std::memcpy(&xf[size_t(b) * 32 + 16], inv_world + size_t(b) * 16, 64);
    }
    // each collider's bounding sphere in world space goes into the infos' spare slots 3, 7, 11
    // (centre) and 15 (radius), where the resolve's prefilter reads it (simulation/mod.rs:1886-1925):
    // the local box's centre through the world matrix (the matrix as given, its rows taken as the
    // columns), radius = half the box diagonal x the largest column length + 0.1
    for (uint32_t b = 0; b < n; ++b) {
        float *I = &s->sdf_info_h[size_t(b) * 20];
        const float *m = world + size_t(b) * 16;
        const float c0 = (I[0] + I[4]) * 0.5f, c1 = (I[1] + I[5]) * 0.5f, c2 = (I[2] + I[6]) * 0.5f;
        const float ex = I[4] - I[0], ey = I[5] - I[1], ez = I[6] - I[2];
        const float r_local = 0.5f * std::sqrt(ex * ex + ey * ey + ez * ez);
        const float cw0 = m[0] * c0 + m[4] * c1 + m[8] * c2 + m[12];
        const float cw1 = m[1] * c0 + m[5] * c1 + m[9] * c2 + m[13];
        const float cw2 = m[2] * c0 + m[6] * c1 + m[10] * c2 + m[14];
        const float sx = std::sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
        const float sy = std::sqrt(m[4] * m[4] + m[5] * m[5] + m[6] * m[6]);
        const float sz = std::sqrt(m[8] * m[8] + m[9] * m[9] + m[10] * m[10]);
        const float max_scale = std::fmax(std::fmax(sx, sy), sz);  // f32::max
        I[3] = cw0; I[7] = cw1; I[11] = cw2;
        I[15] = r_local * max_scale + 0.1f;
    }
    TC_HIP(hipStreamSynchronize(s->stream));
    if (n) {
        TC_HIP(hipMemcpy(s->sdf_xf.p, xf.data(), 4 * xf.size(), hipMemcpyHostToDevice));
        TC_HIP(hipMemcpy(s->sdf_info.p, s->sdf_info_h.data(), 4 * s->sdf_info_h.size(), hipMemcpyHostToDevice));
    }
    s->sdf_xf_h = std::move(xf);
    return TC_OK;
}

int tc_sim_set_mesh_triangles(tc_sim *s, const float *tris, uint32_t nt, const float *attrs, float friction,
                              float thickness, float restitution, int single_sided) {
    if (!s) return fail(TC_ERR_VALUE, "null simulator");
    if (nt > 0 && !tris) return fail(TC_ERR_VALUE, "triangles pointer is null");
    std::vector<tc::MeshTri> T(nt);
    for (uint32_t t = 0; t < nt; ++t) {
        const float *v = tris + 9 * size_t(t);
        tc::MeshTri m{};
        m.p0x = v[0]; m.p0y = v[1]; m.p0z = v[2];
        m.p1x = v[3]; m.p1y = v[4]; m.p1z = v[5];
        m.p2x = v[6]; m.p2y = v[7]; m.p2z = v[8];
        m.friction = attrs ? attrs[4 * t] : friction;
        m.thickness = attrs ? attrs[4 * t + 1] : thickness;
        m.restitution = attrs ? attrs[4 * t + 2] : restitution;
        if (attrs) {  // `row[3].round() as u32` (lib.rs:2210): round half away from zero, saturate
            const float r = std::round(attrs[4 * t + 3]);
            m.flags = r > 0.0f ? (r >= 4294967296.0f ? UINT32_MAX : uint32_t(r)) : 0u;
        } else {
            m.flags = single_sided ? 1u : 0u;
        }
        T[t] = m;
    }
    // Morton order of the triangle centroids (simulation/mod.rs:2607-2632), stable
    std::vector<tc::MeshTri> sorted = T;
    if (nt > 0) {
        float bmin[3] = {3.4028235e38f, 3.4028235e38f, 3.4028235e38f};
        float bmax[3] = {-3.4028235e38f, -3.4028235e38f, -3.4028235e38f};
        std::vector<std::array<float, 3>> cen(nt);
        for (uint32_t t = 0; t < nt; ++t) {
            const tc::MeshTri &m = T[t];
            cen[t] = {(m.p0x + m.p1x + m.p2x) / 3.0f, (m.p0y + m.p1y + m.p2y) / 3.0f, (m.p0z + m.p1z + m.p2z) / 3.0f};
            for (int a = 0; a < 3; ++a) { bmin[a] = std::min(bmin[a], cen[t][a]); bmax[a] = std::max(bmax[a], cen[t][a]); }
        }
        float inv[3];
        for (int a = 0; a < 3; ++a) inv[a] = 1.0f / std::max(bmax[a] - bmin[a], 1e-5f);
        auto expand = [](uint32_t v) {
            v = (v * 0x00010001u) & 0xFF0000FFu;
            v = (v * 0x00000101u) & 0x0F00F00Fu;
            v = (v * 0x00000011u) & 0xC30C30C3u;
            v = (v * 0x00000005u) & 0x49249249u;
            return v;
        };
        auto q = [](float x) -> uint32_t { return x > 0.0f ? uint32_t(x) : 0u; };
        std::vector<uint32_t> code(nt), idx(nt);
        for (uint32_t t = 0; t < nt; ++t) {
            uint32_t c[3];
            for (int a = 0; a < 3; ++a) c[a] = q(std::min(std::max((cen[t][a] - bmin[a]) * inv[a], 0.0f), 1.0f) * 1023.0f);
            code[t] = (expand(c[0]) << 2) | (expand(c[1]) << 1) | expand(c[2]);
            idx[t] = t;
        }
        std::stable_sort(idx.begin(), idx.end(), [&](uint32_t a, uint32_t b) { return code[a] < code[b]; });
        for (uint32_t t = 0; t < nt; ++t) sorted[t] = T[idx[t]];
    }
    // bounding spheres (simulation/mod.rs:2656-2667)
    std::vector<float4> bounds(nt);
    for (uint32_t t = 0; t < nt; ++t) {
        const tc::MeshTri &m = sorted[t];
        const float cx = (m.p0x + m.p1x + m.p2x) / 3.0f, cy = (m.p0y + m.p1y + m.p2y) / 3.0f,
                    cz = (m.p0z + m.p1z + m.p2z) / 3.0f;
        auto d2 = [&](float x, float y, float z) { return (x - cx) * (x - cx) + (y - cy) * (y - cy) + (z - cz) * (z - cz); };
        const float r = std::sqrt(std::max(std::max(d2(m.p0x, m.p0y, m.p0z), d2(m.p1x, m.p1y, m.p1z)),
                                           d2(m.p2x, m.p2y, m.p2z))) + m.thickness;
        bounds[t] = make_float4(cx, cy, cz, r);
    }
    // culling boxes: each triangle's bounding sphere scaled by 1.5 (the kernel's per-triangle test),
    // widened by a relative 1e-6 and 1e-7 m against rounding, gathered into clusters of 32
    // triangles and super-clusters of 32 clusters
    const uint32_t ncl = (nt + 31u) / 32u, nsc = (ncl + 31u) / 32u;
    std::vector<float4> cbox(2 * (size_t(nsc) + ncl));
    auto grow = [](float4 &lo, float4 &hi, float x0, float y0, float z0, float x1, float y1, float z1) {
        lo.x = std::min(lo.x, x0); lo.y = std::min(lo.y, y0); lo.z = std::min(lo.z, z0);
        hi.x = std::max(hi.x, x1); hi.y = std::max(hi.y, y1); hi.z = std::max(hi.z, z1);
    };
    const float4 empty_lo = make_float4(3.4e38f, 3.4e38f, 3.4e38f, 0.0f), empty_hi = make_float4(-3.4e38f, -3.4e38f, -3.4e38f, 0.0f);
    for (uint32_t c = 0; c < nsc + ncl; ++c) { cbox[2 * c] = empty_lo; cbox[2 * c + 1] = empty_hi; }
    for (uint32_t t = 0; t < nt; ++t) {
        const float4 b = bounds[t];
        const float e = 1.5f * b.w;
        const float sx = 1e-6f * (std::fabs(b.x) + e) + 1e-7f, sy = 1e-6f * (std::fabs(b.y) + e) + 1e-7f,
                    sz = 1e-6f * (std::fabs(b.z) + e) + 1e-7f;
        const uint32_t c = t / 32u, sc = c / 32u;
        float4 &clo = cbox[2 * (size_t(nsc) + c)], &chi = cbox[2 * (size_t(nsc) + c) + 1];
        grow(clo, chi, b.x - e - sx, b.y - e - sy, b.z - e - sz, b.x + e + sx, b.y + e + sy, b.z + e + sz);
        float4 &slo = cbox[2 * size_t(sc)], &shi = cbox[2 * size_t(sc) + 1];
        grow(slo, shi, b.x - e - sx, b.y - e - sy, b.z - e - sz, b.x + e + sx, b.y + e + sy, b.z + e + sz);
    }
    s->tris_h = sorted;
    TC_HIP(s->tris.upload(s->tris_h.data(), nt, s->stream));
    TC_HIP(s->tri_bounds.upload(bounds.data(), nt, s->stream));
    TC_HIP(s->tri_cbox.upload(cbox.data(), cbox.size(), s->stream));
    TC_HIP(hipStreamSynchronize(s->stream));
    s->touch();
    return upload_collider_edges(s);
}

int tc_sim_set_gravity(tc_sim *s, float gx, float gy, float gz) {
    if (!s) return fail(TC_ERR_VALUE, "null simulator");
    s->gravity[0] = gx; s->gravity[1] = gy; s->gravity[2] = gz;
    s->touch();
    return TC_OK;
}

int tc_sim_set_damping(tc_sim *s, float d) {
    if (!s) return fail(TC_ERR_VALUE, "null simulator");
    s->damping = std::max(d, 0.0f);
    s->touch();
    return TC_OK;
}

int tc_sim_set_damping_all(tc_sim *s, float t, float c, float sh, float b) {
    if (!s) return fail(TC_ERR_VALUE, "null simulator");
    s->tension_damping = std::max(t, 0.0f);
    s->compression_damping = std::max(c, 0.0f);
    s->shear_damping = std::max(sh, 0.0f);
    s->bending_damping = std::max(b, 0.0f);
    s->touch();
    return TC_OK;
}

int tc_sim_set_solver_iterations(tc_sim *s, uint32_t it) {
    if (!s) return fail(TC_ERR_VALUE, "null simulator");
    s->iterations = it;
    return TC_OK;
}

int tc_sim_set_enable_self_collision(tc_sim *s, int on) {
    if (!s) return fail(TC_ERR_VALUE, "null simulator");
    s->self_collision = on != 0;
    s->touch();
    return TC_OK;
}

int tc_sim_set_self_collision_options(tc_sim *s, float relief, float ratio, int exclude, int untangle,
                                      uint32_t max_it) {
    if (!s) return fail(TC_ERR_VALUE, "null simulator");
    s->normal_untangling = untangle != 0;
    s->relief_factor = relief;
    s->max_disp_ratio = ratio;
    s->enable_relief = relief < 0.999f ? 1 : 0;  // simulation/mod.rs:941
    s->exclude_neighbors = exclude != 0;
    s->max_search_iterations = max_it;
    s->touch();
    return TC_OK;
}

int tc_sim_set_stiffness_all(tc_sim *s, float tension, float compression, float shear, float bending) {
    if (!s) return fail(TC_ERR_VALUE, "null simulator");
    const float tcmp = tc::compliance_of(tension), ccmp = tc::compliance_of(compression);
    const float scmp = tc::compliance_of(shear);
    const float bcmp = bending > 0.0f ? 1.0f / bending : 1e10f;
    tc::Mesh &m = s->mesh;
    for (size_t k = 0; k < m.d_v0.size(); ++k) {
        if (m.d_type[k] == 1u) { m.d_tc[k] = scmp; m.d_cc[k] = scmp; }
        else { m.d_tc[k] = tcmp; m.d_cc[k] = ccmp; }
    }
    for (size_t k = 0; k < m.b_v0.size(); ++k)
        m.b_comp[k] = m.b_pad0[k] > 1.5f ? std::max(bcmp * 10.0f, 0.5f) : bcmp;
    build_records(s);
    return upload_records(s);
}

int tc_sim_set_compliances_raw(tc_sim *s, float stretch, float shear, float bend) {
    if (!s) return fail(TC_ERR_VALUE, "null simulator");
    if (!(stretch >= 0.0f) || !(shear >= 0.0f) || !(bend >= 0.0f)) return fail(TC_ERR_VALUE, "compliances must be >= 0");
    tc::Mesh &m = s->mesh;
    for (size_t k = 0; k < m.d_v0.size(); ++k) {
        const float c = m.d_type[k] == 1u ? shear : stretch;
        m.d_tc[k] = c;
        m.d_cc[k] = c;
    }
    for (size_t k = 0; k < m.b_v0.size(); ++k) m.b_comp[k] = bend;
    build_records(s);
    return upload_records(s);
}

int tc_sim_set_solver_path(tc_sim *s, int path) {
    if (!s) return fail(TC_ERR_VALUE, "null simulator");
    if (path < 0 || path > 2) return fail(TC_ERR_VALUE, "solver path must be 0 (auto), 1 (per colour) or 2 (per iteration)");
    if (path == 2 && !s->fused_ready)
        return fail(TC_ERR_UNSUPPORTED, "one-launch-per-iteration path unavailable: " + s->fused_reason);
    s->solver_path = path;
    s->touch();
    return TC_OK;
}

int tc_sim_set_fast_projection(tc_sim *s, int enable) {
    if (!s) return fail(TC_ERR_VALUE, "null simulator");
    s->fast_projection = enable != 0;
    s->touch();  // captured graphs hold the kernel variant
    return TC_OK;
}

int tc_sim_get_fast_projection(const tc_sim *s, int *enabled) {
    if (!s || !enabled) return fail(TC_ERR_VALUE, "null argument");
    *enabled = s->fast_projection ? 1 : 0;
    return TC_OK;
}

int tc_sim_set_solver_mode(tc_sim *s, uint32_t mode) {
    if (!s) return fail(TC_ERR_VALUE, "null simulator");
    s->solver_mode = mode;  // 1: Jacobi; every other value colours (dispatch.rs:1436)
    s->touch();
    return TC_OK;
}

int tc_sim_get_solver_mode(const tc_sim *s, uint32_t *mode) {
    if (!s || !mode) return fail(TC_ERR_VALUE, "null argument");
    *mode = s->solver_mode;
    return TC_OK;
}

// Device buffers of the pair cache, allocated on first use: slots for 8 pairs per vertex (the
// largest quota), the counters, and the adjacency and island tables the collect kernel reads.
static int pair_cache_alloc(tc_sim *s) {
    if (s->pair_vt.p) return TC_OK;
    const tc::Mesh &m = s->mesh;
    const size_t slots = size_t(std::max<uint32_t>(m.nv, 1u)) * 8u;
    TC_HIP(hipStreamSynchronize(s->stream));
    TC_HIP(s->pair_vt.alloc(slots));
    TC_HIP(s->pair_ee.alloc(slots));
    TC_HIP(s->pair_counters.alloc(4));
    if (!s->adj_off.p) {
        TC_HIP(s->adj_off.upload(m.adj_off.data(), m.adj_off.size(), s->stream));
        TC_HIP(s->adj_idx.upload(m.adj_idx.data(), m.adj_idx.size(), s->stream));
    }
    TC_HIP(s->island.upload(m.island.data(), m.island.size(), s->stream));
    std::vector<uint4> none(slots, make_uint4(0xFFFFFFFFu, 0u, 0u, 0u));
    TC_HIP(hipMemcpyAsync(s->pair_vt.p, none.data(), sizeof(uint4) * slots, hipMemcpyHostToDevice, s->stream));
    TC_HIP(hipMemcpyAsync(s->pair_ee.p, none.data(), sizeof(uint4) * slots, hipMemcpyHostToDevice, s->stream));
    TC_HIP(hipMemsetAsync(s->pair_counters.p, 0, 16, s->stream));
    TC_HIP(hipStreamSynchronize(s->stream));
    return TC_OK;
}

int tc_sim_set_enable_pair_cache(tc_sim *s, int enable) {
    if (!s) return fail(TC_ERR_VALUE, "null simulator");
    if (enable)
        if (int rc = pair_cache_alloc(s)) return rc;
    s->pair_cache = enable != 0;
    s->touch();
    return TC_OK;
}

int tc_sim_get_enable_pair_cache(const tc_sim *s, int *enable) {
    if (!s || !enable) return fail(TC_ERR_VALUE, "null argument");
    *enable = s->pair_cache ? 1 : 0;
    return TC_OK;
}

int tc_sim_set_self_collision_algorithm(tc_sim *s, uint32_t algo) {
    if (!s) return fail(TC_ERR_VALUE, "null simulator");
    if (algo == 2)
        return fail(TC_ERR_UNSUPPORTED, "self_collision_algorithm=2 (virtual-vertex V-V) is not part of this build");
    if (algo == 1) {
        if (int rc = pair_cache_alloc(s)) return rc;
    }
    s->pair_cache = algo == 1;  // simulation/mod.rs:703-715
    s->sc_algorithm = algo;
    s->touch();
    return TC_OK;
}

int tc_sim_get_self_collision_algorithm(const tc_sim *s, uint32_t *algo) {
    if (!s || !algo) return fail(TC_ERR_VALUE, "null argument");
    *algo = s->sc_algorithm;
    return TC_OK;
}

int tc_sim_set_pair_cache_options(tc_sim *s, uint32_t max_vt, uint32_t max_ee, uint32_t margin_mode, float safety,
                                  float horizon_scale, float max_horizon) {
    if (!s) return fail(TC_ERR_VALUE, "null simulator");
    const uint32_t nv = s->mesh.nv;
    auto clampu = [](uint32_t v, uint32_t lo, uint32_t hi) { return std::min(std::max(v, lo), hi); };
    auto clampf = [](float v, float lo, float hi) { return std::min(std::max(v, lo), hi); };
    s->pair_max_vt = clampu(max_vt, 64u, 262144u);
    s->pair_max_ee = clampu(max_ee, 64u, 262144u);
    s->pair_quota_vt = nv ? clampu(s->pair_max_vt / nv, 1u, 8u) : 1u;
    s->pair_quota_ee = nv ? clampu(s->pair_max_ee / nv, 1u, 8u) : 1u;
    s->pair_margin_mode = margin_mode == 0 ? 0u : 1u;
    s->pair_safety = clampf(safety, 0.0f, 0.05f);
    s->pair_hscale = clampf(horizon_scale, 0.0f, 4.0f);
    s->pair_hmax = clampf(max_horizon, 0.0f, 0.1f);
    s->touch();
    return TC_OK;
}

int tc_sim_set_enable_pair_cache_final_fallback(tc_sim *s, int enable) {
    if (!s) return fail(TC_ERR_VALUE, "null simulator");
    s->pair_final_fallback = enable != 0;
    s->touch();
    return TC_OK;
}

int tc_sim_pair_cache_state(const tc_sim *s, double *out) {
    if (!s || !out) return fail(TC_ERR_VALUE, "null argument");
    out[0] = s->pair_max_vt; out[1] = s->pair_max_ee; out[2] = s->pair_quota_vt; out[3] = s->pair_quota_ee;
    out[4] = s->pair_margin_mode; out[5] = s->pair_safety; out[6] = s->pair_hscale; out[7] = s->pair_hmax;
    out[8] = s->pair_final_fallback ? 1.0 : 0.0;
    return TC_OK;
}

int tc_sim_get_active_pairs(tc_sim *s, uint32_t *counts, uint32_t *vt, size_t vt_cap, size_t *n_vt, uint32_t *ee,
                            size_t ee_cap, size_t *n_ee) {
    if (!s || !counts || !n_vt || !n_ee) return fail(TC_ERR_VALUE, "null argument");
    std::memset(counts, 0, 16);
    *n_vt = *n_ee = 0;
    if (!s->pair_vt.p) return TC_OK;
    const uint32_t nv = s->mesh.nv;
    const tc::PairArgs pa = s->pair_args();
    std::vector<uint4> hv(size_t(nv) * pa.quota_vt), he(size_t(nv) * pa.quota_ee);
    TC_HIP(hipStreamSynchronize(s->stream));
    TC_HIP(hipMemcpy(counts, s->pair_counters.p, 16, hipMemcpyDeviceToHost));
    if (!hv.empty()) TC_HIP(hipMemcpy(hv.data(), s->pair_vt.p, sizeof(uint4) * hv.size(), hipMemcpyDeviceToHost));
    if (!he.empty()) TC_HIP(hipMemcpy(he.data(), s->pair_ee.p, sizeof(uint4) * he.size(), hipMemcpyDeviceToHost));
    auto emit = [](const std::vector<uint4> &src, uint32_t *dst, size_t cap, size_t *n) {
        for (const uint4 &p : src) {
            if (p.x == 0xFFFFFFFFu) continue;
            if (dst && *n < cap) { dst[4 * *n] = p.x; dst[4 * *n + 1] = p.y; dst[4 * *n + 2] = p.z; dst[4 * *n + 3] = p.w; }
            ++*n;
        }
    };
    emit(hv, vt, vt_cap, n_vt);
    emit(he, ee, ee_cap, n_ee);
    return TC_OK;
}

int tc_sim_solver_info(tc_sim *s, double *out) {
    if (!s || !out) return fail(TC_ERR_VALUE, "null argument");
    out[0] = s->fused_active() ? 2.0 : 1.0;
    out[1] = s->plan.ntiles;
    out[2] = s->plan.max_ext;
    out[3] = s->plan.ext_vid.size();
    out[4] = double(s->plan.items.size() / 3);
    out[5] = double(s->plan.plain_items);
    out[6] = s->fused_ready ? 1.0 : 0.0;
    out[7] = s->plan.nphases;
    out[8] = s->tile_vertices;
    out[9] = s->fused_block;
    g_err = s->fused_ready ? std::string() : s->fused_reason;
    return TC_OK;
}

int tc_sim_set_hash_cell_size(tc_sim *s, float cell) {
    if (!s) return fail(TC_ERR_VALUE, "null simulator");
    if (!(cell > 0.0f) || !std::isfinite(cell)) return fail(TC_ERR_VALUE, "cell size must be positive and finite");
    s->cell_size = cell;
    s->touch();
    return TC_OK;
}

uint32_t tc_sim_num_vertices(const tc_sim *s) { return s ? s->mesh.nv : 0; }
uint32_t tc_sim_num_distance_constraints(const tc_sim *s) { return s ? uint32_t(s->mesh.d_v0.size()) : 0; }
uint32_t tc_sim_num_bending_constraints(const tc_sim *s) { return s ? uint32_t(s->mesh.b_v0.size()) : 0; }

int tc_sim_build_hashes(tc_sim *s) {
    if (!s) return fail(TC_ERR_VALUE, "null simulator");
    TC_HIP(enqueue_hashes(s, s->Pbuf(s->last_p)));
    TC_HIP(hipStreamSynchronize(s->stream));
    return TC_OK;
}

int tc_sim_get_array(tc_sim *s, int which, void *out, size_t cap, size_t *count) {
    if (!s) return fail(TC_ERR_VALUE, "null simulator");
    const tc::Mesh &m = s->mesh;
    auto host = [&](const void *src, size_t n, size_t elem) -> int {
        if (count) *count = n;
        if (!out) return TC_OK;
        if (cap < n) return fail(TC_ERR_VALUE, "capacity too small");
        if (n) std::memcpy(out, src, n * elem);
        return TC_OK;
    };
    auto dev = [&](const void *src, size_t n, size_t elem) -> int {
        if (count) *count = n;
        if (!out) return TC_OK;
        if (cap < n) return fail(TC_ERR_VALUE, "capacity too small");
        TC_HIP(hipStreamSynchronize(s->stream));
        if (n) TC_HIP(hipMemcpy(out, src, n * elem, hipMemcpyDeviceToHost));
        return TC_OK;
    };
    auto dev_xyz = [&](const float4 *src) -> int {
        const size_t n = size_t(m.nv) * 3;
        if (count) *count = n;
        if (!out) return TC_OK;
        if (cap < n) return fail(TC_ERR_VALUE, "capacity too small");
        std::vector<float4> tmp(m.nv);
        TC_HIP(hipStreamSynchronize(s->stream));
        if (m.nv) TC_HIP(hipMemcpy(tmp.data(), src, sizeof(float4) * m.nv, hipMemcpyDeviceToHost));
        float *o = static_cast<float *>(out);
        for (uint32_t i = 0; i < m.nv; ++i) { o[3 * i] = tmp[i].x; o[3 * i + 1] = tmp[i].y; o[3 * i + 2] = tmp[i].z; }
        return TC_OK;
    };
    switch (which) {
    case 0: return host(m.d_v0.data(), m.d_v0.size(), 4);
    case 1: return host(m.d_v1.data(), m.d_v1.size(), 4);
    case 2: return host(m.d_col.offset.data(), m.d_col.offset.size(), 4);
    case 3: return host(m.d_col.count.data(), m.d_col.count.size(), 4);
    case 4: return host(m.b_v2.data(), m.b_v2.size(), 4);
    case 5: return host(m.b_v3.data(), m.b_v3.size(), 4);
    case 6: return host(m.b_col.offset.data(), m.b_col.offset.size(), 4);
    case 7: return host(m.b_col.count.data(), m.b_col.count.size(), 4);
    case 8: return host(m.d_rest.data(), m.d_rest.size(), 4);
    case 9: return host(m.b_rest.data(), m.b_rest.size(), 4);
    case 10: return dev(s->vkeys.p, m.nv, 4);
    case 11: return dev(s->vstarts.p, size_t(s->table_size) + 1, 4);
    case 12: return dev(s->vsorted.p, m.nv, 4);
    case 13: return dev_xyz(s->Pbuf(s->last_p));
    case 14: return dev_xyz(s->V.p);
    case 15: return dev(s->dlam_buf(s->last_lam), m.d_v0.size(), 4);
    case 16: return dev(s->blam_buf(s->last_lam), m.b_v0.size(), 4);
    case 17: {
        uint32_t sc[4];
        std::memcpy(&sc[0], &s->cell_size, 4);
        sc[1] = s->table_size; sc[2] = m.n_virtual; sc[3] = uint32_t(m.e_v0.size());
        return host(sc, 4, 4);
    }
    case 18: {
        std::vector<uint32_t> e(2 * m.e_v0.size());
        for (size_t k = 0; k < m.e_v0.size(); ++k) { e[2 * k] = m.e_v0[k]; e[2 * k + 1] = m.e_v1[k]; }
        return host(e.data(), e.size(), 4);
    }
    case 19: return dev(s->ekeys.p, m.e_v0.size(), 4);
    case 20: return dev(s->estarts.p, size_t(s->table_size) + 1, 4);
    case 21: return dev(s->esorted.p, m.e_v0.size(), 4);
    case 22: {  // the inverse masses the solver reads (P.w)
        if (count) *count = m.nv;
        if (!out) return TC_OK;
        if (cap < m.nv) return fail(TC_ERR_VALUE, "capacity too small");
        std::vector<float4> tmp(m.nv);
        TC_HIP(hipStreamSynchronize(s->stream));
        if (m.nv) TC_HIP(hipMemcpy(tmp.data(), s->P.p, sizeof(float4) * m.nv, hipMemcpyDeviceToHost));
        float *o = static_cast<float *>(out);
        for (uint32_t i = 0; i < m.nv; ++i) o[i] = tmp[i].w;
        return TC_OK;
    }
    case 23: return dev(s->sdf_tex.p, s->sdf_nb ? size_t(s->sdf_w) * s->sdf_h * s->sdf_d : 0, 4);
    case 24: return host(s->sdf_info_h.data(), s->sdf_info_h.size(), 4);
    case 25: {  // the dynamic bone SDF's skinned positions (xyz)
        const size_t n = s->dyn.ready ? size_t(s->dyn.nv) : 0;
        if (count) *count = 3 * n;
        if (!out) return TC_OK;
        if (cap < 3 * n) return fail(TC_ERR_VALUE, "capacity too small");
        std::vector<float4> tmp(n);
        TC_HIP(hipStreamSynchronize(s->stream));
        if (n) TC_HIP(hipMemcpy(tmp.data(), s->dyn.skinned.p, sizeof(float4) * n, hipMemcpyDeviceToHost));
        float *o = static_cast<float *>(out);
        for (size_t i = 0; i < n; ++i) { o[3 * i] = tmp[i].x; o[3 * i + 1] = tmp[i].y; o[3 * i + 2] = tmp[i].z; }
        return TC_OK;
    }
    case 26: return dev(s->dyn.tris.p, s->dyn.ready ? size_t(s->dyn.nt) * 16 : 0, 4);
    default: return fail(TC_ERR_VALUE, "unknown array id " + std::to_string(which));
    }
}

int tc_sim_residual(tc_sim *s, float dt_sub, double *rel, double *max_rel) {
    if (!s) return fail(TC_ERR_VALUE, "null simulator");
    const tc::Mesh &m = s->mesh;
    double h[4] = {0.0, 0.0, 0.0, 0.0};
    TC_HIP(tc::launch_residual(s->stream, s->Pbuf(s->last_p), s->X.p, s->drec.p, s->dcc.p, s->dlam_buf(s->last_lam),
                               uint32_t(m.d_v0.size()), s->brec.p, s->blam_buf(s->last_lam), uint32_t(m.b_v0.size()),
                               s->args(dt_sub), s->split_compliance, s->resid_part.p, s->resid.p, nullptr, 0.0f));
    TC_HIP(hipMemcpyAsync(h, s->resid.p, sizeof(h), hipMemcpyDeviceToHost, s->stream));
    TC_HIP(hipStreamSynchronize(s->stream));
    if (rel) *rel = h[1] > 0.0 ? std::sqrt(h[0]) / std::sqrt(h[1]) : 0.0;
    if (max_rel) *max_rel = h[2];
    return TC_OK;
}

int tc_mesh_build_arrays(const float *pos, uint32_t nv, const uint32_t *edges, uint32_t ne, const uint32_t *faces,
                         uint32_t nf, const float *inv_masses, const float *thicknesses, const tc_sim_params *params,
                         int which, void *out, size_t cap, size_t *count) {
    return tc_mesh_build_arrays_sewn(pos, nv, edges, ne, faces, nf, inv_masses, thicknesses, nullptr, 0, nullptr,
                                     params, which, out, cap, count);
}

int tc_mesh_build_arrays_sewn(const float *pos, uint32_t nv, const uint32_t *edges, uint32_t ne, const uint32_t *faces,
                              uint32_t nf, const float *inv_masses, const float *thicknesses, const uint32_t *sewing,
                              uint32_t num_sewing, const tc_sewing_params *sewing_params, const tc_sim_params *params,
                              int which, void *out, size_t cap, size_t *count) {
    g_err.clear();
    if (nv > 0 && !pos) return fail(TC_ERR_VALUE, "positions is null");
    if (ne > 0 && !edges) return fail(TC_ERR_VALUE, "edges is null");
    for (size_t e = 0; e < size_t(ne) * 2; ++e)
        if (edges[e] >= nv) return fail(TC_ERR_VALUE, "edges contain an index >= number of vertices");
    if (faces)
        for (size_t f = 0; f < size_t(nf) * 3; ++f)
            if (faces[f] >= nv) return fail(TC_ERR_VALUE, "faces contain an index >= number of vertices");
    tc_sim_params def;
    tc_sim_params_default(&def);
    const tc_sim_params &pp = params ? *params : def;
    tc::MeshInput in;
    in.positions = pos; in.nv = nv; in.edges = edges; in.ne = ne;
    in.faces = (faces && nf) ? faces : nullptr; in.nf = (faces && nf) ? nf : 0;
    in.inv_masses = inv_masses; in.thicknesses = thicknesses;
    in.layer_id = pp.layer_id; in.thickness = pp.thickness;
    in.tension_k = pp.stiffness;
    in.compression_k = pp.compression_stiffness >= 0.0f ? pp.compression_stiffness : pp.stiffness;
    in.shear_k = pp.shear_stiffness >= 0.0f ? pp.shear_stiffness : pp.stiffness * 0.5f;
    in.bending_k = pp.bending_stiffness;
    if (int rc = sewing_input(nv, sewing, num_sewing, sewing_params, in)) return rc;
    tc::Mesh m;
    const std::string err = tc::build_mesh(in, m);
    if (!err.empty()) return fail(TC_ERR_VALUE, err);
    auto host = [&](const void *src, size_t n) -> int {
        if (count) *count = n;
        if (!out) return TC_OK;
        if (cap < n) return fail(TC_ERR_VALUE, "capacity too small");
        if (n) std::memcpy(out, src, n * 4);
        return TC_OK;
    };
    std::vector<uint32_t> tmp;
    switch (which) {
    case 0: return host(m.d_v0.data(), m.d_v0.size());
    case 1: return host(m.d_v1.data(), m.d_v1.size());
    case 2: return host(m.d_col.offset.data(), m.d_col.offset.size());
    case 3: return host(m.d_col.count.data(), m.d_col.count.size());
    case 4: return host(m.b_v2.data(), m.b_v2.size());
    case 5: return host(m.b_v3.data(), m.b_v3.size());
    case 6: return host(m.b_col.offset.data(), m.b_col.offset.size());
    case 7: return host(m.b_col.count.data(), m.b_col.count.size());
    case 8: return host(m.d_rest.data(), m.d_rest.size());
    case 9: return host(m.b_rest.data(), m.b_rest.size());
    case 17: {
        uint32_t sc[4];
        std::memcpy(&sc[0], &m.cell_size, 4);
        sc[1] = m.table_size; sc[2] = m.n_virtual; sc[3] = uint32_t(m.e_v0.size());
        return host(sc, 4);
    }
    case 18:
        for (size_t k = 0; k < m.e_v0.size(); ++k) { tmp.push_back(m.e_v0[k]); tmp.push_back(m.e_v1[k]); }
        return host(tmp.data(), tmp.size());
    case 22: return host(m.th_off.data(), m.th_off.size());
    case 23: return host(m.th_idx.data(), m.th_idx.size());
    case 24: return host(m.th_rest.data(), m.th_rest.size());
    case 25: return host(m.local_len.data(), m.local_len.size());
    case 26: return host(m.star_off.data(), m.star_off.size());
    case 27:
        for (size_t k = 0; k < m.star_v0.size(); ++k) { tmp.push_back(m.star_v0[k]); tmp.push_back(m.star_v1[k]); }
        return host(tmp.data(), tmp.size());
    case 28: return host(m.e_bbits.data(), m.e_bbits.size());
    case 29: return host(m.island.data(), m.island.size());
    case 30: return host(m.edge_to_constraint.data(), m.edge_to_constraint.size());
    case 31: return host(m.d_type.data(), m.d_type.size());
    case 32: return host(m.d_tc.data(), m.d_tc.size());
    case 33: return host(m.d_cc.data(), m.d_cc.size());
    case 34: return host(m.b_comp.data(), m.b_comp.size());
    case 35: return host(m.adj_off.data(), m.adj_off.size());
    case 36: return host(m.adj_idx.data(), m.adj_idx.size());
    case 37: return host(m.layer.data(), m.layer.size());
    case 38: return host(m.b_v0.data(), m.b_v0.size());
    case 39: return host(m.b_v1.data(), m.b_v1.size());
    case 40: return host(m.b_pad0.data(), m.b_pad0.size());
    case 41: return host(m.b_pad1.data(), m.b_pad1.size());
    case 42: return host(m.s_v0.data(), m.s_v0.size());
    case 43: return host(m.s_v1.data(), m.s_v1.size());
    case 44: return host(m.s_rest.data(), m.s_rest.size());
    case 45: return host(m.s_col.offset.data(), m.s_col.offset.size());
    case 46: return host(m.s_col.count.data(), m.s_col.count.size());
    default: return fail(TC_ERR_VALUE, "unknown array id " + std::to_string(which));
    }
}

int tc_color_edge_pairs(uint32_t nv, const uint32_t *pairs, uint32_t np, uint32_t *colors_out) {
    if (np && (!pairs || !colors_out)) return fail(TC_ERR_VALUE, "null argument");
    const std::vector<uint32_t> c = tc::color_edge_pairs(nv, pairs, np);
    if (np) std::memcpy(colors_out, c.data(), sizeof(uint32_t) * np);
    return TC_OK;
}

int tc_compute_inv_masses(const float *pos, uint32_t nv, const uint32_t *faces, uint32_t nf, const float *pins,
                          float density, float *out) {
    if (nv && (!pos || !out)) return fail(TC_ERR_VALUE, "null argument");
    if (faces)
        for (size_t k = 0; k < size_t(nf) * 3; ++k)
            if (faces[k] >= nv) return fail(TC_ERR_VALUE, "faces contain an index >= number of vertices");
    tc::areal_inv_masses(pos, nv, faces, faces ? nf : 0, density, pins, out);
    return TC_OK;
}

const char *tc_last_error(void) { return g_err.c_str(); }
const char *tc_version(void) { return "taremincloth_amd 0.1 (gfx950)"; }

}  // extern "C"
