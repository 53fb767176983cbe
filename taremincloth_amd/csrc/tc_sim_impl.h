// Internal definition of the simulator object shared by the translation units
// of the C ABI (tc_sim.hip: simulator, tc_ws_api.hip: learned warm start).
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "../../include/tc_sim.h"
#include "tc_kernels.h"
#include "tc_mesh.h"
#include "tc_sdf.h"
#include "tc_tiles.h"
#include "tc_ws.h"

namespace tc {

extern thread_local std::string g_err;

inline int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}

#define TC_HIP(call)                                                                              \
    do {                                                                                          \
        hipError_t e_ = (call);                                                                   \
        if (e_ != hipSuccess)                                                                     \
            return fail(TC_ERR_RUNTIME, std::string(#call) + ": " + hipGetErrorString(e_));       \
    } while (0)

template <typename T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    hipError_t alloc(size_t count) {
        release();
        n = count;
        return hipMalloc(reinterpret_cast<void **>(&p), sizeof(T) * std::max<size_t>(count, 1));
    }
    hipError_t upload(const T *src, size_t count, hipStream_t s) {
        hipError_t e = alloc(count);
        if (e != hipSuccess || !count) return e;
        return hipMemcpyAsync(p, src, sizeof(T) * count, hipMemcpyHostToDevice, s);
    }
};

}  // namespace tc

using tc::DevBuf;
using tc::fail;

// Weights, optimiser state and training scratch of one warm-start network.
struct tc_ws_net {
    tc::ws::NetLayout nl;
    DevBuf<float> theta, m, v, grad;
    DevBuf<tc::ws::bf16> w, wt;
    uint32_t step = 0;
    hipStream_t stream = nullptr;
    // training scratch for up to rows_cap rows
    uint32_t rows_cap = 0;
    DevBuf<tc::ws::bf16> h[5], x_r[4], a_r[4], a_enc, a_dec, dyb, dzb, feat;
    DevBuf<float> y, dy, dh0, dh1, dx, partial;
    DevBuf<double> loss_part;
    DevBuf<float4> tgt;
    DevBuf<uint32_t> goff, gnbr;
    ~tc_ws_net() {
        if (stream) (void)hipStreamDestroy(stream);
    }
    tc::ws::NetDev dev() const { return {theta.p, w.p, wt.p}; }
};

// Pass timing with HIP events under the reference's pass names (simulation/profile.rs): a
// scope's start and end events are recorded on the simulator's stream around the pass, with its
// nesting depth; take() waits for the stream and turns the frame's events into milliseconds.
struct TcProfiler {
    struct Scope {
        const char *name;
        uint32_t depth;
        size_t ev0, ev1;
    };
    struct Entry {
        std::string name;
        uint32_t depth;
        float start_ms, ms;
    };
    bool enabled = false;
    bool active = false;           // inside a frame of run_frame (other entry points record nothing)
    std::vector<hipEvent_t> pool;  // events of the frame being recorded, reused frame after frame
    size_t used = 0;
    uint32_t depth = 0;
    std::vector<Scope> cur;        // scopes of the last frame enqueued
    std::vector<Entry> last;       // the last frame taken
    ~TcProfiler() {
        for (hipEvent_t e : pool) (void)hipEventDestroy(e);
    }
    void frame_begin() {
        used = 0;
        depth = 0;
        cur.clear();
    }
    // Records a start event; returns the scope index (or -1 when disabled / out of events).
    int begin(hipStream_t s, const char *name) {
        if (!enabled || !active) return -1;
        const size_t e = next();
        if (e == SIZE_MAX || hipEventRecord(pool[e], s) != hipSuccess) return -1;
        cur.push_back({name, depth++, e, e});
        return int(cur.size() - 1);
    }
    void end(hipStream_t s, int idx) {
        if (!enabled || !active || idx < 0) return;
        const size_t e = next();
        if (depth) --depth;
        if (e == SIZE_MAX || hipEventRecord(pool[e], s) != hipSuccess) return;
        cur[size_t(idx)].ev1 = e;
    }
    size_t next() {
        if (used == pool.size()) {
            hipEvent_t e;
            if (hipEventCreate(&e) != hipSuccess) return SIZE_MAX;
            pool.push_back(e);
        }
        return used++;
    }
};

struct tc_sim {
    tc::Mesh mesh;
    hipStream_t stream = nullptr;
    uint32_t block = 256;

    // vertex streams (P has a second buffer for the one-launch-per-iteration path)
    DevBuf<float4> X, P, V, Palt;
    std::vector<float4> X0h, P0h;  // initial state for reset()

    // dynamic pins (vertex -> target, weight) and inverse masses: inv_orig is the base inverse mass
    // of each vertex (set_vertex_inv_masses changes it), inv_eff what P.w holds (orig * (1 - w) when
    // pinned); the device table lists the pins in ascending vertex order
    struct PinRec { float target[3]; float weight; };
    std::map<uint32_t, PinRec> pins;
    std::vector<float> inv_orig, inv_eff;
    DevBuf<uint32_t> pin_v;
    DevBuf<float4> pin_t;  // target xyz, weight
    uint32_t npins = 0, pin_cap = 0;
    DevBuf<uint32_t> scat_idx;
    DevBuf<float> scat_val;

    // constraints
    DevBuf<uint4> drec, brec;
    DevBuf<float> dcc, dlam, blam, dlam2, blam2;
    DevBuf<uint4> bgate;               // seam-gated bending: {v0, v1, gate distance bits, 1} per constraint (meshes with seams)
    // sewing (mesh.rs:51-60, sewing.wgsl, sewing_shrink.wgsl): {v0, v1, current rest bits, target rest bits}
    DevBuf<uint4> srec;
    DevBuf<float> slam, slam2;         // sewing lambdas; the second buffer for the one-launch-per-iteration path
    float sew_stiffness = 10000.0f;    // simulator fields (mod.rs:637-640); the constructor's stiffness and
    bool sew_lock = true;              // lock flag are kept only in the records, which the solver does not read
    float sew_lock_dist = 0.02f;
    float sew_speed = 1.0f;
    // sewing priority (mod.rs:1262-1397): a one-way latch, advanced on the host from positions the
    // caller holds, that keeps gravity and the wrinkle field's blend at zero until enough sewing pairs
    // have closed (or max_frames have passed), then ramps them back over ramp_frames by smoothstep
    bool prio_on = false;
    float prio_threshold = 0.9f, prio_merge = 0.005f;
    uint32_t prio_ramp = 3, prio_max = 600;
    bool prio_latched = false;
    uint32_t prio_frame = 0;
    float prio_ramp_t = 0.0f, prio_ratio = 0.0f;
    float priority_scale() const {  // mod.rs:1364-1376
        if (!prio_on || mesh.s_v0.empty()) return 1.0f;
        if (!prio_latched) return 0.0f;
        if (prio_ramp == 0) return 1.0f;
        const float x = std::min(std::max(prio_ramp_t, 0.0f), 1.0f);
        return x * x * (3.0f - 2.0f * x);
    }
    bool split_compliance = false;
    std::vector<uint4> drec_h, brec_h;
    std::vector<float> dcc_h;

    // colliders
    std::vector<tc::Collider> cols_h;
    DevBuf<tc::Collider> cols;
    std::vector<tc::MeshTri> tris_h;
    DevBuf<tc::MeshTri> tris;
    DevBuf<float4> tri_bounds;
    DevBuf<float4> tri_cbox;           // super-cluster then cluster boxes (lo, hi) of the Morton-ordered triangles
    // wrinkle field (tc_wrinkle.hip): parameters and the RGBA8 texture
    bool wrinkle_on = false;
    tc::WrinkleArgs wrinkle;
    DevBuf<uchar4> wrinkle_tex;
    uint32_t wrinkle_w = 0, wrinkle_h = 0;
    // collider-edge collision (tc_edge_collision.hip): the triangles' unique edges and their grid
    bool edge_collision = false;
    float edge_margin_scale = 1.0f, edge_margin_offset = 0.0f;
    DevBuf<tc::ColliderEdge> cedges;
    DevBuf<uint32_t> cedge_starts, cedge_sorted;
    uint32_t ncedges = 0;
    DevBuf<uint32_t> col_order;        // vertices in Morton order of their rest positions (collider lanes)
    int single_sided_recovery = 1;

    // SDF colliders (set_bone_sdf_colliders); host copies of the infos and of the (world, inverse
    // world) transforms, 32 floats per collider, for update_bone_transforms and the dynamic bake
    DevBuf<uint32_t> sdf_tex;
    DevBuf<float> sdf_info, sdf_xf;
    std::vector<float> sdf_info_h, sdf_xf_h;
    uint32_t sdf_w = 0, sdf_h = 0, sdf_d = 0, sdf_nb = 0;
    // dynamic bone SDF (setup_dynamic_bone_sdf, simulation/mod.rs:1998-2312): every update_interval-th
    // frame the rest mesh is skinned, its triangles moved into their bones' frames and the bones'
    // atlas tiles re-baked straight into sdf_tex (tc_dyn_sdf.hip)
    struct DynSdf {
        bool ready = false;     // dynamic_sdf_setup is Some
        bool enabled = false;   // enable_dynamic_bone_sdf
        bool baked = false;     // dynamic_sdf_initial_baked
        uint32_t interval = 1, counter = 0;
        uint32_t res = 0, nbones = 0, nv = 0, nt = 0;
        std::vector<float> bind_inv;                   // 16 per bone, row-major
        std::vector<tc::bone::Params> params;          // bake_params, one per bone
        std::vector<tc::bone::Params> params_dev;      // what the device's parameter buffer holds
        bool dirty_set = false;                        // active_dirty_bone_indices is Some
        std::vector<uint32_t> dirty;
        DevBuf<float4> rest;                           // 4 per vertex: position, normal, indices (bits), weights
        DevBuf<float4> skinned;                        // skinned_positions
        DevBuf<uint4> src;                             // 2 per triangle: (i0, i1, i2, bone), (w0, w1, w2, mask) bits
        DevBuf<tc::bone::BakeTri> tris;                // bone-space bake triangles
        DevBuf<tc::bone::Params> dparams;
        DevBuf<float> mats;                            // skin matrices, then inverse world matrices (16 per bone)
        float *mats_host = nullptr;                    // pinned staging of `mats`
        hipEvent_t mats_ev = nullptr;                  // the last upload from mats_host
        uint32_t bakes = 0, skips = 0;                 // counters: tile bakes launched, frames skipped
        void clear() {
            ready = enabled = baked = dirty_set = false;
            counter = 0;
            dirty.clear();
        }
    } dyn;
    tc::SdfArgs sdf_args() const {
        tc::SdfArgs a;
        a.tex = sdf_tex.p; a.w = sdf_w; a.h = sdf_h; a.d = sdf_d;
        a.info = sdf_info.p; a.xf = sdf_xf.p; a.nb = sdf_nb;
        return a;
    }

    // self collision
    DevBuf<float> local_len, th_rest;
    DevBuf<uint32_t> th_off, th_idx, star_off, edge_bbits;
    DevBuf<uint2> star, edges;
    DevBuf<uint32_t> vkeys, vcounts, vstarts, vsorted;
    DevBuf<uint32_t> ekeys, ecounts, estarts, esorted;
    DevBuf<tc::Accum> accum;
    DevBuf<float> readback;
    DevBuf<double> resid, resid_part;

    // parameters (simulation/mod.rs:607-691 defaults)
    float gravity[3] = {0.0f, 0.0f, -9.81f};
    float damping = 1.0f;
    float tension_damping = 0.0f, compression_damping = 0.0f, shear_damping = 0.0f, bending_damping = 0.0f;
    uint32_t iterations = 2;
    bool self_collision = true;
    float relief_factor = 0.2f, max_disp_ratio = 0.2f;
    int enable_relief = 1, exclude_neighbors = 1;
    bool normal_untangling = true;  // only effective for multi-layer meshes (pipelines.rs:895-896)
    uint32_t max_search_iterations = 256;
    float cell_size = 0.01f;
    uint32_t table_size = 1024;

    // one-launch-per-iteration plan (tc_tiles.cpp)
    int solver_path = 0;          // 0 auto, 1 one launch per colour, 2 one launch per iteration
    uint32_t solver_mode = 0;     // 1: distance constraints by fixed-point Jacobi sums (tc_jacobi.hip)

    // pair cache (self_collision_algorithm 1, simulation/mod.rs:583-597, 703-870): the first
    // substep of a frame collects per-vertex top-K candidate pairs, the others solve only those;
    // with the final fallback the last substep takes the direct path on fresh hash tables
    uint32_t sc_algorithm = 0;
    bool pair_cache = false, pair_final_fallback = true;
    uint32_t pair_max_vt = 8192, pair_max_ee = 8192, pair_quota_vt = 1, pair_quota_ee = 1;
    uint32_t pair_margin_mode = 1;
    float pair_safety = 0.005f, pair_hscale = 1.3f, pair_hmax = 0.02f;
    float pair_dt_frame = 1.0f / 60.0f;  // frame length of the step being recorded (velocity horizon)
    bool sub_first = true, sub_last = true;  // position of the substep being recorded in its frame
    DevBuf<uint4> pair_vt, pair_ee;     // nv * 8 slots each (quota slots in use per vertex)
    DevBuf<uint32_t> pair_counters;     // 4: V-T / E-E evaluated, V-T / E-E dropped
    DevBuf<uint32_t> adj_off, adj_idx, island;
    DevBuf<uint32_t> layer_d;          // multi-layer cloth: layer ids, and the vertex normals of the
    DevBuf<float4> normals;            // substep's start positions (compute_normals.wgsl)
    DevBuf<int4> jacobi_acc;      // per-vertex sums of solver_mode 1, zero between iterations
    bool fused_ready = false;     // plan built and uploaded
    std::string fused_reason;     // why the plan is unavailable
    tc::TilePlan plan;
    DevBuf<uint32_t> pl_ext_off, pl_ext_vid, pl_n_own, pl_ph_off, pl_items;
    uint32_t fused_block = 1024;
    bool fast_projection = false;  // the tile path's refined-reciprocal projection (opt-in)
    uint32_t tile_vertices = 1024;  // set per mesh at creation; TCLOTH_TILE_VERTICES overrides; tc_sim_autotune picks per mesh
    uint32_t last_p = 0, last_lam = 0;  // buffers holding P and lambda after the last substep

    // learned warm start (tc_ws.h, tc_ws_api.hip)
    int ws_mode = 0;                   // 0 off, 1 previous correction only, 2 network
    tc_ws_net *ws_net = nullptr;
    uint32_t ws_iters = 8, ws_fallback = 56;
    float ws_threshold = 1e-4f;
    float ws_beta = 0.5f;              // weight of the previous correction in the warm start
    bool ws_ready = false;             // incidence arrays and buffers on the device
    DevBuf<uint32_t> ws_doff, ws_dnbr, ws_boff, ws_bnbr, ws_goff, ws_gnbr;
    DevBuf<float> ws_drest, ws_brest;
    std::vector<uint32_t> ws_goff_h, ws_gnbr_h;
    uint32_t ws_gmaxdeg = 0;
    float ws_lscale = 0.01f;
    DevBuf<float4> ws_ppred, ws_D;
    DevBuf<tc::ws::bf16> ws_feat, ws_h0, ws_h1;
    DevBuf<float> ws_y;
    DevBuf<uint32_t> ws_gate;          // [fallback flag, substeps checked, fallbacks, fallback chunks run]
    DevBuf<double> ws_res;             // residual at the gate: sum h^2, sum L0^2, max, relative
    // training samples recorded from plain substeps
    uint32_t rec_cap = 0, rec_stride = 1, rec_count = 0, rec_tick = 0;
    int rec_slot = -1;                 // slot the substep being enqueued writes (-1: none)
    DevBuf<tc::ws::bf16> rec_feat;
    DevBuf<float4> rec_T, rec_ppred, rec_D;
    bool ws_active() const { return ws_mode != 0 && ws_ready && (ws_mode == 1 || ws_net != nullptr); }
    bool ws_tracking() const { return ws_active() || rec_cap > 0; }
    uint32_t iters_effective() const { return ws_active() ? ws_iters : (eff_iters ? eff_iters : iterations); }

    // frame graph cache
    hipGraphExec_t graph_exec = nullptr;
    hipGraph_t graph = nullptr;
    uint64_t graph_key[4] = {0, 0, 0, 0};
    // Warm start with a fallback, replayed from graphs: the host reads the gate after the warm
    // iterations and launches fallback chunks only when they run (graphs of the substep's parts).
    struct PartGraph {
        hipGraph_t g = nullptr;
        hipGraphExec_t x = nullptr;
        uint64_t key[4] = {0, 0, 0, 0};
        void clear() {
            if (x) (void)hipGraphExecDestroy(x);
            if (g) (void)hipGraphDestroy(g);
            x = nullptr;
            g = nullptr;
        }
    };
    PartGraph pg_head, pg_chunk, pg_last, pg_tail;
    uint32_t *gate_host = nullptr;  // pinned host copy of the gate's fallback flag
    hipEvent_t gate_ev[3] = {nullptr, nullptr, nullptr};  // completion of the head and of queued chunks
    uint64_t epoch = 1;  // bumped whenever a parameter or buffer baked into the graph changes
    bool use_graph = true;  // TCLOTH_USE_GRAPH=0 launches kernel by kernel (profilers)
    TcProfiler prof;        // set_profiling_enabled: frames launch kernel by kernel with pass events

    // Adaptive substeps and iterations (dispatch.rs:477-607, simulation/mod.rs:332-348, 723-790):
    // state of the host policy and of the feedback it reads from the last frame's positions.
    bool adaptive = false;
    uint32_t min_substeps = 4, max_substeps = 64, base_substeps = 20;
    uint32_t eff_substeps = 0;         // 0: no history (the next step takes its target directly)
    uint32_t eff_iters = 0;            // iterations of the current frame when the policy sets them (0: iterations)
    bool auto_compensate = true, strain_adaptive = true, compact_readback = true;
    float strain_tolerance = 0.008f;
    float cloth_thickness = 0.005f, mesh_char_len = 0.01f, mesh_avg_edge_len = 0.02f;
    uint32_t mesh_diameter_hops = 10;
    float pending_external = 0.0f, last_max_disp = 0.0f, last_eff_disp = 0.0f, last_max_strain = 0.0f;
    std::vector<float> prev_step_pos;  // 3N, positions of the last frame read back
    // pinned host buffers for the positions copied after a frame (two each, so a copy in flight
    // is never overwritten): the policy's feedback, step_async, and the buffered ring's download
    struct HostReadback {
        float *host = nullptr;
        hipEvent_t ev = nullptr;
        bool pending = false;
    };
    HostReadback fb[2], as[2];
    uint32_t fb_next = 0, as_next = 0;
    bool async_pending = false;
    uint32_t as_ready = 0;             // buffer holding the positions of the last step_async
    DevBuf<float> ring;                // step_buffered: up to ring_frames frames of 3N floats on the device
    uint32_t ring_frames = 8, ring_count = 0;
    float *ring_host = nullptr;
    int prof_begin(const char *name) { return prof.begin(stream, name); }
    void prof_end(int idx) { prof.end(stream, idx); }

    ~tc_sim() {
        if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
        if (graph) (void)hipGraphDestroy(graph);
        pg_head.clear();
        pg_chunk.clear();
        pg_last.clear();
        pg_tail.clear();
        if (gate_host) (void)hipHostFree(gate_host);
        for (hipEvent_t &ev : gate_ev)
            if (ev) (void)hipEventDestroy(ev);
        for (HostReadback *b : {&fb[0], &fb[1], &as[0], &as[1]}) {
            if (b->host) (void)hipHostFree(b->host);
            if (b->ev) (void)hipEventDestroy(b->ev);
        }
        if (ring_host) (void)hipHostFree(ring_host);
        if (dyn.mats_ev) (void)hipEventDestroy(dyn.mats_ev);
        if (dyn.mats_host) (void)hipHostFree(dyn.mats_host);
        if (stream) (void)hipStreamDestroy(stream);
    }
    void touch() { ++epoch; }

    tc::SubstepArgs args(float dt_sub) const {
        tc::SubstepArgs a;
        a.dt = dt_sub;
        const float k = priority_scale();  // dispatch.rs:13-21
        a.gx = gravity[0] * k; a.gy = gravity[1] * k; a.gz = gravity[2] * k;
        a.damping = damping;
        a.tension_damping = tension_damping; a.compression_damping = compression_damping;
        a.shear_damping = shear_damping; a.bending_damping = bending_damping;
        return a;
    }
    tc::SelfColArgs sc_args() const {
        tc::SelfColArgs a;
        a.cell = cell_size; a.table = table_size;
        a.nv = mesh.nv; a.ne = uint32_t(mesh.e_v0.size());
        a.relief_factor = relief_factor; a.max_disp_ratio = max_disp_ratio;
        a.enable_relief = enable_relief; a.exclude_neighbors = exclude_neighbors;
        a.max_search_iterations = max_search_iterations;
        a.untangle = untangle_active() ? 1 : 0;
        return a;
    }
    tc::SelfColTopo topo() const {
        tc::SelfColTopo t;
        t.local_len = local_len.p; t.th_off = th_off.p; t.th_idx = th_idx.p; t.th_rest = th_rest.p;
        t.star_off = star_off.p; t.star = star.p; t.edges = edges.p; t.edge_bbits = edge_bbits.p;
        t.layer = layer_d.p; t.normals = normals.p; t.adj_off = adj_off.p; t.adj_idx = adj_idx.p;
        return t;
    }
    // effective_normal_untangling (simulation/mod.rs:885-890): on, and the cloth has several layers
    bool untangle_active() const { return normal_untangling && mesh.has_multiple_layers && layer_d.p && normals.p; }
    bool fused_active() const {
        if (!fused_ready) return false;
        if (solver_mode == 1 || solver_path == 1) return false;
        // one launch per iteration at every size: on the 1,985-vertex curtain it runs 2.5x the
        // substeps of one launch per colour (16 colour launches per iteration)
        return true;
    }
    float4 *Pbuf(uint32_t k) const { return k ? Palt.p : P.p; }
    float *dlam_buf(uint32_t k) const { return k ? dlam2.p : dlam.p; }
    float *blam_buf(uint32_t k) const { return k ? blam2.p : blam.p; }
    float *slam_buf(uint32_t k) const { return k ? slam2.p : slam.p; }
    tc::FusedArgs fused_args() const {
        tc::FusedArgs f{};
        f.ntiles = plan.ntiles; f.nphases = plan.nphases; f.nd_phases = plan.nd_phases; f.max_ext = plan.max_ext;
        f.fast = fast_projection ? 1u : 0u;
        f.ns_phases = plan.ns_phases;
        f.sew_comp = sew_stiffness >= 5000.0f ? 0.0f : 1.0f / std::max(sew_stiffness, 1.0f);  // dispatch.rs:29-33
        f.sew_lock = sew_lock ? 1.0f : 0.0f;
        f.sew_lock_dist = sew_lock_dist;
        for (size_t c = 0; c < plan.dcls_tc.size(); ++c) { f.dtc[c] = plan.dcls_tc[c]; f.dcc[c] = plan.dcls_cc[c]; }
        for (size_t c = 0; c < plan.bcls.size(); ++c) f.bcomp[c] = plan.bcls[c];
        return f;
    }
    tc::FusedPlanDev plan_dev() const {
        return {pl_ext_off.p, pl_ext_vid.p, pl_n_own.p, pl_ph_off.p, pl_items.p, srec.p, brec.p, bgate.p};
    }
    tc::PairArgs pair_args() const {
        tc::PairArgs a;
        a.cell = cell_size; a.table = table_size; a.nv = mesh.nv;
        a.safety_margin = pair_safety; a.exclude_neighbors = exclude_neighbors;
        a.mode_f = float(pair_margin_mode);
        a.dt_frame = pair_dt_frame; a.horizon_scale = pair_hscale; a.max_horizon = pair_hmax;
        a.quota_vt = std::min<uint32_t>(std::max<uint32_t>(pair_quota_vt, 1u), 8u);
        a.quota_ee = std::min<uint32_t>(std::max<uint32_t>(pair_quota_ee, 1u), 8u);
        return a;
    }
    tc::HashBufs vhash() const { return {vkeys.p, vcounts.p, vstarts.p, vsorted.p}; }
    tc::HashBufs ehash() const { return {ekeys.p, ecounts.p, estarts.p, esorted.p}; }
};

// Wrinkle field after predict (tc_wrinkle.hip); nothing when off or without a texture.
// sewing (tc_sewing.hip): the per-substep shrink (also clears the sewing lambdas) and one pass over the colours
hipError_t enqueue_sew_shrink(tc_sim *s, float dt_sub);
hipError_t enqueue_sewing(tc_sim *s, float4 *P, float dt_sub, float *lam);
int upload_sewing(tc_sim *s);
hipError_t enqueue_wrinkle(tc_sim *s, float dt_sub);

// Collider-edge collision (tc_edge_collision.hip): the edges' rebuild after the triangles change,
// and the pass (kernel + apply) on P.
int upload_collider_edges(tc_sim *s);
hipError_t enqueue_edge_collision(tc_sim *s, float4 *P);

// Dynamic pins (tc_pins.hip): the pin projection of the current table on P, the table's upload,
// and the inverse mass a vertex takes when its state is set from outside.
hipError_t enqueue_pins(tc_sim *s, float4 *P, const tc::SubstepArgs &a);
int upload_pins(tc_sim *s);
float effective_inv_mass(const tc_sim *s, uint32_t i);

// One frame of `substeps` substeps on the simulator's stream (graph replay when it applies).
int tc_run_frame(tc_sim *s, float dt, uint32_t substeps);
// The dynamic bone SDF's update at the start of a step / step_async / step_buffered frame
// (dispatch_dynamic_bone_sdf_update, simulation/mod.rs:2320-2457), enqueued on the simulator's
// stream ahead of the frame (tc_dyn_sdf.hip).
int dyn_sdf_frame(tc_sim *s);
// The policy's mesh constants (simulation/mod.rs:380-420), set once at creation (tc_frames.hip).
void frames_init(tc_sim *s);

// Learned warm start hooks around the projection loop of one substep (tc_ws_api.hip).
// ws_pre runs after predict: features, network and apply (or only the copy of the
// prediction when recording plain substeps); ws_post runs after the projection loop.
hipError_t ws_pre(tc_sim *s, float dt_sub);
hipError_t ws_post(tc_sim *s, float dt_sub, const float4 *Pc);
float ws_scale(const tc_sim *s, float dt_sub);
// Zeroes the previous-substep correction (after the state is set from outside).
int ws_clear_history(tc_sim *s);
// `count` projection iterations of the active schedule (tc_sim.hip); pc / lc track the
// buffers holding P and lambda on the per-iteration path.
hipError_t enqueue_gated_fallback(tc_sim *s, const tc::SubstepArgs &a, float dt_sub, uint32_t &pc, uint32_t &lc);
hipError_t enqueue_iterations(tc_sim *s, const tc::SubstepArgs &a, uint32_t count, bool first, uint32_t &pc,
                              uint32_t &lc);
