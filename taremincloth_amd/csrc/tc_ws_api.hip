// C ABI of the learned warm start (include/tc_sim.h, tc_ws_* and tc_sim_ws_*):
// network objects, the hooks that run it inside a substep, recording of
// training samples from plain substeps, and training on the GPU.
// No reference counterpart (the network sits between predict and the projection
// loop, crates/cloth_core/src/simulation/dispatch.rs:137 -> 169).
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>

#include "tc_sim_impl.h"

using tc::ws::bf16;

namespace {

uint64_t splitmix64(uint64_t &x) {
    uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
float uniform01(uint64_t &x) { return float(splitmix64(x) >> 40) * (1.0f / 16777216.0f); }

// float -> bf16, round to nearest even (finite inputs)
uint16_t f2bf(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return uint16_t(u >> 16);
}
float bf2f(uint16_t b) {
    const uint32_t u = uint32_t(b) << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

// CSR of the pairs (a[k], b[k]) in both directions, in pair order.
void build_csr(uint32_t nv, const std::vector<uint32_t> &a, const std::vector<uint32_t> &b, const std::vector<float> *rest,
               const std::vector<uint32_t> *type, std::vector<uint32_t> &off, std::vector<uint32_t> &nbr,
               std::vector<float> &rr) {
    off.assign(size_t(nv) + 1, 0);
    auto keep = [&](size_t k) { return !type || (*type)[k] == 0u; };
    for (size_t k = 0; k < a.size(); ++k)
        if (keep(k)) { ++off[a[k] + 1]; ++off[b[k] + 1]; }
    for (uint32_t i = 0; i < nv; ++i) off[i + 1] += off[i];
    nbr.assign(off[nv], 0);
    rr.assign(off[nv], 0.0f);
    std::vector<uint32_t> fill(off.begin(), off.end() - 1);
    for (size_t k = 0; k < a.size(); ++k) {
        if (!keep(k)) continue;
        const float L = rest ? (*rest)[k] : 0.0f;
        nbr[fill[a[k]]] = b[k]; rr[fill[a[k]]++] = L;
        nbr[fill[b[k]]] = a[k]; rr[fill[b[k]]++] = L;
    }
}

int ws_prepare(tc_sim *s) {
    if (s->ws_ready) return TC_OK;
    const tc::Mesh &m = s->mesh;
    const uint32_t nv = m.nv;
    // Each upload reads its own host array, and every array lives until the stream is synchronized
    // below: an asynchronous copy from pageable memory may still be reading it after the call returns.
    std::vector<uint32_t> doff, dnbr, boff, bnbr;
    std::vector<float> drr, brr, grr;
    auto up32 = [&](DevBuf<uint32_t> &d, const std::vector<uint32_t> &h) -> hipError_t {
        hipError_t e = d.alloc(std::max<size_t>(h.size(), 1));
        if (e != hipSuccess || h.empty()) return e;
        return hipMemcpyAsync(d.p, h.data(), 4 * h.size(), hipMemcpyHostToDevice, s->stream);
    };
    auto upf = [&](DevBuf<float> &d, const std::vector<float> &h) -> hipError_t {
        hipError_t e = d.alloc(std::max<size_t>(h.size(), 1));
        if (e != hipSuccess || h.empty()) return e;
        return hipMemcpyAsync(d.p, h.data(), 4 * h.size(), hipMemcpyHostToDevice, s->stream);
    };
    build_csr(nv, m.d_v0, m.d_v1, &m.d_rest, nullptr, doff, dnbr, drr);
    TC_HIP(up32(s->ws_doff, doff));
    TC_HIP(up32(s->ws_dnbr, dnbr));
    TC_HIP(upf(s->ws_drest, drr));
    build_csr(nv, m.b_v2, m.b_v3, &m.b_rest, nullptr, boff, bnbr, brr);
    TC_HIP(up32(s->ws_boff, boff));
    TC_HIP(up32(s->ws_bnbr, bnbr));
    TC_HIP(upf(s->ws_brest, brr));
    build_csr(nv, m.d_v0, m.d_v1, nullptr, &m.d_type, s->ws_goff_h, s->ws_gnbr_h, grr);
    TC_HIP(up32(s->ws_goff, s->ws_goff_h));
    s->ws_gmaxdeg = tc::ws::max_degree(s->ws_goff_h.data(), nv);
    TC_HIP(up32(s->ws_gnbr, s->ws_gnbr_h));
    s->ws_lscale = m.mean_edge > 0.0f ? m.mean_edge : 0.01f;
    TC_HIP(s->ws_ppred.alloc(nv));
    TC_HIP(s->ws_D.alloc(nv));
    TC_HIP(hipMemsetAsync(s->ws_D.p, 0, sizeof(float4) * std::max<size_t>(nv, 1), s->stream));
    TC_HIP(s->ws_feat.alloc(size_t(nv) * tc::ws::F));
    TC_HIP(s->ws_h0.alloc(tc::ws::padded_rows(nv) * tc::ws::H));
    TC_HIP(s->ws_h1.alloc(tc::ws::padded_rows(nv) * tc::ws::H));
    TC_HIP(s->ws_y.alloc(size_t(nv) * tc::ws::O));
    TC_HIP(s->ws_gate.alloc(8));
    TC_HIP(hipMemsetAsync(s->ws_gate.p, 0, 32, s->stream));
    TC_HIP(s->ws_res.alloc(4));
    TC_HIP(hipMemsetAsync(s->ws_res.p, 0, 32, s->stream));
    TC_HIP(tc::ws::set_lds_limits());
    TC_HIP(hipStreamSynchronize(s->stream));
    s->ws_ready = true;
    return TC_OK;
}

tc::ws::FeatArgs feat_args(const tc_sim *s, float dt_sub) {
    tc::ws::FeatArgs a;
    a.P = s->P.p;
    a.D = s->ws_D.p;
    a.d_off = s->ws_doff.p; a.d_nbr = s->ws_dnbr.p; a.d_rest = s->ws_drest.p;
    a.b_off = s->ws_boff.p; a.b_nbr = s->ws_bnbr.p; a.b_rest = s->ws_brest.p;
    a.cols = s->cols.p;
    a.ncols = uint32_t(s->cols_h.size());
    a.X = s->X.p;
    a.nv = s->mesh.nv;
    a.s = ws_scale(s, dt_sub);
    a.lscale = s->ws_lscale;
    a.beta = s->ws_beta;
    return a;
}

tc::ws::FwdBufs infer_bufs(tc_sim *s, const bf16 *feat) {
    tc::ws::FwdBufs b{};
    b.feat = feat;
    for (int r = 0; r < 8; ++r) b.h[r] = (r & 1) ? s->ws_h1.p : s->ws_h0.p;
    b.y = s->ws_y.p;
    b.training = false;
    return b;
}

int net_alloc_scratch(tc_ws_net *n, uint32_t rows) {
    if (rows <= n->rows_cap) return TC_OK;
    const size_t R = rows, H = tc::ws::H;
    const size_t Rp = tc::ws::padded_rows(rows);
    for (uint32_t r = 0; r <= n->nl.rounds; ++r) TC_HIP(n->h[r].alloc(Rp * H));
    for (uint32_t r = 0; r < n->nl.rounds; ++r) {
        TC_HIP(n->x_r[r].alloc(Rp * 2 * H));
        TC_HIP(n->a_r[r].alloc(Rp * H));
    }
    TC_HIP(n->a_enc.alloc(R * H));
    TC_HIP(n->a_dec.alloc(R * H));
    TC_HIP(n->dyb.alloc(R * H));
    TC_HIP(n->dzb.alloc(R * H));
    TC_HIP(n->feat.alloc(R * tc::ws::F));
    TC_HIP(n->y.alloc(R * tc::ws::O));
    TC_HIP(n->dy.alloc(R * tc::ws::O));
    TC_HIP(n->dh0.alloc(R * H));
    TC_HIP(n->dh1.alloc(R * H));
    TC_HIP(n->dx.alloc(R * 2 * H));
    TC_HIP(n->tgt.alloc(R));
    TC_HIP(n->loss_part.alloc(tc::ws::loss_blocks(rows)));
    if (!n->partial.p) TC_HIP(n->partial.alloc(tc::ws::partial_floats(n->nl, 256)));
    n->rows_cap = rows;
    return TC_OK;
}

tc::ws::FwdBufs train_bufs(tc_ws_net *n, const bf16 *feat) {
    tc::ws::FwdBufs b{};
    b.feat = feat;
    for (uint32_t r = 0; r <= n->nl.rounds; ++r) b.h[r] = n->h[r].p;
    b.y = n->y.p;
    b.a_enc = n->a_enc.p;
    b.a_dec = n->a_dec.p;
    for (uint32_t r = 0; r < n->nl.rounds; ++r) { b.x_r[r] = n->x_r[r].p; b.a_r[r] = n->a_r[r].p; }
    b.training = true;
    return b;
}

tc::ws::FwdBufs infer_like(tc_ws_net *n, const bf16 *feat) {
    tc::ws::FwdBufs b = train_bufs(n, feat);
    b.training = false;
    return b;
}

tc::ws::BwdBufs bwd_bufs(tc_ws_net *n, uint32_t rows) {
    tc::ws::BwdBufs b{};
    b.dy = n->dy.p;
    b.dh0 = n->dh0.p;
    b.dh1 = n->dh1.p;
    b.dx = n->dx.p;
    b.dyb = n->dyb.p;
    b.dzb = n->dzb.p;
    b.partial = n->partial.p;
    b.partial_blocks = tc::ws::gemm_blocks(rows);
    b.grad = n->grad.p;
    b.loss_part = n->loss_part.p;
    return b;
}

int adam_step(tc_ws_net *n, float lr) {
    ++n->step;
    tc::ws::AdamArgs a;
    a.n = n->nl.nparams;
    a.lr = lr;
    a.beta1 = 0.9f;
    a.beta2 = 0.999f;
    a.eps = 1e-8f;
    a.bc1 = float(1.0 - std::pow(0.9, double(n->step)));
    a.bc2 = float(1.0 - std::pow(0.999, double(n->step)));
    TC_HIP(tc::ws::launch_adam(n->stream, n->nl, a, n->theta.p, n->m.p, n->v.p, n->grad.p, n->w.p, n->wt.p));
    return TC_OK;
}

double read_loss(tc_ws_net *n, uint32_t rows, hipError_t *err) {
    std::vector<double> h(tc::ws::loss_blocks(rows));
    *err = hipMemcpyAsync(h.data(), n->loss_part.p, 8 * h.size(), hipMemcpyDeviceToHost, n->stream);
    if (*err == hipSuccess) *err = hipStreamSynchronize(n->stream);
    double s = 0.0;
    for (double v : h) s += v;
    return s;
}

}  // namespace

float ws_scale(const tc_sim *s, float dt_sub) {
    const float g = std::sqrt(s->gravity[0] * s->gravity[0] + s->gravity[1] * s->gravity[1] +
                              s->gravity[2] * s->gravity[2]);
    const float dt = std::max(dt_sub, 1e-8f);
    return std::max(g, 1.0f) * dt * dt;
}

hipError_t ws_pre(tc_sim *s, float dt_sub) {
    const uint32_t nv = s->mesh.nv;
    const int slot = s->rec_slot;
    const bool active = s->ws_active();
    const bool need_feat = (active && s->ws_mode == 2) || slot >= 0;
    bf16 *feat = slot >= 0 ? s->rec_feat.p + size_t(slot) * nv * tc::ws::F : s->ws_feat.p;
    const float sc = ws_scale(s, dt_sub);
    hipError_t e = hipSuccess;
    if (need_feat && (e = tc::ws::launch_features(s->stream, feat_args(s, dt_sub), feat)) != hipSuccess) return e;
    if (active && s->ws_mode == 2) {
        const tc::ws::Graph g{s->ws_goff.p, s->ws_gnbr.p, s->ws_gmaxdeg};
        if ((e = tc::ws::forward(s->stream, s->ws_net->nl, s->ws_net->dev(), g, nv, infer_bufs(s, feat))) != hipSuccess)
            return e;
        e = tc::ws::launch_apply(s->stream, s->P.p, s->ws_ppred.p, s->ws_D.p, s->ws_y.p, sc, s->ws_beta, nv);
    } else if (active) {
        e = tc::ws::launch_apply(s->stream, s->P.p, s->ws_ppred.p, s->ws_D.p, nullptr, sc, s->ws_beta, nv);
    } else {
        e = tc::ws::launch_copy_pred(s->stream, s->P.p, s->ws_ppred.p, nv);
    }
    if (e != hipSuccess) return e;
    if (slot >= 0) {
        if ((e = hipMemcpyAsync(s->rec_ppred.p + size_t(slot) * nv, s->ws_ppred.p, sizeof(float4) * nv,
                                hipMemcpyDeviceToDevice, s->stream)) != hipSuccess) return e;
        e = hipMemcpyAsync(s->rec_D.p + size_t(slot) * nv, s->ws_D.p, sizeof(float4) * nv, hipMemcpyDeviceToDevice,
                           s->stream);
    }
    return e;
}

hipError_t ws_post(tc_sim *s, float dt_sub, const float4 *Pc) {
    const uint32_t nv = s->mesh.nv;
    float4 *T = s->rec_slot >= 0 ? s->rec_T.p + size_t(s->rec_slot) * nv : nullptr;
    return tc::ws::launch_update_prev(s->stream, Pc, s->ws_ppred.p, s->ws_D.p, T, ws_scale(s, dt_sub), s->ws_beta, nv);
}

int ws_clear_history(tc_sim *s) {
    if (!s->ws_ready) return TC_OK;
    TC_HIP(hipMemsetAsync(s->ws_D.p, 0, sizeof(float4) * std::max<size_t>(s->mesh.nv, 1), s->stream));
    return TC_OK;
}

// ------------------------------------------------------------------ networks
int tc_ws_net_create(uint32_t rounds, tc_ws_net **out) {
    if (!out) return fail(TC_ERR_VALUE, "null output pointer");
    *out = nullptr;
    if (rounds < 1 || rounds > 4) return fail(TC_ERR_VALUE, "rounds must be 1..4");
    tc_ws_net *n = new tc_ws_net();
    n->nl = tc::ws::make_layout(rounds);
    const size_t P = n->nl.nparams;
    hipError_t e;
    if ((e = hipStreamCreateWithFlags(&n->stream, hipStreamNonBlocking)) != hipSuccess || (e = n->theta.alloc(P)) != hipSuccess ||
        (e = n->m.alloc(P)) != hipSuccess || (e = n->v.alloc(P)) != hipSuccess || (e = n->grad.alloc(P)) != hipSuccess ||
        (e = n->w.alloc(P)) != hipSuccess || (e = n->wt.alloc(P)) != hipSuccess ||
        (e = hipMemsetAsync(n->theta.p, 0, 4 * P, n->stream)) != hipSuccess ||
        (e = hipMemsetAsync(n->m.p, 0, 4 * P, n->stream)) != hipSuccess ||
        (e = hipMemsetAsync(n->v.p, 0, 4 * P, n->stream)) != hipSuccess ||
        (e = hipMemsetAsync(n->grad.p, 0, 4 * P, n->stream)) != hipSuccess ||
        (e = hipMemsetAsync(n->w.p, 0, 2 * P, n->stream)) != hipSuccess ||
        (e = hipMemsetAsync(n->wt.p, 0, 2 * P, n->stream)) != hipSuccess ||
        (e = tc::ws::set_lds_limits()) != hipSuccess || (e = hipStreamSynchronize(n->stream)) != hipSuccess) {
        delete n;
        return fail(TC_ERR_RUNTIME, std::string("tc_ws_net_create: ") + hipGetErrorString(e));
    }
    *out = n;
    return TC_OK;
}

void tc_ws_net_destroy(tc_ws_net *n) {
    if (!n) return;
    if (n->stream) (void)hipStreamSynchronize(n->stream);
    delete n;
}

uint32_t tc_ws_net_num_params(const tc_ws_net *n) { return n ? n->nl.nparams : 0; }
uint32_t tc_ws_net_rounds(const tc_ws_net *n) { return n ? n->nl.rounds : 0; }

int tc_ws_net_set_params(tc_ws_net *n, const float *params, size_t count) {
    if (!n || !params) return fail(TC_ERR_VALUE, "null argument");
    if (count != n->nl.nparams)
        return fail(TC_ERR_VALUE, "expected " + std::to_string(n->nl.nparams) + " parameters, got " + std::to_string(count));
    for (size_t i = 0; i < count; ++i)
        if (!std::isfinite(params[i])) return fail(TC_ERR_VALUE, "parameters must be finite");
    const size_t P = n->nl.nparams;
    TC_HIP(hipMemcpyAsync(n->theta.p, params, 4 * P, hipMemcpyHostToDevice, n->stream));
    TC_HIP(hipMemsetAsync(n->m.p, 0, 4 * P, n->stream));
    TC_HIP(hipMemsetAsync(n->v.p, 0, 4 * P, n->stream));
    TC_HIP(tc::ws::launch_refresh(n->stream, n->nl, n->theta.p, n->w.p, n->wt.p));
    TC_HIP(hipStreamSynchronize(n->stream));
    n->step = 0;
    return TC_OK;
}

int tc_ws_net_get_params(tc_ws_net *n, float *params, size_t count) {
    if (!n || !params) return fail(TC_ERR_VALUE, "null argument");
    if (count < n->nl.nparams) return fail(TC_ERR_VALUE, "output buffer too small");
    TC_HIP(hipMemcpyAsync(params, n->theta.p, 4 * size_t(n->nl.nparams), hipMemcpyDeviceToHost, n->stream));
    TC_HIP(hipStreamSynchronize(n->stream));
    return TC_OK;
}

int tc_ws_net_init(tc_ws_net *n, uint64_t seed) {
    if (!n) return fail(TC_ERR_VALUE, "null network");
    std::vector<float> p(n->nl.nparams, 0.0f);
    uint64_t st = seed;
    for (uint32_t li = 0; li < n->nl.nlayers; ++li) {
        const tc::ws::Layer L = n->nl.L[li];
        float bound = std::sqrt(6.0f / float(L.K));      // He-uniform for layers followed by relu
        const bool residual_out = li >= 2 && li < n->nl.dec1() && (li - 2) % 2 == 1;
        if (residual_out || li == n->nl.enc2()) bound *= 0.25f;
        if (li == n->nl.dec2()) bound = 0.0f;            // the untrained network leaves D_prev unchanged
        for (uint32_t k = 0; k < L.K * L.N; ++k) p[L.off + k] = (2.0f * uniform01(st) - 1.0f) * bound;
    }
    return tc_ws_net_set_params(n, p.data(), p.size());
}

int tc_ws_net_forward_backward(tc_ws_net *n, const float *feat, uint32_t rows, const uint32_t *graph_off,
                               const uint32_t *graph_nbr, const float *targets, float *out_y, float *out_grad,
                               double *out_loss) {
    if (!n || !feat || !graph_off || (rows && !graph_nbr && graph_off[rows] > 0))
        return fail(TC_ERR_VALUE, "null argument");
    if (rows == 0) return fail(TC_ERR_VALUE, "rows must be positive");
    const uint32_t nn = graph_off[rows];
    for (uint32_t i = 0; i < rows; ++i)
        if (graph_off[i] > graph_off[i + 1]) return fail(TC_ERR_VALUE, "graph offsets must be non-decreasing");
    for (uint32_t k = 0; k < nn; ++k)
        if (graph_nbr[k] >= rows) return fail(TC_ERR_VALUE, "graph neighbour index out of range");
    for (uint32_t i = 0; i < rows; ++i)  // the backward pass assumes a symmetric graph
        for (uint32_t k = graph_off[i]; k < graph_off[i + 1]; ++k) {
            const uint32_t j = graph_nbr[k];
            uint32_t c_ij = 0, c_ji = 0;
            for (uint32_t q = graph_off[i]; q < graph_off[i + 1]; ++q) c_ij += graph_nbr[q] == j;
            for (uint32_t q = graph_off[j]; q < graph_off[j + 1]; ++q) c_ji += graph_nbr[q] == i;
            if (c_ij != c_ji) return fail(TC_ERR_VALUE, "graph must be symmetric");
        }
    int r = net_alloc_scratch(n, rows);
    if (r != TC_OK) return r;
    std::vector<uint16_t> fb(size_t(rows) * tc::ws::F);
    for (size_t i = 0; i < fb.size(); ++i) fb[i] = f2bf(feat[i]);
    TC_HIP(hipMemcpyAsync(n->feat.p, fb.data(), 2 * fb.size(), hipMemcpyHostToDevice, n->stream));
    TC_HIP(n->goff.alloc(size_t(rows) + 1));
    TC_HIP(n->gnbr.alloc(std::max<uint32_t>(nn, 1)));
    TC_HIP(hipMemcpyAsync(n->goff.p, graph_off, 4 * (size_t(rows) + 1), hipMemcpyHostToDevice, n->stream));
    if (nn) TC_HIP(hipMemcpyAsync(n->gnbr.p, graph_nbr, 4 * size_t(nn), hipMemcpyHostToDevice, n->stream));
    const tc::ws::Graph g{n->goff.p, n->gnbr.p, tc::ws::max_degree(graph_off, rows)};
    // without targets there is no backward pass, so the forward keeps no activations for it
    const tc::ws::FwdBufs fw = targets ? train_bufs(n, n->feat.p) : infer_like(n, n->feat.p);
    TC_HIP(tc::ws::forward(n->stream, n->nl, n->dev(), g, rows, fw));
    if (targets) {
        uint32_t nfree = 0;
        for (uint32_t i = 0; i < rows; ++i) nfree += targets[4 * size_t(i) + 3] > 0.0f;
        TC_HIP(hipMemcpyAsync(n->tgt.p, targets, 16 * size_t(rows), hipMemcpyHostToDevice, n->stream));
        TC_HIP(tc::ws::loss_and_backward(n->stream, n->nl, n->dev(), g, rows, fw, n->tgt.p,
                                         1.0f / float(std::max<uint32_t>(nfree, 1)), bwd_bufs(n, rows)));
        hipError_t le = hipSuccess;
        const double loss = read_loss(n, rows, &le);
        TC_HIP(le);
        if (out_loss) *out_loss = loss;
        if (out_grad) TC_HIP(hipMemcpyAsync(out_grad, n->grad.p, 4 * size_t(n->nl.nparams), hipMemcpyDeviceToHost, n->stream));
    }
    if (out_y) TC_HIP(hipMemcpyAsync(out_y, n->y.p, 4 * size_t(rows) * tc::ws::O, hipMemcpyDeviceToHost, n->stream));
    TC_HIP(hipStreamSynchronize(n->stream));
    return TC_OK;
}

int tc_ws_net_adam_step(tc_ws_net *n, float lr) {
    if (!n) return fail(TC_ERR_VALUE, "null network");
    if (!(lr >= 0.0f) || !std::isfinite(lr)) return fail(TC_ERR_VALUE, "learning rate must be finite and >= 0");
    int r = adam_step(n, lr);
    if (r != TC_OK) return r;
    TC_HIP(hipStreamSynchronize(n->stream));
    return TC_OK;
}

// ------------------------------------------------------------------ simulator side
int tc_sim_set_warm_start(tc_sim *s, tc_ws_net *net, int mode, uint32_t iterations, uint32_t fallback_iterations,
                          float residual_threshold, float history_weight) {
    if (!s) return fail(TC_ERR_VALUE, "null simulator");
    if (mode < 0 || mode > 2) return fail(TC_ERR_VALUE, "mode must be 0 (off), 1 (previous correction) or 2 (network)");
    if (mode == 2 && !net) return fail(TC_ERR_VALUE, "mode 2 needs a network");
    if (mode != 0 && iterations == 0) return fail(TC_ERR_VALUE, "the warm-started solve needs at least one iteration");
    if (!(residual_threshold >= 0.0f) || !std::isfinite(residual_threshold))
        return fail(TC_ERR_VALUE, "residual threshold must be finite and >= 0");
    if (!(history_weight >= 0.0f && history_weight <= 1.0f)) return fail(TC_ERR_VALUE, "history weight must be in [0, 1]");
    if (mode != 0) {
        int r = ws_prepare(s);
        if (r != TC_OK) return r;
    }
    const bool was_active = s->ws_active();
    s->ws_mode = mode;
    s->ws_net = mode == 2 ? net : nullptr;
    s->ws_iters = iterations;
    s->ws_fallback = fallback_iterations + (fallback_iterations & 1u);  // whole pairs: see enqueue_substep
    s->ws_threshold = residual_threshold;
    s->ws_beta = history_weight;
    if (mode != 0 && !was_active && s->rec_cap == 0) {
        int r = ws_clear_history(s);
        if (r != TC_OK) return r;
    }
    s->touch();
    return TC_OK;
}

int tc_sim_warm_start_stats(tc_sim *s, double *out) {
    if (!s || !out) return fail(TC_ERR_VALUE, "null argument");
    for (int k = 0; k < 8; ++k) out[k] = 0.0;
    out[4] = s->ws_mode;
    out[5] = s->ws_iters;
    if (!s->ws_ready) return TC_OK;
    uint32_t g[8];
    double r[4];
    TC_HIP(hipMemcpyAsync(g, s->ws_gate.p, 32, hipMemcpyDeviceToHost, s->stream));
    TC_HIP(hipMemcpyAsync(r, s->ws_res.p, 32, hipMemcpyDeviceToHost, s->stream));
    TC_HIP(hipStreamSynchronize(s->stream));
    out[0] = g[1];
    out[1] = g[2];
    out[2] = r[3];
    out[3] = r[2];
    out[6] = g[3];
    out[7] = g[4];
    return TC_OK;
}

int tc_sim_ws_reset_stats(tc_sim *s) {
    if (!s) return fail(TC_ERR_VALUE, "null simulator");
    if (!s->ws_ready) return TC_OK;
    TC_HIP(hipMemsetAsync(s->ws_gate.p, 0, 32, s->stream));
    TC_HIP(hipMemsetAsync(s->ws_res.p, 0, 32, s->stream));
    TC_HIP(hipStreamSynchronize(s->stream));
    return TC_OK;
}

int tc_sim_ws_record(tc_sim *s, uint32_t capacity, uint32_t stride, float history_weight) {
    if (!s) return fail(TC_ERR_VALUE, "null simulator");
    if (stride == 0) return fail(TC_ERR_VALUE, "stride must be >= 1");
    if (!(history_weight >= 0.0f && history_weight <= 1.0f)) return fail(TC_ERR_VALUE, "history weight must be in [0, 1]");
    if (s->ws_active() && history_weight != s->ws_beta)
        return fail(TC_ERR_VALUE, "the history weight must match the active warm start's");
    TC_HIP(hipStreamSynchronize(s->stream));
    s->rec_feat.release();
    s->rec_T.release();
    s->rec_ppred.release();
    s->rec_D.release();
    s->rec_cap = 0;
    s->rec_count = 0;
    s->rec_tick = 0;
    s->rec_slot = -1;
    if (capacity == 0) return TC_OK;
    int r = ws_prepare(s);
    if (r != TC_OK) return r;
    const size_t nv = s->mesh.nv;
    TC_HIP(s->rec_feat.alloc(size_t(capacity) * nv * tc::ws::F));
    TC_HIP(s->rec_T.alloc(size_t(capacity) * nv));
    TC_HIP(s->rec_ppred.alloc(size_t(capacity) * nv));
    TC_HIP(s->rec_D.alloc(size_t(capacity) * nv));
    if (!s->ws_active()) {
        r = ws_clear_history(s);
        if (r != TC_OK) return r;
    }
    s->rec_cap = capacity;
    s->rec_stride = stride;
    s->ws_beta = history_weight;
    s->touch();
    return TC_OK;
}

uint32_t tc_sim_ws_num_records(const tc_sim *s) { return s ? s->rec_count : 0; }

int tc_sim_ws_get_record(tc_sim *s, uint32_t idx, float *feat, float *target, float *ppred, float *dprev) {
    if (!s) return fail(TC_ERR_VALUE, "null simulator");
    if (idx >= s->rec_count) return fail(TC_ERR_VALUE, "record index out of range");
    const size_t nv = s->mesh.nv;
    TC_HIP(hipStreamSynchronize(s->stream));
    if (feat) {
        std::vector<uint16_t> h(nv * tc::ws::F);
        TC_HIP(hipMemcpy(h.data(), s->rec_feat.p + idx * nv * tc::ws::F, 2 * h.size(), hipMemcpyDeviceToHost));
        for (size_t i = 0; i < h.size(); ++i) feat[i] = bf2f(h[i]);
    }
    if (target) TC_HIP(hipMemcpy(target, s->rec_T.p + idx * nv, 16 * nv, hipMemcpyDeviceToHost));
    if (ppred) TC_HIP(hipMemcpy(ppred, s->rec_ppred.p + idx * nv, 16 * nv, hipMemcpyDeviceToHost));
    if (dprev) TC_HIP(hipMemcpy(dprev, s->rec_D.p + idx * nv, 16 * nv, hipMemcpyDeviceToHost));
    return TC_OK;
}

int tc_sim_ws_last_output(tc_sim *s, float *y, size_t count) {
    if (!s || !y) return fail(TC_ERR_VALUE, "null argument");
    if (!s->ws_ready) return fail(TC_ERR_VALUE, "warm start was never enabled on this simulator");
    if (count < size_t(s->mesh.nv) * tc::ws::O) return fail(TC_ERR_VALUE, "output buffer too small");
    TC_HIP(hipStreamSynchronize(s->stream));
    TC_HIP(hipMemcpy(y, s->ws_y.p, 4 * size_t(s->mesh.nv) * tc::ws::O, hipMemcpyDeviceToHost));
    return TC_OK;
}

int tc_ws_train(tc_ws_net *n, tc_sim *const *sims, uint32_t nsims, uint32_t steps, float lr, uint64_t seed,
                float *losses) {
    if (!n || (!sims && nsims)) return fail(TC_ERR_VALUE, "null argument");
    if (!(lr >= 0.0f) || !std::isfinite(lr)) return fail(TC_ERR_VALUE, "learning rate must be finite and >= 0");
    std::vector<uint32_t> nfree(nsims, 0);
    for (uint32_t k = 0; k < nsims; ++k) {
        tc_sim *s = sims[k];
        if (!s) return fail(TC_ERR_VALUE, "null simulator in the list");
        if (s->rec_count == 0) return fail(TC_ERR_VALUE, "a simulator in the list holds no recorded samples");
        TC_HIP(hipStreamSynchronize(s->stream));
        for (float w : s->mesh.inv_mass) nfree[k] += w > 0.0f;
        int r = net_alloc_scratch(n, s->mesh.nv);
        if (r != TC_OK) return r;
    }
    if (!nsims) return TC_OK;
    uint64_t st = seed;
    for (uint32_t it = 0; it < steps; ++it) {
        const uint32_t k = uint32_t(splitmix64(st) % nsims);
        tc_sim *s = sims[k];
        const uint32_t idx = uint32_t(splitmix64(st) % s->rec_count), rows = s->mesh.nv;
        const bf16 *feat = s->rec_feat.p + size_t(idx) * rows * tc::ws::F;
        const tc::ws::Graph g{s->ws_goff.p, s->ws_gnbr.p, s->ws_gmaxdeg};
        const tc::ws::FwdBufs fw = train_bufs(n, feat);
        TC_HIP(tc::ws::forward(n->stream, n->nl, n->dev(), g, rows, fw));
        TC_HIP(tc::ws::loss_and_backward(n->stream, n->nl, n->dev(), g, rows, fw, s->rec_T.p + size_t(idx) * rows,
                                         1.0f / float(std::max<uint32_t>(nfree[k], 1)), bwd_bufs(n, rows)));
        if (losses) {
            hipError_t le = hipSuccess;
            losses[it] = float(read_loss(n, rows, &le));
            TC_HIP(le);
        }
        int r = adam_step(n, lr);
        if (r != TC_OK) return r;
    }
    TC_HIP(hipStreamSynchronize(n->stream));
    return TC_OK;
}

int tc_sim_ws_history(tc_sim *s, float *get_xyz, const float *set_xyz, uint32_t n) {
    if (!s) return fail(TC_ERR_VALUE, "null simulator");
    if (n != s->mesh.nv) return fail(TC_ERR_VALUE, "expected one xyz triple per vertex");
    if (!s->ws_ready) {
        int r = ws_prepare(s);
        if (r != TC_OK) return r;
    }
    std::vector<float4> h(n);
    TC_HIP(hipStreamSynchronize(s->stream));
    if (get_xyz) {
        TC_HIP(hipMemcpy(h.data(), s->ws_D.p, 16 * size_t(n), hipMemcpyDeviceToHost));
        for (uint32_t i = 0; i < n; ++i) { get_xyz[3 * i] = h[i].x; get_xyz[3 * i + 1] = h[i].y; get_xyz[3 * i + 2] = h[i].z; }
    }
    if (set_xyz) {
        for (uint32_t i = 0; i < n; ++i) h[i] = make_float4(set_xyz[3 * i], set_xyz[3 * i + 1], set_xyz[3 * i + 2], 0.0f);
        TC_HIP(hipMemcpy(s->ws_D.p, h.data(), 16 * size_t(n), hipMemcpyHostToDevice));
    }
    return TC_OK;
}

int tc_ws_eval(tc_ws_net *n, tc_sim *const *sims, uint32_t nsims, double *mean_loss) {
    if (!n || !mean_loss || (!sims && nsims)) return fail(TC_ERR_VALUE, "null argument");
    double total = 0.0;
    uint64_t count = 0;
    for (uint32_t k = 0; k < nsims; ++k) {
        tc_sim *s = sims[k];
        if (!s) return fail(TC_ERR_VALUE, "null simulator in the list");
        TC_HIP(hipStreamSynchronize(s->stream));
        uint32_t nfree = 0;
        for (float w : s->mesh.inv_mass) nfree += w > 0.0f;
        const uint32_t rows = s->mesh.nv;
        int r = net_alloc_scratch(n, rows);
        if (r != TC_OK) return r;
        const tc::ws::Graph g{s->ws_goff.p, s->ws_gnbr.p, s->ws_gmaxdeg};
        for (uint32_t idx = 0; idx < s->rec_count; ++idx) {
            const tc::ws::FwdBufs fw = infer_like(n, s->rec_feat.p + size_t(idx) * rows * tc::ws::F);
            TC_HIP(tc::ws::forward(n->stream, n->nl, n->dev(), g, rows, fw));
            TC_HIP(tc::ws::loss_only(n->stream, rows, n->y.p, s->rec_T.p + size_t(idx) * rows,
                                     1.0f / float(std::max<uint32_t>(nfree, 1)), n->dy.p, n->loss_part.p));
            hipError_t le = hipSuccess;
            total += read_loss(n, rows, &le);
            TC_HIP(le);
            ++count;
        }
    }
    *mean_loss = count ? total / double(count) : 0.0;
    return TC_OK;
}

int tc_sim_ws_time(tc_sim *s, float dt_sub, uint32_t reps, double *out) {
    if (!s || !out) return fail(TC_ERR_VALUE, "null argument");
    for (int k = 0; k < 16; ++k) out[k] = 0.0;
    if (!s->ws_active()) return fail(TC_ERR_VALUE, "warm start is not enabled on this simulator");
    if (s->mesh.nv == 0 || reps == 0) return TC_OK;
    const tc::Mesh &m = s->mesh;
    const uint32_t nv = m.nv, B = s->block, R = s->ws_net ? s->ws_net->nl.rounds : 0;
    const bool net = s->ws_mode == 2;
    std::vector<hipEvent_t> ev(32, nullptr);
    for (auto &e : ev) TC_HIP(hipEventCreate(&e));
    struct Guard {
        std::vector<hipEvent_t> &v;
        ~Guard() { for (auto e : v) if (e) (void)hipEventDestroy(e); }
    } guard{ev};
    const float sc = ws_scale(s, dt_sub);
    double acc[16] = {0};
    s->rec_slot = -1;
    for (uint32_t rep = 0; rep < reps; ++rep) {
        const tc::SubstepArgs a = s->args(dt_sub);
        TC_HIP(hipEventRecord(ev[0], s->stream));
        TC_HIP(tc::launch_predict(s->stream, s->P.p, s->X.p, s->V.p, nv, a, B));
        TC_HIP(hipEventRecord(ev[1], s->stream));
        if (net) {
            TC_HIP(tc::ws::launch_features(s->stream, feat_args(s, dt_sub), s->ws_feat.p));
            TC_HIP(hipEventRecord(ev[2], s->stream));
            const tc::ws::Graph g{s->ws_goff.p, s->ws_gnbr.p, s->ws_gmaxdeg};
            TC_HIP(tc::ws::forward(s->stream, s->ws_net->nl, s->ws_net->dev(), g, nv, infer_bufs(s, s->ws_feat.p),
                                   &ev[3]));  // ev[3 .. 3 + R + 2]
        } else {
            TC_HIP(hipEventRecord(ev[2], s->stream));
        }
        const int fe = net ? 3 + int(R) + 2 : 2;  // event after the network
        TC_HIP(tc::ws::launch_apply(s->stream, s->P.p, s->ws_ppred.p, s->ws_D.p, net ? s->ws_y.p : nullptr, sc,
                                    s->ws_beta, nv));
        TC_HIP(hipEventRecord(ev[20], s->stream));
        uint32_t pc = 0, lc = 0;
        TC_HIP(enqueue_iterations(s, a, s->ws_iters, true, pc, lc));
        TC_HIP(hipEventRecord(ev[21], s->stream));
        TC_HIP(hipEventRecord(ev[22], s->stream));  // gate and fallback chunks are timed together
        if (s->ws_fallback > 0) TC_HIP(enqueue_gated_fallback(s, a, dt_sub, pc, lc));
        TC_HIP(hipEventRecord(ev[23], s->stream));
        float4 *Pc = s->Pbuf(pc);
        s->last_p = pc;
        s->last_lam = lc;
        TC_HIP(tc::ws::launch_update_prev(s->stream, Pc, s->ws_ppred.p, s->ws_D.p, nullptr, sc, s->ws_beta, nv));
        TC_HIP(hipEventRecord(ev[24], s->stream));
        if (!s->cols_h.empty() || !s->tris_h.empty() || s->sdf_nb > 0)
            TC_HIP(tc::launch_collision(s->stream, Pc, s->X.p, nv, s->col_order.p, s->cols.p, uint32_t(s->cols_h.size()), s->tris.p,
                                        s->tri_bounds.p, uint32_t(s->tris_h.size()), s->tri_cbox.p, s->single_sided_recovery, dt_sub,
                                        B, s->sdf_args()));
        if (s->self_collision) {
            TC_HIP(tc::launch_hashes(s->stream, Pc, nv, s->edges.p, uint32_t(m.e_v0.size()), s->cell_size,
                                     s->table_size, s->vhash(), s->ehash()));
            const tc::SelfColArgs sa = s->sc_args();
            const tc::SelfColTopo t = s->topo();
            if (sa.untangle) TC_HIP(tc::launch_compute_normals(s->stream, s->X.p, t, s->normals.p, nv));
            TC_HIP(tc::launch_sc_solve(s->stream, Pc, s->X.p, t, s->vhash(), s->accum.p, sa));
            TC_HIP(tc::launch_sc_untangle(s->stream, Pc, s->X.p, t, s->accum.p, sa));
            TC_HIP(tc::launch_sc_solve_ee(s->stream, Pc, s->X.p, t, s->ehash(), s->accum.p, sa));
            TC_HIP(tc::launch_sc_apply(s->stream, Pc, s->local_len.p, s->accum.p, sa, B));
        }
        TC_HIP(hipEventRecord(ev[25], s->stream));
        TC_HIP(tc::launch_update_vel(s->stream, Pc, s->X.p, s->V.p, nv, dt_sub, B));
        TC_HIP(hipEventRecord(ev[26], s->stream));
        TC_HIP(hipStreamSynchronize(s->stream));
        auto el = [&](int i, int j) -> double {
            float ms = 0.0f;
            (void)hipEventElapsedTime(&ms, ev[i], ev[j]);
            return ms;
        };
        acc[0] += el(0, 1);
        acc[1] += el(1, 2);
        if (net) {
            acc[2] += el(3, 4);
            for (uint32_t r = 0; r < R; ++r) acc[3] += el(4 + r, 5 + r);
            acc[4] += el(4 + R, 5 + R);
        }
        acc[5] += el(fe, 20);
        acc[6] += el(20, 21);
        acc[7] += el(21, 22);
        acc[8] += el(22, 23);
        acc[9] += el(23, 24);
        acc[10] += el(24, 25);
        acc[11] += el(25, 26);
        acc[12] += el(0, 26);
    }
    for (int k = 0; k < 13; ++k) out[k] = acc[k] / double(reps);
    out[13] = R;
    out[14] = s->ws_iters;
    return TC_OK;
}

// ------------------------------------------------------------------ forward pass without allocation
// tc_ws_net_forward_noalloc: the inference half of tc_ws_net_forward_backward with every device
// buffer carved from one file-scope device array, for callers that cannot allocate (graph capture,
// serving loops).  Host code never names `ws_scratch` in a launch or a copy: each device's address
// of it comes from hipGetSymbolAddress once, on the first call on that device, and only that
// address is used.
namespace {

constexpr size_t NA_ALIGN = 256;
constexpr size_t na_round(size_t b) { return (b + NA_ALIGN - 1) / NA_ALIGN * NA_ALIGN; }
constexpr size_t NA_ROWS = TC_WS_NOALLOC_MAX_ROWS;
constexpr size_t NA_NBR = TC_WS_NOALLOC_MAX_NBR;
// layout: fp32 feature staging | BF16 features | hidden state h0 | h1 | outputs y | graph offsets | neighbours
constexpr size_t NA_F32 = 0;
constexpr size_t NA_FEAT = NA_F32 + na_round(NA_ROWS * tc::ws::F * sizeof(float));
constexpr size_t NA_H0 = NA_FEAT + na_round(NA_ROWS * tc::ws::F * sizeof(bf16));
constexpr size_t NA_H1 = NA_H0 + na_round(tc::ws::padded_rows(NA_ROWS) * tc::ws::H * sizeof(bf16));
constexpr size_t NA_Y = NA_H1 + na_round(tc::ws::padded_rows(NA_ROWS) * tc::ws::H * sizeof(bf16));
constexpr size_t NA_OFF = NA_Y + na_round(NA_ROWS * tc::ws::O * sizeof(float));
constexpr size_t NA_NBR_OFF = NA_OFF + na_round((NA_ROWS + 1) * sizeof(uint32_t));
constexpr size_t NA_BYTES = NA_NBR_OFF + na_round(NA_NBR * sizeof(uint32_t));
static_assert(NA_BYTES == TC_WS_NOALLOC_SCRATCH_BYTES, "the header states the scratch size");

__device__ __align__(256) unsigned char ws_scratch[NA_BYTES];

constexpr int NA_MAX_DEVICES = 64;
std::mutex na_mutex;
unsigned char *na_ptr[NA_MAX_DEVICES] = {};

// This device's address of ws_scratch (looked up once per device).
int na_scratch(unsigned char **out) {
    int dev = 0;
    TC_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= NA_MAX_DEVICES) return fail(TC_ERR_RUNTIME, "device ordinal beyond the scratch table");
    std::lock_guard<std::mutex> lock(na_mutex);
    if (!na_ptr[dev]) {
        void *p = nullptr;
        TC_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(ws_scratch)));
        if (!p) return fail(TC_ERR_RUNTIME, "hipGetSymbolAddress returned null for the scratch array");
        na_ptr[dev] = static_cast<unsigned char *>(p);
    }
    *out = na_ptr[dev];
    return TC_OK;
}

// fp32 -> BF16, round to nearest even: the same bit operation as f2bf on the host.
__global__ void __launch_bounds__(256) k_f32_to_bf16(const float *__restrict__ in, uint16_t *__restrict__ out,
                                                     size_t n) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t u = __float_as_uint(in[i]);
    u += 0x7FFFu + ((u >> 16) & 1u);
    out[i] = uint16_t(u >> 16);
}

}  // namespace

int tc_ws_net_forward_noalloc(tc_ws_net *n, const float *feat, uint32_t rows, const uint32_t *graph_off,
                              const uint32_t *graph_nbr, float *out_y) {
    if (!n || !feat || !graph_off || !out_y) return fail(TC_ERR_VALUE, "null argument");
    if (rows == 0) return fail(TC_ERR_VALUE, "rows must be positive");
    if (rows > NA_ROWS)
        return fail(TC_ERR_VALUE, "rows (" + std::to_string(rows) + ") exceed the allocation-free limit of " +
                                      std::to_string(NA_ROWS));
    const uint32_t nn = graph_off[rows];
    if (nn > NA_NBR)
        return fail(TC_ERR_VALUE, "neighbour entries (" + std::to_string(nn) + ") exceed the allocation-free limit of " +
                                      std::to_string(NA_NBR));
    if (nn && !graph_nbr) return fail(TC_ERR_VALUE, "null argument");
    for (uint32_t i = 0; i < rows; ++i)
        if (graph_off[i] > graph_off[i + 1]) return fail(TC_ERR_VALUE, "graph offsets must be non-decreasing");
    for (uint32_t k = 0; k < nn; ++k)
        if (graph_nbr[k] >= rows) return fail(TC_ERR_VALUE, "graph neighbour index out of range");
    unsigned char *base = nullptr;
    int r = na_scratch(&base);
    if (r != TC_OK) return r;
    float *f32 = reinterpret_cast<float *>(base + NA_F32);
    bf16 *fb = reinterpret_cast<bf16 *>(base + NA_FEAT);
    uint32_t *goff = reinterpret_cast<uint32_t *>(base + NA_OFF);
    uint32_t *gnbr = reinterpret_cast<uint32_t *>(base + NA_NBR_OFF);
    const size_t nf = size_t(rows) * tc::ws::F;
    TC_HIP(hipMemcpyAsync(f32, feat, 4 * nf, hipMemcpyHostToDevice, n->stream));
    hipLaunchKernelGGL(k_f32_to_bf16, dim3(unsigned((nf + 255) / 256)), dim3(256), 0, n->stream, f32,
                       reinterpret_cast<uint16_t *>(fb), nf);
    TC_HIP(hipGetLastError());
    TC_HIP(hipMemcpyAsync(goff, graph_off, 4 * (size_t(rows) + 1), hipMemcpyHostToDevice, n->stream));
    if (nn) TC_HIP(hipMemcpyAsync(gnbr, graph_nbr, 4 * size_t(nn), hipMemcpyHostToDevice, n->stream));
    tc::ws::FwdBufs b{};
    b.feat = fb;
    bf16 *h0 = reinterpret_cast<bf16 *>(base + NA_H0), *h1 = reinterpret_cast<bf16 *>(base + NA_H1);
    for (int k = 0; k < 8; ++k) b.h[k] = (k & 1) ? h1 : h0;
    b.y = reinterpret_cast<float *>(base + NA_Y);
    b.training = false;
    const tc::ws::Graph g{goff, gnbr, tc::ws::max_degree(graph_off, rows)};
    TC_HIP(tc::ws::forward(n->stream, n->nl, n->dev(), g, rows, b));
    TC_HIP(hipMemcpyAsync(out_y, b.y, 4 * size_t(rows) * tc::ws::O, hipMemcpyDeviceToHost, n->stream));
    TC_HIP(hipStreamSynchronize(n->stream));
    return TC_OK;
}
