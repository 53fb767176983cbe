// Cone-tile planner (see tc_tiles.h).
#include "tc_tiles.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>

namespace tc {
namespace {

// Recursive coordinate bisection of the rest positions: split at the median of
// the longest extent until a part holds at most `leaf` vertices.  Parts come
// out depth first, so consecutive parts are spatial neighbours.
void rcb(const std::vector<float> &pos, uint32_t *idx, size_t n, uint32_t leaf, std::vector<uint32_t> &order,
         std::vector<uint32_t> &part_size) {
    if (n <= leaf) {
        order.insert(order.end(), idx, idx + n);
        part_size.push_back(uint32_t(n));
        return;
    }
    float lo[3] = {3.4e38f, 3.4e38f, 3.4e38f}, hi[3] = {-3.4e38f, -3.4e38f, -3.4e38f};
    for (size_t k = 0; k < n; ++k)
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::min(lo[a], pos[3 * size_t(idx[k]) + a]);
            hi[a] = std::max(hi[a], pos[3 * size_t(idx[k]) + a]);
        }
    int ax = 0;
    for (int a = 1; a < 3; ++a)
        if (hi[a] - lo[a] > hi[ax] - lo[ax]) ax = a;
    // split into parts whose sizes are multiples of the leaf size where possible
    const size_t leaves = (n + leaf - 1) / leaf;
    const size_t left = ((leaves + 1) / 2) * size_t(leaf) < n ? ((leaves + 1) / 2) * size_t(leaf) : n / 2;
    std::nth_element(idx, idx + left, idx + n, [&](uint32_t a, uint32_t b) {
        const float pa = pos[3 * size_t(a) + ax], pb = pos[3 * size_t(b) + ax];
        return pa < pb || (pa == pb && a < b);
    });
    rcb(pos, idx, left, leaf, order, part_size);
    rcb(pos, idx + left, n - left, leaf, order, part_size);
}

}  // namespace

std::string build_tile_plan(const Mesh &m, uint32_t tile_vertices, uint32_t max_lds_vertices, TilePlan &P) {
    P = TilePlan();
    const uint32_t nv = m.nv;
    const size_t nd = m.d_v0.size(), nb = m.b_v0.size();
    if (nv == 0) return "empty mesh";
    const size_t ns = m.s_v0.size();
    if (nd >= (1u << 26) || nb >= (1u << 26) || ns >= (1u << 26)) return "too many constraints for the item encoding";
    // seam-gated bending (gated only where the mesh has seams, as in the per-colour sweep) also reads
    // its hinge vertices v0, v1: their values enter the cone, they are not written
    auto gated = [&](size_t k) { return m.has_seams && m.b_pad0[k] > 0.5f && m.b_pad0[k] < 1.5f; };
    for (size_t k = 0; k < nb; ++k)
        if (gated(k) && (m.b_v0[k] >= nv || m.b_v1[k] >= nv)) return "seam-gated bending with a hinge vertex out of range";

    // compliance classes
    std::vector<uint32_t> dcls(nd), bcls(nb);
    for (size_t k = 0; k < nd; ++k) {
        uint32_t c = 0;
        for (; c < P.dcls_tc.size(); ++c)
            if (P.dcls_tc[c] == m.d_tc[k] && P.dcls_cc[c] == m.d_cc[k]) break;
        if (c == P.dcls_tc.size()) {
            if (c == 16) return "more than 16 distinct distance compliances";
            P.dcls_tc.push_back(m.d_tc[k]);
            P.dcls_cc.push_back(m.d_cc[k]);
        }
        dcls[k] = c;
    }
    for (size_t k = 0; k < nb; ++k) {
        uint32_t c = 0;
        for (; c < P.bcls.size(); ++c)
            if (P.bcls[c] == m.b_comp[k]) break;
        if (c == P.bcls.size()) {
            if (c == 16) return "more than 16 distinct bending compliances";
            P.bcls.push_back(m.b_comp[k]);
        }
        bcls[k] = c;
    }

    // phases of one iteration (distance colours, sewing colours, bending colours, the order of the
    // per-colour sweep) and, per phase, the constraint writing each vertex
    const uint32_t ndp = uint32_t(m.d_col.count.size()), nsp = uint32_t(m.s_col.count.size()),
                   nbp = uint32_t(m.b_col.count.size());
    const uint32_t nds = ndp + nsp, L = nds + nbp;
    P.nphases = L;
    P.nd_phases = ndp;
    P.ns_phases = nsp;
    std::vector<uint32_t> touch(size_t(L) * nv, UINT32_MAX);
    for (uint32_t c = 0; c < ndp; ++c)
        for (uint32_t k = m.d_col.offset[c]; k < m.d_col.offset[c] + m.d_col.count[c]; ++k) {
            touch[size_t(c) * nv + m.d_v0[k]] = k;
            touch[size_t(c) * nv + m.d_v1[k]] = k;
        }
    for (uint32_t c = 0; c < nsp; ++c)
        for (uint32_t k = m.s_col.offset[c]; k < m.s_col.offset[c] + m.s_col.count[c]; ++k) {
            touch[size_t(ndp + c) * nv + m.s_v0[k]] = k;
            touch[size_t(ndp + c) * nv + m.s_v1[k]] = k;
        }
    for (uint32_t c = 0; c < nbp; ++c)
        for (uint32_t k = m.b_col.offset[c]; k < m.b_col.offset[c] + m.b_col.count[c]; ++k) {
            touch[size_t(nds + c) * nv + m.b_v2[k]] = k;
            touch[size_t(nds + c) * nv + m.b_v3[k]] = k;
        }

    const uint32_t T = std::max<uint32_t>(1, tile_vertices);
    std::vector<uint32_t> order, part_size, idx(nv);
    std::iota(idx.begin(), idx.end(), 0u);
    if (m.part_off.size() >= 2 && m.part_off.back() == nv) {
        // independent instances: tiles never mix two of them (their rest positions may overlap)
        for (size_t k = 0; k + 1 < m.part_off.size(); ++k)
            if (m.part_off[k + 1] > m.part_off[k])
                rcb(m.pos0, idx.data() + m.part_off[k], m.part_off[k + 1] - m.part_off[k], T, order, part_size);
    } else {
        rcb(m.pos0, idx.data(), nv, T, order, part_size);
    }
    P.ntiles = uint32_t(part_size.size());
    std::vector<uint32_t> part_begin(P.ntiles + 1, 0), tile_of(nv);
    for (uint32_t t = 0; t < P.ntiles; ++t) part_begin[t + 1] = part_begin[t] + part_size[t];
    for (uint32_t t = 0; t < P.ntiles; ++t)
        for (uint32_t r = part_begin[t]; r < part_begin[t + 1]; ++r) tile_of[order[r]] = t;

    std::vector<uint32_t> vstamp(nv, UINT32_MAX), local(nv, 0);
    std::vector<uint32_t> dsel(nd, UINT32_MAX), ssel(ns, UINT32_MAX), bsel(nb, UINT32_MAX);
    std::vector<std::vector<uint32_t>> phase_sel(L);
    std::vector<uint32_t> cone;
    P.ext_off.assign(P.ntiles + 1, 0);
    P.n_own.assign(P.ntiles, 0);
    P.ph_off.assign(size_t(P.ntiles) * (L + 1), 0);
    P.plain_items = uint64_t(nd) + ns + nb;

    for (uint32_t t = 0; t < P.ntiles; ++t) {
        cone.clear();
        const uint32_t r0 = part_begin[t], r1 = part_begin[t + 1];
        for (uint32_t r = r0; r < r1; ++r) {
            const uint32_t v = order[r];
            vstamp[v] = t;
            local[v] = uint32_t(cone.size());
            cone.push_back(v);
        }
        P.n_own[t] = r1 - r0;
        auto need = [&](uint32_t v) {
            if (vstamp[v] != t) {
                vstamp[v] = t;
                local[v] = uint32_t(cone.size());
                cone.push_back(v);
            }
        };
        for (uint32_t c = L; c-- > 0;) {
            auto &sel = phase_sel[c];
            sel.clear();
            const uint32_t kind = c < ndp ? 0u : c < nds ? 1u : 2u;
            const size_t base = size_t(c) * nv;
            const size_t n_now = cone.size();  // vertices needed after phase c
            for (size_t q = 0; q < n_now; ++q) {
                const uint32_t k = touch[base + cone[q]];
                if (k == UINT32_MAX) continue;
                uint32_t &st = kind == 0u ? dsel[k] : kind == 1u ? ssel[k] : bsel[k];
                if (st == t) continue;
                st = t;
                sel.push_back(k);
            }
            for (uint32_t k : sel) {
                if (kind == 0u) { need(m.d_v0[k]); need(m.d_v1[k]); }
                else if (kind == 1u) { need(m.s_v0[k]); need(m.s_v1[k]); }
                else {
                    need(m.b_v2[k]); need(m.b_v3[k]);
                    if (gated(k)) { need(m.b_v0[k]); need(m.b_v1[k]); }
                }
            }
        }
        if (cone.size() > max_lds_vertices || cone.size() > 65535)
            return "dependency cone of a tile exceeds the LDS budget (" + std::to_string(cone.size()) + " vertices)";
        P.max_ext = std::max<uint32_t>(P.max_ext, uint32_t(cone.size()));
        P.ext_vid.insert(P.ext_vid.end(), cone.begin(), cone.end());
        P.ext_off[t + 1] = uint32_t(P.ext_vid.size());
        for (uint32_t c = 0; c < L; ++c) {
            auto &sel = phase_sel[c];
            std::sort(sel.begin(), sel.end());
            P.ph_off[size_t(t) * (L + 1) + c] = uint32_t(P.items.size() / 3);
            const uint32_t kind = c < ndp ? 0u : c < nds ? 1u : 2u;
            for (uint32_t k : sel) {
                const uint32_t a = kind == 0u ? m.d_v0[k] : kind == 1u ? m.s_v0[k] : m.b_v2[k];
                const uint32_t b = kind == 0u ? m.d_v1[k] : kind == 1u ? m.s_v1[k] : m.b_v3[k];
                const bool owner = tile_of[a] == t;
                uint32_t w1 = k | (owner ? 0x80000000u : 0u), w2 = 0u;
                if (kind == 0u) {
                    std::memcpy(&w2, &m.d_rest[k], 4);
                    w1 |= (dcls[k] << 26) | (m.d_type[k] == 1u ? 1u << 30 : 0u);
                } else if (kind == 2u) {
                    w1 |= bcls[k] << 26;
                    if (gated(k)) {  // the rest length is read from the bending record
                        w1 |= 1u << 30;
                        w2 = local[m.b_v0[k]] | (local[m.b_v1[k]] << 16);
                    } else {
                        std::memcpy(&w2, &m.b_rest[k], 4);
                    }
                }  // sewing: the current rest length lives in the sewing record (it shrinks and locks)
                P.items.push_back(local[a] | (local[b] << 16));
                P.items.push_back(w1);
                P.items.push_back(w2);
            }
        }
        P.ph_off[size_t(t) * (L + 1) + L] = uint32_t(P.items.size() / 3);
    }
    return std::string();
}

}  // namespace tc
