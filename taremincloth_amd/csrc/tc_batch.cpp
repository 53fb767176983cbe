// Batches of independent cloth instances: the disjoint union of their meshes.
#include <algorithm>
#include <cstring>

#include "tc_mesh.h"

namespace tc {

namespace {

// Appends a CSR (off[n + 1], idx) whose indices are vertex ids, shifting both.
void append_csr(std::vector<uint32_t> &off, std::vector<uint32_t> &idx, const std::vector<uint32_t> &poff,
                const std::vector<uint32_t> &pidx, uint32_t voff, bool shift_idx) {
    const uint32_t base = uint32_t(idx.size());
    if (off.empty()) off.push_back(0);
    for (size_t i = 1; i < poff.size(); ++i) off.push_back(base + poff[i]);
    for (uint32_t v : pidx) idx.push_back(shift_idx ? v + voff : v);
}

template <typename T>
void append(std::vector<T> &dst, const std::vector<T> &src) {
    dst.insert(dst.end(), src.begin(), src.end());
}

}  // namespace

void merge_meshes(const std::vector<Mesh> &parts, Mesh &m) {
    m = Mesh();
    size_t nd_colours = 0, nb_colours = 0;
    for (const Mesh &p : parts) {
        nd_colours = std::max(nd_colours, p.d_col.count.size());
        nb_colours = std::max(nb_colours, p.b_col.count.size());
    }
    std::vector<uint32_t> voff(parts.size() + 1, 0), eoff(parts.size() + 1, 0);
    for (size_t k = 0; k < parts.size(); ++k) {
        voff[k + 1] = voff[k] + parts[k].nv;
        eoff[k + 1] = eoff[k] + uint32_t(parts[k].e_v0.size());
    }
    m.nv = voff.back();
    m.part_off = voff;

    // distance constraints, colour-major; map[k][i] = union index of part k's constraint i
    std::vector<std::vector<uint32_t>> dmap(parts.size()), bmap(parts.size());
    for (size_t k = 0; k < parts.size(); ++k) {
        dmap[k].assign(parts[k].d_v0.size(), UINT32_MAX);
        bmap[k].assign(parts[k].b_v0.size(), UINT32_MAX);
    }
    for (size_t c = 0; c < nd_colours; ++c) {
        m.d_col.offset.push_back(uint32_t(m.d_v0.size()));
        for (size_t k = 0; k < parts.size(); ++k) {
            const Mesh &p = parts[k];
            if (c >= p.d_col.count.size()) continue;
            for (uint32_t i = p.d_col.offset[c]; i < p.d_col.offset[c] + p.d_col.count[c]; ++i) {
                dmap[k][i] = uint32_t(m.d_v0.size());
                m.d_v0.push_back(p.d_v0[i] + voff[k]);
                m.d_v1.push_back(p.d_v1[i] + voff[k]);
                m.d_type.push_back(p.d_type[i]);
                m.d_rest.push_back(p.d_rest[i]);
                m.d_tc.push_back(p.d_tc[i]);
                m.d_cc.push_back(p.d_cc[i]);
            }
        }
        m.d_col.count.push_back(uint32_t(m.d_v0.size()) - m.d_col.offset.back());
    }
    for (size_t c = 0; c < nb_colours; ++c) {
        m.b_col.offset.push_back(uint32_t(m.b_v0.size()));
        for (size_t k = 0; k < parts.size(); ++k) {
            const Mesh &p = parts[k];
            if (c >= p.b_col.count.size()) continue;
            for (uint32_t i = p.b_col.offset[c]; i < p.b_col.offset[c] + p.b_col.count[c]; ++i) {
                bmap[k][i] = uint32_t(m.b_v0.size());
                m.b_v0.push_back(p.b_v0[i] + voff[k]);
                m.b_v1.push_back(p.b_v1[i] + voff[k]);
                m.b_v2.push_back(p.b_v2[i] + voff[k]);
                m.b_v3.push_back(p.b_v3[i] + voff[k]);
                m.b_rest.push_back(p.b_rest[i]);
                m.b_comp.push_back(p.b_comp[i]);
                m.b_pad0.push_back(p.b_pad0[i]);
                m.b_pad1.push_back(p.b_pad1[i]);
            }
        }
        m.b_col.count.push_back(uint32_t(m.b_v0.size()) - m.b_col.offset.back());
    }

    uint32_t island_off = 0;
    double edge_len_sum = 0.0;
    size_t edge_cnt = 0;
    uint64_t n_virtual = 0;
    for (size_t k = 0; k < parts.size(); ++k) {
        const Mesh &p = parts[k];
        const uint32_t vo = voff[k];
        append(m.pos0, p.pos0);
        append(m.inv_mass, p.inv_mass);
        append(m.thick, p.thick);
        append(m.layer, p.layer);
        for (uint32_t e : p.edge_to_constraint) m.edge_to_constraint.push_back(e == UINT32_MAX ? UINT32_MAX : dmap[k][e]);
        append_csr(m.adj_off, m.adj_idx, p.adj_off, p.adj_idx, vo, true);
        const size_t th_base = m.th_idx.size();
        append_csr(m.th_off, m.th_idx, p.th_off, p.th_idx, vo, true);
        (void)th_base;
        append(m.th_rest, p.th_rest);
        append(m.local_len, p.local_len);
        {
            const uint32_t base = uint32_t(m.star_v0.size());
            if (m.star_off.empty()) m.star_off.push_back(0);
            for (size_t i = 1; i < p.star_off.size(); ++i) m.star_off.push_back(base + p.star_off[i]);
            for (uint32_t v : p.star_v0) m.star_v0.push_back(v + vo);
            for (uint32_t v : p.star_v1) m.star_v1.push_back(v + vo);
        }
        for (uint32_t v : p.tri) m.tri.push_back(v + vo);
        for (uint32_t v : p.e_v0) m.e_v0.push_back(v + vo);
        for (uint32_t v : p.e_v1) m.e_v1.push_back(v + vo);
        uint32_t nisl = 0;
        for (uint32_t isl : p.island) {
            m.island.push_back(isl + island_off);
            nisl = std::max(nisl, isl + 1);
        }
        island_off += nisl;
        edge_len_sum += double(p.mean_edge) * double(p.e_v0.size());
        edge_cnt += p.e_v0.size();
        m.max_edge = std::max(m.max_edge, p.max_edge);
        n_virtual += p.n_virtual;
        m.has_multiple_layers = m.has_multiple_layers || p.has_multiple_layers;
    }
    // open-boundary bits per unique edge, re-packed at the union's edge offsets
    m.e_bbits.assign((m.e_v0.size() + 31) / 32, 0u);
    for (size_t k = 0; k < parts.size(); ++k)
        for (uint32_t e = 0; e < parts[k].e_v0.size(); ++e)
            if (e / 32 < parts[k].e_bbits.size() && ((parts[k].e_bbits[e >> 5] >> (e & 31)) & 1u)) {
                const uint32_t u = eoff[k] + e;
                m.e_bbits[u >> 5] |= 1u << (u & 31);
            }
    if (m.adj_off.empty()) m.adj_off.push_back(0);
    if (m.th_off.empty()) m.th_off.push_back(0);
    if (m.star_off.empty()) m.star_off.push_back(0);
    m.mean_edge = edge_cnt ? float(edge_len_sum / double(edge_cnt)) : 0.0f;
    m.n_virtual = uint32_t(std::min<uint64_t>(n_virtual, 150000));
    // hash sizing as for a single mesh
    const float th = m.nv ? m.thick[0] : 0.005f;
    const float edge_based = std::min(std::max(m.max_edge * 0.85f, m.mean_edge * 1.0f), 0.025f);
    m.cell_size = std::max(std::max(th * 4.0f, edge_based), 0.01f);
    uint32_t ts = 1;
    while (ts < (m.nv + m.n_virtual) * 4u) ts <<= 1;
    m.table_size = std::min<uint32_t>(std::max<uint32_t>(ts, 1024), 65536);
}

}  // namespace tc
