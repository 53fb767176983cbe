// Host-side mesh build for the XPBD substep library.
//
// Restates ClothMesh::from_raw_with_opts (crates/cloth_core/src/mesh.rs:407-1377),
// the Welsh-Powell colouring (crates/cloth_core/src/coloring.rs:6-173), the
// virtual-sample count that sizes the hash table
// (crates/cloth_core/src/virtual_mesh.rs:75-200) and the hash cell/table rule
// (crates/cloth_core/src/simulation/pipelines.rs:899-912).  Sewing springs give
// seam-gated bending (mesh.rs:706-831), sewing constraints (mesh.rs:909-947) and
// the contracted self-collision topology (mesh.rs:991-1191, 1291-1344, 1382-1490).
// The optional coarse (2-hop) constraints are not part of this build.
//
// Constraint order decides the Gauss-Seidel order inside a colour only up to
// independence, but the colour each constraint gets depends on the order the
// constraints are generated in, so the generation order below is the
// reference's: stretch in input-edge order, then shear per sorted triangle-edge
// key followed by its cross pair, then bending per sorted original-edge key.
#include "tc_layout.h"
#include "tc_mesh.h"

#include <algorithm>
#include <array>
#include <map>
#include <cmath>
#include <cstring>
#include <deque>
#include <numeric>
#include <unordered_map>
#include <unordered_set>

namespace tc {
namespace {

inline uint64_t key2(uint32_t a, uint32_t b) { return (uint64_t(a) << 32) | b; }
inline uint64_t ukey(uint32_t a, uint32_t b) { return a < b ? key2(a, b) : key2(b, a); }

inline float dist_p0_minus_p1(const float *pos, uint32_t a, uint32_t b) {
    const float dx = pos[3 * a] - pos[3 * b];
    const float dy = pos[3 * a + 1] - pos[3 * b + 1];
    const float dz = pos[3 * a + 2] - pos[3 * b + 2];
    return std::sqrt(dx * dx + dy * dy + dz * dz);
}
inline float dist_p1_minus_p0(const float *pos, uint32_t a, uint32_t b) {
    const float dx = pos[3 * b] - pos[3 * a];
    const float dy = pos[3 * b + 1] - pos[3 * a + 1];
    const float dz = pos[3 * b + 2] - pos[3 * a + 2];
    return std::sqrt(dx * dx + dy * dy + dz * dz);
}

// Generic Welsh-Powell: constraints are visited by (degree desc, index asc);
// each takes the smallest colour not used by a constraint sharing a listed vertex.
struct WPInput {
    // registered vertices (build the vertex -> constraint lists)
    std::vector<std::vector<uint32_t>> reg;
    // vertices whose list sizes sum to the degree
    std::vector<std::vector<uint32_t>> deg;
    // vertices whose lists provide the used colours
    std::vector<std::vector<uint32_t>> use;
};

std::vector<uint32_t> welsh_powell(uint32_t nv, const WPInput &in, uint32_t &ncolors) {
    const size_t nc = in.reg.size();
    std::vector<std::vector<uint32_t>> v2c(nv);
    for (size_t i = 0; i < nc; ++i)
        for (uint32_t v : in.reg[i]) v2c[v].push_back(uint32_t(i));
    std::vector<uint32_t> degree(nc, 0);
    for (size_t i = 0; i < nc; ++i)
        for (uint32_t v : in.deg[i]) degree[i] += uint32_t(v2c[v].size());
    std::vector<uint32_t> order(nc);
    std::iota(order.begin(), order.end(), 0u);
    std::sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
        if (degree[a] != degree[b]) return degree[a] > degree[b];
        return a < b;
    });
    std::vector<uint32_t> color(nc, UINT32_MAX);
    std::vector<uint32_t> mark;  // mark[c] == visit stamp -> colour c taken
    ncolors = 0;
    uint32_t stamp = 0;
    for (uint32_t i : order) {
        ++stamp;
        for (uint32_t v : in.use[i])
            for (uint32_t j : v2c[v]) {
                if (j == i || color[j] == UINT32_MAX) continue;
                if (color[j] >= mark.size()) mark.resize(color[j] + 1, 0);
                mark[color[j]] = stamp;
            }
        uint32_t c = 0;
        while (c < mark.size() && mark[c] == stamp) ++c;
        color[i] = c;
        ncolors = std::max(ncolors, c + 1);
    }
    return color;
}

// Stable bucket permutation: sorted position -> original index.
std::vector<uint32_t> bucket_perm(const std::vector<uint32_t> &color, uint32_t ncolors, ColorSet &cs) {
    cs.count.assign(ncolors, 0);
    for (uint32_t c : color) cs.count[c]++;
    cs.offset.assign(ncolors, 0);
    for (uint32_t c = 1; c < ncolors; ++c) cs.offset[c] = cs.offset[c - 1] + cs.count[c - 1];
    std::vector<uint32_t> fill(ncolors, 0), perm(color.size());
    for (uint32_t i = 0; i < color.size(); ++i) perm[cs.offset[color[i]] + fill[color[i]]++] = i;
    return perm;
}

uint32_t expand_bits_10(uint32_t v) {  // mesh.rs:314-320
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}
uint32_t as_u32_sat(float x) {  // Rust `f32 as u32`
    if (!(x > 0.0f)) return 0;
    if (x >= 4294967296.0f) return UINT32_MAX;
    return uint32_t(x);
}
uint32_t morton(const float *p, const float *bmin, const float *inv) {  // mesh.rs:322-327
    uint32_t q[3];
    for (int a = 0; a < 3; ++a)
        q[a] = as_u32_sat(std::min(std::max((p[a] - bmin[a]) * inv[a], 0.0f), 1.0f) * 1023.0f);
    return (expand_bits_10(q[0]) << 2) | (expand_bits_10(q[1]) << 1) | expand_bits_10(q[2]);
}

// Union-find over the sewing pairs; every vertex maps to the smallest vertex of its group.
std::vector<uint32_t> sew_groups(uint32_t nv, const uint32_t *sew, uint32_t nsew) {
    std::vector<uint32_t> parent(nv);
    std::iota(parent.begin(), parent.end(), 0u);
    auto find = [&](uint32_t i) {
        uint32_t r = i;
        while (r != parent[r]) r = parent[r];
        while (i != r) { const uint32_t n = parent[i]; parent[i] = r; i = n; }
        return r;
    };
    for (uint32_t k = 0; k < nsew; ++k) {
        const uint32_t u = sew[2 * k], v = sew[2 * k + 1];
        if (u >= nv || v >= nv || u == v) continue;
        const uint32_t ru = find(u), rv = find(v);
        if (ru != rv) parent[std::max(ru, rv)] = std::min(ru, rv);
    }
    std::vector<uint32_t> rep(nv);
    for (uint32_t i = 0; i < nv; ++i) rep[i] = find(i);
    return rep;
}

inline void sub3(const float *a, const float *b, float *o) { o[0] = a[0] - b[0]; o[1] = a[1] - b[1]; o[2] = a[2] - b[2]; }
inline float norm3(const float *v) { return std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); }

// Distance of two opposite vertices once their triangles are unfolded flat about a shared seam
// axis: vertex q over edge (e0, e1) at along-axis s and height h, likewise r over (f0, f1);
// returns sqrt((s_q - s_r)^2 + (h_q + h_r)^2), or -1 when an edge is shorter than 1e-6
// (mesh.rs:789-822, 1436-1458).
float unfolded_length(const float *pos, uint32_t e0, uint32_t e1, uint32_t qv, uint32_t f0, uint32_t f1, uint32_t rv) {
    float ea[3], eb[3];
    sub3(pos + 3 * e1, pos + 3 * e0, ea);
    sub3(pos + 3 * f1, pos + 3 * f0, eb);
    const float la = norm3(ea), lb = norm3(eb);
    if (!(la > 1e-6f && lb > 1e-6f)) return -1.0f;
    const float ua[3] = {ea[0] / la, ea[1] / la, ea[2] / la};
    float da[3];
    sub3(pos + 3 * qv, pos + 3 * e0, da);
    const float sa = da[0] * ua[0] + da[1] * ua[1] + da[2] * ua[2];
    const float pa[3] = {da[0] - sa * ua[0], da[1] - sa * ua[1], da[2] - sa * ua[2]};
    const float ha = norm3(pa);
    const float ub[3] = {eb[0] / lb, eb[1] / lb, eb[2] / lb};
    float db[3];
    sub3(pos + 3 * rv, pos + 3 * f0, db);
    const float sb = db[0] * ub[0] + db[1] * ub[1] + db[2] * ub[2];
    const float pb[3] = {db[0] - sb * ub[0], db[1] - sb * ub[1], db[2] - sb * ub[2]};
    const float hb = norm3(pb);
    const float ds = sa - sb, dh = ha + hb;
    return std::sqrt(ds * ds + dh * dh);
}

// Rest length of a true two-hop pair that only the sewing joins (compute_seam_two_hop_rest_length,
// mesh.rs:1382-1490): over original neighbours m1 of i and m2 of nbr in one sewing group, the
// direct distance when m1 == m2, else the unfolded length over a shared seam edge (m1-k1, m2-k2
// with k1, k2 in one group), else |i - m1| + |nbr - m2|; the smallest, or 0 if none exceeds 1e-5.
float seam_two_hop_rest(uint32_t i, uint32_t nbr, const float *pos, const std::vector<std::vector<uint32_t>> &orig,
                        const std::vector<uint32_t> &rep) {
    const uint32_t ri = rep[i], rn = rep[nbr];
    float best = 3.4028235e38f;
    for (uint32_t m1 : orig[i]) {
        const uint32_t rm1 = rep[m1];
        if (rm1 == ri || rm1 == rn) continue;
        for (uint32_t m2 : orig[nbr]) {
            if (rep[m2] != rm1) continue;
            if (m1 == m2) {
                best = std::min(best, dist_p0_minus_p1(pos, i, nbr));
                continue;
            }
            bool found = false;
            for (uint32_t k1 : orig[m1]) {
                const uint32_t rk1 = rep[k1];
                if (rk1 == rm1) continue;
                for (uint32_t k2 : orig[m2]) {
                    if (rep[k2] != rk1) continue;
                    const float fl = unfolded_length(pos, m1, k1, i, m2, k2, nbr);
                    if (fl >= 0.0f && fl < best) { best = fl; found = true; }
                }
            }
            if (!found) {
                const float slit = dist_p0_minus_p1(pos, i, m1) + dist_p0_minus_p1(pos, nbr, m2);
                best = std::min(best, slit);
            }
        }
    }
    return (best < 3.4028235e38f && std::isfinite(best) && best > 1e-5f) ? best : 0.0f;
}

}  // namespace

std::vector<uint32_t> color_edge_pairs(uint32_t nv, const uint32_t *pairs, uint32_t np) {
    WPInput in;
    in.reg.resize(np);
    for (uint32_t i = 0; i < np; ++i) {
        const uint32_t a = pairs[2 * i], b = pairs[2 * i + 1];
        if (a < nv && b < nv) in.reg[i] = {a, b};
    }
    in.deg = in.reg;
    in.use = in.reg;
    uint32_t ncol = 0;
    std::vector<uint32_t> col = welsh_powell(nv, in, ncol);
    for (uint32_t i = 0; i < np; ++i)
        if (in.reg[i].empty()) col[i] = 0;
    return col;
}

void areal_inv_masses(const float *pos, uint32_t nv, const uint32_t *faces, uint32_t nf, float density,
                      const float *pin_weights, float *out) {
    std::vector<float> mass(nv, 0.0f);
    if (density > 0.0f && faces) {
        for (uint32_t f = 0; f < nf; ++f) {
            const uint32_t a = faces[3 * f], b = faces[3 * f + 1], c = faces[3 * f + 2];
            if (a >= nv || b >= nv || c >= nv || a == b || b == c || c == a) continue;
            const float *pa = pos + 3 * a, *pb = pos + 3 * b, *pc = pos + 3 * c;
            const float ab[3] = {pb[0] - pa[0], pb[1] - pa[1], pb[2] - pa[2]};
            const float ac[3] = {pc[0] - pa[0], pc[1] - pa[1], pc[2] - pa[2]};
            const float cr[3] = {ab[1] * ac[2] - ab[2] * ac[1], ab[2] * ac[0] - ab[0] * ac[2],
                                 ab[0] * ac[1] - ab[1] * ac[0]};
            const float area = 0.5f * std::sqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]);
            const float share = area * density / 3.0f;
            mass[a] += share;
            mass[b] += share;
            mass[c] += share;
        }
    }
    for (uint32_t i = 0; i < nv; ++i) {
        const float w = pin_weights ? std::min(std::max(pin_weights[i], 0.0f), 1.0f) : 0.0f;
        out[i] = w >= 1.0f ? 0.0f : (1.0f - w) / std::max(mass[i], 1e-9f);
    }
}

std::string build_mesh(const MeshInput &in, Mesh &m) {
    const uint32_t nv = in.nv;
    const float *pos = in.positions;
    m = Mesh();
    m.nv = nv;
    m.pos0.assign(pos, pos + 3 * size_t(nv));
    m.inv_mass.resize(nv);
    m.thick.resize(nv);
    m.layer.resize(nv);
    for (uint32_t i = 0; i < nv; ++i) {
        m.inv_mass[i] = in.inv_masses ? in.inv_masses[i] : 1.0f;
        m.thick[i] = in.thicknesses ? in.thicknesses[i] : in.thickness;
        m.layer[i] = in.layer_ids ? in.layer_ids[i] : in.layer_id;
    }
    const float tc = compliance_of(in.tension_k), cc = compliance_of(in.compression_k);
    const float sc = compliance_of(in.shear_k);
    auto valid_edge = [&](uint32_t a, uint32_t b) { return a < nv && b < nv && a != b; };

    // --- distance constraints (generation order) ---
    struct DC { uint32_t v0, v1, type; float rest, tc, cc; };
    std::vector<DC> dcs;
    dcs.reserve(size_t(in.ne) * 2);
    std::vector<uint32_t> raw_edge_to_dist(in.ne, UINT32_MAX);
    std::unordered_set<uint64_t> original;
    original.reserve(size_t(in.ne) * 2);
    for (uint32_t e = 0; e < in.ne; ++e) {
        const uint32_t a = in.edges[2 * e], b = in.edges[2 * e + 1];
        if (!valid_edge(a, b)) continue;
        raw_edge_to_dist[e] = uint32_t(dcs.size());
        dcs.push_back({a, b, 0u, dist_p0_minus_p1(pos, a, b), tc, cc});
        original.insert(ukey(a, b));
    }

    // triangle edge -> opposite vertices in face order (mesh.rs:514-527)
    std::unordered_map<uint64_t, std::vector<uint32_t>> edge_opp;
    if (in.faces && in.nf) {
        edge_opp.reserve(size_t(in.nf) * 2);
        for (uint32_t f = 0; f < in.nf; ++f) {
            const uint32_t f0 = in.faces[3 * f], f1 = in.faces[3 * f + 1], f2 = in.faces[3 * f + 2];
            if (f0 >= nv || f1 >= nv || f2 >= nv) continue;
            edge_opp[ukey(f0, f1)].push_back(f2);
            edge_opp[ukey(f1, f2)].push_back(f0);
            edge_opp[ukey(f2, f0)].push_back(f1);
        }
    }
    std::vector<uint64_t> sorted_keys;
    sorted_keys.reserve(edge_opp.size());
    for (const auto &kv : edge_opp) sorted_keys.push_back(kv.first);
    std::sort(sorted_keys.begin(), sorted_keys.end());

    // shear (mesh.rs:508-560): non-original triangle edges with two opposite
    // vertices, each followed by the cross pair of those vertices
    {
        std::unordered_set<uint64_t> existing = original;
        for (uint64_t k : sorted_keys) {
            if (original.count(k)) continue;
            const auto &o = edge_opp[k];
            if (o.size() != 2) continue;
            const uint64_t pairs[2] = {k, ukey(o[0], o[1])};
            for (uint64_t p : pairs) {
                const uint32_t a = uint32_t(p >> 32), b = uint32_t(p);
                if (a != b && existing.insert(p).second) dcs.push_back({a, b, 1u, dist_p0_minus_p1(pos, a, b), sc, sc});
            }
        }
    }

    // colour distance constraints (coloring.rs:6-79)
    {
        WPInput wp;
        wp.reg.resize(dcs.size());
        for (size_t i = 0; i < dcs.size(); ++i) wp.reg[i] = {dcs[i].v0, dcs[i].v1};
        wp.deg = wp.reg;
        wp.use = wp.reg;
        uint32_t ncol = 0;
        const std::vector<uint32_t> col = welsh_powell(nv, wp, ncol);
        const std::vector<uint32_t> perm = bucket_perm(col, ncol, m.d_col);
        std::vector<uint32_t> remap(dcs.size());
        const size_t n = dcs.size();
        m.d_v0.resize(n); m.d_v1.resize(n); m.d_type.resize(n);
        m.d_rest.resize(n); m.d_tc.resize(n); m.d_cc.resize(n);
        for (size_t s = 0; s < n; ++s) {
            const DC &c = dcs[perm[s]];
            remap[perm[s]] = uint32_t(s);
            m.d_v0[s] = c.v0; m.d_v1[s] = c.v1; m.d_type[s] = c.type;
            m.d_rest[s] = c.rest; m.d_tc[s] = c.tc; m.d_cc[s] = c.cc;
        }
        m.edge_to_constraint.resize(in.ne);
        for (uint32_t e = 0; e < in.ne; ++e)
            m.edge_to_constraint[e] = raw_edge_to_dist[e] == UINT32_MAX ? UINT32_MAX : remap[raw_edge_to_dist[e]];
    }

    // --- bending (mesh.rs:648-705) and strand bending (mesh.rs:833-892) ---
    struct BC { uint32_t v0, v1, v2, v3; float rest, comp, pad0, pad1; };
    std::vector<BC> bcs;
    const float bend_c = in.bending_k > 0.0f ? 1.0f / in.bending_k : 1e10f;
    for (uint64_t k : sorted_keys) {
        if (!original.count(k)) continue;
        const auto &o = edge_opp[k];
        if (o.size() != 2 || o[0] == o[1]) continue;
        bcs.push_back({uint32_t(k >> 32), uint32_t(k), o[0], o[1], dist_p0_minus_p1(pos, o[0], o[1]), bend_c, 0.f, 0.f});
    }
    const bool sewing = in.sew && in.nsew;
    // seam-gated bending (mesh.rs:706-831): two boundary edges whose end points are sewn pairwise
    // (one group per end) bend about the seam: the opposite vertices keep their unfolded distance,
    // only while closer than it and while the seam's first pair is within half of it (>= 1 cm)
    if (sewing && in.faces && in.nf) {
        const std::vector<uint32_t> rep = sew_groups(nv, in.sew, in.nsew);
        std::vector<uint8_t> has_sew(nv, 0);
        for (uint32_t k = 0; k < in.nsew; ++k) {
            const uint32_t u = in.sew[2 * k], v = in.sew[2 * k + 1];
            if (u < nv && v < nv && u != v) has_sew[u] = has_sew[v] = 1;
        }
        // boundary edges (one opposite vertex) in ascending key order, grouped by their pair of
        // sewing groups, each aligned so that its first end is in the smaller group
        std::map<uint64_t, std::vector<std::array<uint32_t, 3>>> seam;
        for (uint64_t k : sorted_keys) {
            const auto &o = edge_opp[k];
            if (o.size() != 1) continue;
            const uint32_t ea = uint32_t(k >> 32), eb = uint32_t(k);
            if (!has_sew[ea] || !has_sew[eb]) continue;
            const uint32_t ra = rep[ea], rb = rep[eb];
            if (ra == rb) continue;
            if (ra < rb) seam[key2(ra, rb)].push_back({ea, eb, o[0]});
            else seam[key2(rb, ra)].push_back({eb, ea, o[0]});
        }
        for (const auto &kv : seam) {
            if (kv.second.size() != 2) continue;
            const auto &A = kv.second[0], &B = kv.second[1];
            const uint32_t a0 = A[0], a1 = A[1], v2 = A[2], b0 = B[0], b1 = B[1], v3 = B[2];
            if (a0 == b0 || a1 == b1 || v2 == v3 || rep[v2] == rep[v3]) continue;
            const float flat = unfolded_length(pos, a0, a1, v2, b0, b1, v3);
            if (flat < 0.0f || flat <= 1e-5f) continue;
            bcs.push_back({a0, b0, v2, v3, flat, bend_c, 1.0f, std::max(flat * 0.5f, 0.01f)});
            m.has_seams = true;
        }
    }
    {
        std::unordered_set<uint64_t> face_edges;
        if (in.faces)
            for (uint32_t f = 0; f < in.nf; ++f) {
                const uint32_t f0 = in.faces[3 * f], f1 = in.faces[3 * f + 1], f2 = in.faces[3 * f + 2];
                face_edges.insert(ukey(f0, f1));
                face_edges.insert(ukey(f1, f2));
                face_edges.insert(ukey(f2, f0));
            }
        std::unordered_map<uint32_t, std::vector<uint32_t>> strand;
        for (uint32_t e = 0; e < in.ne; ++e) {
            const uint32_t a = in.edges[2 * e], b = in.edges[2 * e + 1];
            if (valid_edge(a, b) && !face_edges.count(ukey(a, b))) {
                strand[a].push_back(b);
                strand[b].push_back(a);
            }
        }
        if (!strand.empty()) {
            std::unordered_set<uint64_t> bend_pairs;
            for (const BC &b : bcs) bend_pairs.insert(ukey(b.v2, b.v3));
            std::vector<uint32_t> us;
            for (const auto &kv : strand) us.push_back(kv.first);
            std::sort(us.begin(), us.end());
            const float strand_c = std::max(bend_c * 10.0f, 0.5f);
            for (uint32_t u : us) {
                const auto &nb = strand[u];
                for (size_t i = 0; i < nb.size(); ++i)
                    for (size_t j = i + 1; j < nb.size(); ++j) {
                        const uint32_t a = nb[i], b = nb[j];
                        if (a != b && bend_pairs.insert(ukey(a, b)).second)
                            bcs.push_back({u, u, a, b, dist_p0_minus_p1(pos, a, b), strand_c, 2.0f, 0.0f});
                    }
            }
        }
    }
    // colour bending (coloring.rs:82-173): the two opposite vertices, plus the
    // hinge vertices for gated constraints (_pad0 > 0.5)
    {
        WPInput wp;
        const size_t n = bcs.size();
        wp.reg.resize(n); wp.deg.resize(n); wp.use.resize(n);
        for (size_t i = 0; i < n; ++i) {
            const BC &c = bcs[i];
            const bool gated = c.pad0 > 0.5f;
            wp.reg[i] = {c.v2, c.v3};
            wp.deg[i] = {c.v2, c.v3};
            if (gated) {
                if (c.v0 < nv && c.v0 != c.v2 && c.v0 != c.v3) wp.reg[i].push_back(c.v0);
                if (c.v1 < nv && c.v1 != c.v2 && c.v1 != c.v3 && c.v1 != c.v0) wp.reg[i].push_back(c.v1);
                if (c.v0 < nv) wp.deg[i].push_back(c.v0);
                if (c.v1 < nv) wp.deg[i].push_back(c.v1);
                wp.use[i] = {c.v0, c.v1, c.v2, c.v3};
            } else {
                wp.use[i] = {c.v2, c.v3};
            }
            wp.use[i].erase(std::remove_if(wp.use[i].begin(), wp.use[i].end(), [&](uint32_t v) { return v >= nv; }),
                            wp.use[i].end());
        }
        uint32_t ncol = 0;
        const std::vector<uint32_t> col = welsh_powell(nv, wp, ncol);
        const std::vector<uint32_t> perm = bucket_perm(col, ncol, m.b_col);
        m.b_v0.resize(n); m.b_v1.resize(n); m.b_v2.resize(n); m.b_v3.resize(n);
        m.b_rest.resize(n); m.b_comp.resize(n); m.b_pad0.resize(n); m.b_pad1.resize(n);
        for (size_t s = 0; s < n; ++s) {
            const BC &c = bcs[perm[s]];
            m.b_v0[s] = c.v0; m.b_v1[s] = c.v1; m.b_v2[s] = c.v2; m.b_v3[s] = c.v3;
            m.b_rest[s] = c.rest; m.b_comp[s] = c.comp; m.b_pad0[s] = c.pad0; m.b_pad1[s] = c.pad1;
        }
    }

    // --- sewing constraints (mesh.rs:909-947) and their colouring (coloring.rs:236-300) ---
    m.sew_shrink_speed = std::max(in.sew_shrink_speed, 0.0f);
    if (sewing) {
        const float comp = compliance_of(in.sew_stiffness);
        std::vector<uint32_t> sv0, sv1;
        std::vector<float> srest;
        for (uint32_t k = 0; k < in.nsew; ++k) {
            const uint32_t a = in.sew[2 * k], b = in.sew[2 * k + 1];
            if (!valid_edge(a, b)) continue;
            sv0.push_back(a);
            sv1.push_back(b);
            srest.push_back(dist_p0_minus_p1(pos, a, b));
        }
        WPInput wp;
        wp.reg.resize(sv0.size());
        for (size_t i = 0; i < sv0.size(); ++i) wp.reg[i] = {sv0[i], sv1[i]};
        wp.deg = wp.reg;
        wp.use = wp.reg;
        uint32_t ncol = 0;
        const std::vector<uint32_t> col = welsh_powell(nv, wp, ncol);
        const std::vector<uint32_t> perm = bucket_perm(col, ncol, m.s_col);
        const size_t n = sv0.size();
        m.s_v0.resize(n); m.s_v1.resize(n); m.s_rest.resize(n); m.s_target.assign(n, 0.0f);
        m.s_speed.assign(n, m.sew_shrink_speed); m.s_comp.assign(n, comp); m.s_lock.assign(n, in.sew_lock ? 1.0f : 0.0f);
        for (size_t q = 0; q < n; ++q) {
            m.s_v0[q] = sv0[perm[q]];
            m.s_v1[q] = sv1[perm[q]];
            m.s_rest[q] = srest[perm[q]];
        }
    }

    // --- adjacency, local edge lengths (mesh.rs:949-1075) ---
    std::vector<std::vector<uint32_t>> adj(nv);
    std::vector<float> len_sum(nv, 0.0f);
    std::vector<uint32_t> len_cnt(nv, 0);
    float total = 0.0f, mx = 0.0f;
    uint32_t nvalid = 0;
    for (uint32_t e = 0; e < in.ne; ++e) {
        const uint32_t a = in.edges[2 * e], b = in.edges[2 * e + 1];
        if (!valid_edge(a, b)) continue;
        const float len = dist_p0_minus_p1(pos, a, b);
        total += len;
        if (len > mx) mx = len;
        ++nvalid;
        adj[a].push_back(b);
        adj[b].push_back(a);
        len_sum[a] += len; len_cnt[a]++;
        len_sum[b] += len; len_cnt[b]++;
    }
    m.mean_edge = nvalid ? total / float(nvalid) : 0.0f;
    m.max_edge = mx;
    // the original adjacency (edge order, before sewing) and the sewing groups
    const std::vector<std::vector<uint32_t>> orig = adj;
    std::vector<uint32_t> rep(nv);
    std::iota(rep.begin(), rep.end(), 0u);
    if (sewing) {
        // sewn vertices are one node (mesh.rs:991-1063): a group's vertices are neighbours of each
        // other, and every neighbour of a group vertex is a neighbour of the whole group (both ways)
        rep = sew_groups(nv, in.sew, in.nsew);
        std::vector<std::vector<uint32_t>> group(nv);
        for (uint32_t i = 0; i < nv; ++i) group[rep[i]].push_back(i);
        for (const auto &g : group)
            if (g.size() > 1)
                for (uint32_t u : g)
                    for (uint32_t v : g)
                        if (u != v) adj[u].push_back(v);
        const std::vector<std::vector<uint32_t>> base = adj;
        for (const auto &g : group) {
            if (g.size() < 2) continue;
            std::vector<uint32_t> shared;
            for (uint32_t u : g)
                for (uint32_t n : base[u])
                    if (rep[n] != rep[u]) shared.push_back(n);
            std::sort(shared.begin(), shared.end());
            shared.erase(std::unique(shared.begin(), shared.end()), shared.end());
            for (uint32_t u : g)
                for (uint32_t n : shared) {
                    adj[u].push_back(n);
                    adj[n].push_back(u);
                }
        }
    }
    m.local_len.resize(nv);
    for (uint32_t i = 0; i < nv; ++i)
        m.local_len[i] = len_cnt[i] ? len_sum[i] / float(len_cnt[i]) : std::max(in.thickness, 0.01f);
    m.adj_off.assign(nv + 1, 0);
    for (uint32_t i = 0; i < nv; ++i) {
        auto &l = adj[i];
        std::sort(l.begin(), l.end());
        l.erase(std::unique(l.begin(), l.end()), l.end());
        m.adj_off[i + 1] = m.adj_off[i] + uint32_t(l.size());
    }
    m.adj_idx.reserve(m.adj_off[nv]);
    for (uint32_t i = 0; i < nv; ++i) m.adj_idx.insert(m.adj_idx.end(), adj[i].begin(), adj[i].end());

    // two-hop CSR with rest lengths of true two-hop face-vertex pairs (mesh.rs:1087-1191)
    std::vector<uint8_t> is_face(nv, 0);
    if (in.faces)
        for (uint32_t k = 0; k < 3 * in.nf; ++k)
            if (in.faces[k] < nv) is_face[in.faces[k]] = 1;
    // one-hop neighbours of the contracted graph, as group representatives (mesh.rs:1103-1118)
    std::vector<std::vector<uint32_t>> c1(sewing ? nv : 0);
    if (sewing) {
        for (uint32_t e = 0; e < in.ne; ++e) {
            const uint32_t a = in.edges[2 * e], b = in.edges[2 * e + 1];
            if (!valid_edge(a, b) || rep[a] == rep[b]) continue;
            c1[rep[a]].push_back(rep[b]);
            c1[rep[b]].push_back(rep[a]);
        }
        for (auto &l : c1) {
            std::sort(l.begin(), l.end());
            l.erase(std::unique(l.begin(), l.end()), l.end());
        }
    }
    std::unordered_map<uint64_t, float> seam_memo;
    m.th_off.assign(nv + 1, 0);
    std::vector<uint32_t> nbh, onb;
    for (uint32_t i = 0; i < nv; ++i) {
        nbh.clear();
        nbh.push_back(i);
        for (uint32_t u : adj[i]) {
            nbh.push_back(u);
            nbh.insert(nbh.end(), adj[u].begin(), adj[u].end());
        }
        std::sort(nbh.begin(), nbh.end());
        nbh.erase(std::unique(nbh.begin(), nbh.end()), nbh.end());
        if (sewing) {  // two-hop neighbours of the original mesh
            onb.clear();
            onb.push_back(i);
            for (uint32_t u : orig[i]) {
                onb.push_back(u);
                onb.insert(onb.end(), orig[u].begin(), orig[u].end());
            }
            std::sort(onb.begin(), onb.end());
            onb.erase(std::unique(onb.begin(), onb.end()), onb.end());
        }
        for (uint32_t j : nbh) {
            m.th_idx.push_back(j);
            float rest = 0.0f;
            if (!sewing) {
                const bool two_hop = j != i && is_face[i] && is_face[j] &&
                                     !std::binary_search(adj[i].begin(), adj[i].end(), j);
                if (two_hop) rest = dist_p0_minus_p1(pos, i, j);
            } else {
                // a true two-hop pair of the contracted graph: face vertices of different groups,
                // not joined by an original edge, their groups not neighbours (mesh.rs:1160-1171)
                const bool two_hop = j != i && is_face[i] && is_face[j] && rep[i] != rep[j] &&
                                     std::find(orig[i].begin(), orig[i].end(), j) == orig[i].end() &&
                                     !std::binary_search(c1[rep[i]].begin(), c1[rep[i]].end(), rep[j]);
                if (two_hop) {
                    if (std::binary_search(onb.begin(), onb.end(), j)) {
                        rest = dist_p0_minus_p1(pos, i, j);
                    } else {
                        const uint64_t key = ukey(i, j);
                        auto it = seam_memo.find(key);
                        if (it == seam_memo.end())
                            it = seam_memo.emplace(key, seam_two_hop_rest(i, j, pos, orig, rep)).first;
                        rest = it->second;
                    }
                }
            }
            m.th_rest.push_back(rest);
        }
        m.th_off[i + 1] = uint32_t(m.th_idx.size());
    }

    // star pairs and triangles (mesh.rs:1193-1226)
    std::vector<std::vector<std::pair<uint32_t, uint32_t>>> star(nv);
    if (in.faces)
        for (uint32_t f = 0; f < in.nf; ++f) {
            const uint32_t f0 = in.faces[3 * f], f1 = in.faces[3 * f + 1], f2 = in.faces[3 * f + 2];
            if (f0 >= nv || f1 >= nv || f2 >= nv) continue;
            star[f0].push_back({f1, f2});
            star[f1].push_back({f2, f0});
            star[f2].push_back({f0, f1});
            m.tri.insert(m.tri.end(), {f0, f1, f2});
        }
    m.star_off.assign(nv + 1, 0);
    for (uint32_t i = 0; i < nv; ++i) {
        m.star_off[i + 1] = m.star_off[i] + uint32_t(star[i].size());
        for (const auto &p : star[i]) { m.star_v0.push_back(p.first); m.star_v1.push_back(p.second); }
    }

    // islands by BFS (mesh.rs:1228-1255)
    m.island.assign(nv, UINT32_MAX);
    uint32_t nisl = 0;
    for (uint32_t s = 0; s < nv; ++s) {
        if (m.island[s] != UINT32_MAX) continue;
        std::deque<uint32_t> q{s};
        m.island[s] = nisl;
        while (!q.empty()) {
            const uint32_t u = q.front();
            q.pop_front();
            for (uint32_t v : adj[u])
                if (m.island[v] == UINT32_MAX) { m.island[v] = nisl; q.push_back(v); }
        }
        ++nisl;
    }

    // unique edges in first-seen order, stably sorted by the Morton code of the midpoint (mesh.rs:1257-1289)
    {
        std::unordered_set<uint64_t> seen;
        std::vector<std::pair<uint32_t, uint32_t>> es;
        for (uint32_t e = 0; e < in.ne; ++e) {
            const uint32_t a = in.edges[2 * e], b = in.edges[2 * e + 1];
            if (valid_edge(a, b) && seen.insert(ukey(a, b)).second) es.push_back({std::min(a, b), std::max(a, b)});
        }
        if (!es.empty() && nv) {
            float bmin[3] = {3.4028235e38f, 3.4028235e38f, 3.4028235e38f};
            float bmax[3] = {-3.4028235e38f, -3.4028235e38f, -3.4028235e38f};
            for (uint32_t i = 0; i < nv; ++i)
                for (int a = 0; a < 3; ++a) {
                    bmin[a] = std::min(bmin[a], pos[3 * i + a]);
                    bmax[a] = std::max(bmax[a], pos[3 * i + a]);
                }
            float inv[3];
            for (int a = 0; a < 3; ++a) inv[a] = 1.0f / std::max(bmax[a] - bmin[a], 1e-5f);
            std::vector<uint32_t> code(es.size());
            for (size_t k = 0; k < es.size(); ++k) {
                const float *p0 = pos + 3 * es[k].first, *p1 = pos + 3 * es[k].second;
                const float mid[3] = {(p0[0] + p1[0]) * 0.5f, (p0[1] + p1[1]) * 0.5f, (p0[2] + p1[2]) * 0.5f};
                code[k] = morton(mid, bmin, inv);
            }
            std::vector<uint32_t> idx(es.size());
            std::iota(idx.begin(), idx.end(), 0u);
            std::stable_sort(idx.begin(), idx.end(), [&](uint32_t a, uint32_t b) { return code[a] < code[b]; });
            for (uint32_t k : idx) { m.e_v0.push_back(es[k].first); m.e_v1.push_back(es[k].second); }
        }
        // open-boundary bits (mesh.rs:1291-1344): an edge in no triangle, or in one triangle whose
        // contracted edge (sewing groups) is in one triangle too; edges inside a group are never open
        const size_t ne = m.e_v0.size();
        m.e_bbits.assign((ne + 31) / 32, 0u);
        std::unordered_map<uint64_t, uint32_t> tri_count, ctri_count;
        if (in.faces && in.nf)
            for (uint32_t f = 0; f < in.nf; ++f) {
                const uint32_t f0 = in.faces[3 * f], f1 = in.faces[3 * f + 1], f2 = in.faces[3 * f + 2];
                const uint32_t pr[3][2] = {{f0, f1}, {f1, f2}, {f2, f0}};
                for (const auto &p : pr)
                    if (p[0] < nv && p[1] < nv && p[0] != p[1]) {
                        tri_count[ukey(p[0], p[1])]++;
                        if (rep[p[0]] != rep[p[1]]) ctri_count[ukey(rep[p[0]], rep[p[1]])]++;
                    }
            }
        for (size_t k = 0; k < ne; ++k) {
            const uint32_t a = m.e_v0[k], b = m.e_v1[k];
            if (rep[a] == rep[b]) continue;
            bool open = true;
            if (in.faces && in.nf) {
                auto it = tri_count.find(key2(a, b));
                const uint32_t raw = it == tri_count.end() ? 0u : it->second;
                if (raw == 1) {
                    auto ct = ctri_count.find(ukey(rep[a], rep[b]));
                    open = ct != ctri_count.end() && ct->second == 1;
                } else {
                    open = raw == 0;
                }
            }
            if (open) m.e_bbits[k >> 5] |= 1u << (k & 31);
        }
    }

    // island layers by mean z when all vertices share one layer (pipelines.rs:131-166)
    if (nv) {
        const uint32_t first = m.layer[0];
        const bool same = std::all_of(m.layer.begin(), m.layer.end(), [&](uint32_t l) { return l == first; });
        if (same && nisl > 1) {
            std::vector<float> zs(nisl, 0.0f);
            std::vector<uint32_t> zc(nisl, 0);
            for (uint32_t i = 0; i < nv; ++i) { zs[m.island[i]] += pos[3 * i + 2]; zc[m.island[i]]++; }
            std::vector<float> mz(nisl);
            for (uint32_t k = 0; k < nisl; ++k) mz[k] = zs[k] / float(std::max<uint32_t>(zc[k], 1));
            std::vector<uint32_t> ord(nisl);
            std::iota(ord.begin(), ord.end(), 0u);
            std::stable_sort(ord.begin(), ord.end(), [&](uint32_t a, uint32_t b) { return mz[a] < mz[b]; });
            std::vector<uint32_t> to_layer(nisl);
            for (uint32_t r = 0; r < nisl; ++r) to_layer[ord[r]] = r;
            for (uint32_t i = 0; i < nv; ++i) m.layer[i] = to_layer[m.island[i]];
        }
        m.has_multiple_layers =
            std::any_of(m.layer.begin(), m.layer.end(), [&](uint32_t l) { return l != m.layer[0]; });
    }

    // number of virtual samples (virtual_mesh.rs:75-200); it only sizes the hash table
    {
        uint64_t total_v = 0;
        const float spacing = nv ? m.thick[0] * 1.5f : 0.0f;
        const size_t ntri = m.tri.size() / 3;
        if (spacing > 1e-6f && ntri > 0) {
            std::unordered_set<uint64_t> tri_edges;
            for (size_t f = 0; f < ntri; ++f) {
                const uint32_t a = m.tri[3 * f], b = m.tri[3 * f + 1], c = m.tri[3 * f + 2];
                tri_edges.insert(ukey(a, b));
                tri_edges.insert(ukey(b, c));
                tri_edges.insert(ukey(c, a));
            }
            for (size_t k = 0; k < m.d_v0.size(); ++k) {
                const uint32_t u = m.d_v0[k], v = m.d_v1[k];
                if (!tri_edges.count(ukey(u, v))) continue;
                uint32_t ke = uint32_t(std::ceil(dist_p1_minus_p0(pos, u, v) / spacing));
                ke = std::min<uint32_t>(ke, 16);
                if (ke > 1) total_v += ke - 1;
            }
            for (size_t f = 0; f < ntri; ++f) {
                const uint32_t a = m.tri[3 * f], b = m.tri[3 * f + 1], c = m.tri[3 * f + 2];
                const float lmax = std::max(std::max(dist_p1_minus_p0(pos, a, b), dist_p1_minus_p0(pos, b, c)),
                                            dist_p1_minus_p0(pos, c, a));
                uint32_t k = std::min<uint32_t>(uint32_t(std::ceil(lmax / spacing)), 16);
                if (k >= 3) total_v += uint64_t(k - 1) * (k - 2) / 2;
            }
        }
        m.n_virtual = uint32_t(std::min<uint64_t>(total_v, 150000));
    }
    {
        const float th = nv ? m.thick[0] : 0.005f;
        const float edge_based = std::min(std::max(m.max_edge * 0.85f, m.mean_edge * 1.0f), 0.025f);
        m.cell_size = std::max(std::max(th * 4.0f, edge_based), 0.01f);
        uint32_t ts = 1;
        while (ts < (nv + m.n_virtual) * 4u) ts <<= 1;
        m.table_size = std::min<uint32_t>(std::max<uint32_t>(ts, 1024), 65536);
    }
    return std::string();
}

}  // namespace tc
