// Host-side cloth mesh build: constraint generation, Welsh-Powell colouring
// and the topology tables the self-collision kernels read.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace tc {

struct MeshInput {
    const float *positions = nullptr;  // 3 * nv
    uint32_t nv = 0;
    const uint32_t *edges = nullptr;   // 2 * ne
    uint32_t ne = 0;
    const uint32_t *faces = nullptr;   // 3 * nf (may be null)
    uint32_t nf = 0;
    const float *inv_masses = nullptr;
    const float *thicknesses = nullptr;
    const uint32_t *layer_ids = nullptr;
    uint32_t layer_id = 0;
    float thickness = 0.005f;
    float tension_k = 1000.f, compression_k = 1000.f, shear_k = 500.f, bending_k = 10.f;
    // sewing springs (vertex pairs, 2 * nsew); with sewing the sewn vertices count as one node of
    // the self-collision topology and seams get gated bending constraints
    const uint32_t *sew = nullptr;
    uint32_t nsew = 0;
    float sew_shrink_speed = 1.0f;  // m/s
    float sew_stiffness = 10000.f;  // compliance stored per constraint (the solver uses the simulator's)
    bool sew_lock = true;           // lock flag stored per constraint (likewise)
};

struct ColorSet {
    std::vector<uint32_t> offset, count;  // per colour
};

struct Mesh {
    uint32_t nv = 0;
    std::vector<float> pos0;           // 3 * nv
    std::vector<float> inv_mass, thick;
    std::vector<uint32_t> layer;

    // distance constraints in colour order
    std::vector<uint32_t> d_v0, d_v1, d_type;
    std::vector<float> d_rest, d_tc, d_cc;
    ColorSet d_col;
    std::vector<uint32_t> edge_to_constraint;  // input edge -> colour-order index (UINT32_MAX invalid)

    // bending constraints in colour order
    std::vector<uint32_t> b_v0, b_v1, b_v2, b_v3;
    std::vector<float> b_rest, b_comp, b_pad0, b_pad1;
    ColorSet b_col;

    // sewing constraints in colour order: current (shrinking) and target rest lengths, shrink speed,
    // stored compliance and lock flag (GpuSewingConstraint, mesh.rs:51-60)
    std::vector<uint32_t> s_v0, s_v1;
    std::vector<float> s_rest, s_target, s_speed, s_comp, s_lock;
    ColorSet s_col;
    float sew_shrink_speed = 1.0f;
    bool has_seams = false;  // some bending constraint is seam-gated (_pad0 in (0.5, 1.5))

    // topology for self-collision
    std::vector<uint32_t> adj_off, adj_idx;
    std::vector<uint32_t> th_off, th_idx;
    std::vector<float> th_rest;
    std::vector<float> local_len;
    std::vector<uint32_t> star_off, star_v0, star_v1;
    std::vector<uint32_t> tri;                  // 3 * ntri
    std::vector<uint32_t> e_v0, e_v1, e_bbits;  // Morton-sorted unique edges
    std::vector<uint32_t> island;
    float mean_edge = 0.f, max_edge = 0.f;
    uint32_t n_virtual = 0;
    float cell_size = 0.01f;
    uint32_t table_size = 1024;
    bool has_multiple_layers = false;

    // independent parts (instances of a batch): vertex offsets, nparts + 1 entries; empty = one part
    std::vector<uint32_t> part_off;
};

// Disjoint union of independently built meshes (a batch of cloth instances).  Vertices are
// concatenated; colour c of the union holds colour c of every instance, instance by instance, so
// each instance keeps its own colouring and its own Gauss-Seidel order (constraints of one colour
// share no vertex, so their order inside the colour does not matter).
void merge_meshes(const std::vector<Mesh> &parts, Mesh &out);

// Builds the mesh. Returns an empty string on success, else an error message.
std::string build_mesh(const MeshInput &in, Mesh &out);

// Welsh-Powell colouring of generic vertex pairs (coloring.rs:180-...).
std::vector<uint32_t> color_edge_pairs(uint32_t nv, const uint32_t *pairs, uint32_t np);

// Areal inverse masses (mesh.rs:237-312, faces only).
void areal_inv_masses(const float *pos, uint32_t nv, const uint32_t *faces, uint32_t nf, float density,
                      const float *pin_weights, float *out);

inline float compliance_of(float k) {  // mesh.rs:450-472
    if (k >= 5000.0f) return 0.0f;
    if (k > 0.0f) return 1.0f / k;
    return 1e10f;
}

}  // namespace tc
