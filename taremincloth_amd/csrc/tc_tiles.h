// Cone-tile plan for running one full solver iteration (every distance colour,
// every sewing colour, then every bending colour) in a single kernel launch.
//
// The vertices are cut into tiles of spatially close vertices (recursive
// coordinate bisection of the rest positions).  For each tile the planner walks the iteration's colour
// phases backwards and collects the backward dependency cone of the tile's
// vertices: the constraints of each phase that touch a vertex whose value is
// needed later, and the vertices those constraints read.  A workgroup loads
// the cone's vertices into LDS, replays the selected constraints phase by
// phase with a block barrier between phases, and writes back its own
// vertices.  Every value a tile writes back is computed from exactly the same
// inputs in the same order as the colour-by-colour sweep, so the result is
// identical to launching one kernel per colour.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "tc_mesh.h"

namespace tc {

struct TilePlan {
    uint32_t ntiles = 0;
    uint32_t nphases = 0;       // distance colours + sewing colours + bending colours
    uint32_t nd_phases = 0;     // the first nd_phases phases are distance colours
    uint32_t ns_phases = 0;     // then ns_phases sewing colours
    uint32_t max_ext = 0;       // largest cone (vertices) -> LDS size
    std::vector<uint32_t> ext_off;   // ntiles + 1
    std::vector<uint32_t> ext_vid;   // cone vertex ids; a tile's own vertices come first
    std::vector<uint32_t> n_own;     // own vertices per tile
    std::vector<uint32_t> ph_off;    // ntiles * (nphases + 1) offsets into items
    std::vector<uint32_t> items;     // 3 words per item (see tc_fused.hip)
    std::vector<float> dcls_tc, dcls_cc, bcls;  // compliance classes
    uint64_t plain_items = 0;        // constraints per iteration without redundancy
};

// Builds the plan; returns "" on success or the reason the mesh cannot use it
// (too many compliance classes, cones too large for LDS, indices too wide).
std::string build_tile_plan(const Mesh &m, uint32_t tile_vertices, uint32_t max_lds_vertices, TilePlan &plan);

}  // namespace tc
