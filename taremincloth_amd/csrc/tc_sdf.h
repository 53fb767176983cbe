// Mesh SDF bake: host-side preparation (domain, 20-float collider info, bake
// triangles, open-boundary closing) and the device bakes (tc_sdf_bake.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

namespace tc {
namespace sdf {

// Bake triangle (GpuBakeMeshTriangle, sdf_baker.rs): three corners and the unit face normal.
struct BakeTri {
    float p0x, p0y, p0z, pad0;
    float p1x, p1y, p1z, pad1;
    float p2x, p2y, p2z, pad2;
    float nx, ny, nz, pad3;
};

static_assert(sizeof(BakeTri) == 64, "a cluster of 32 triangles is staged as 128 float4");

struct Domain {
    float min[3], max[3], size[3];
    uint32_t w, h, d;
};

// close_open_boundary_loops (sdf_baker.rs:244-462): caps closed boundary loops of a solid mesh;
// with `weights` (one vector per bone) each new cap centre takes the loop's mean weight.
// Returns the number of loops closed.
size_t close_open_boundary_loops(std::vector<float> &verts, std::vector<int32_t> &tris,
                                 std::vector<std::vector<float>> *weights = nullptr);
// mesh_sdf_domain (sdf_baker.rs:939-980)
Domain mesh_sdf_domain(const std::vector<float> &verts, float voxel_size, float margin, float thickness);
// mesh_sdf_bone_info (sdf_baker.rs:984-1013)
void mesh_sdf_bone_info(const Domain &dm, float friction, float thickness, float restitution, float out[20]);
// collect_mesh_bake_triangles (sdf_baker.rs:1016-1065): invalid indices skipped, dummy when empty.
std::vector<BakeTri> collect_bake_triangles(const std::vector<float> &verts, const std::vector<int32_t> &tris);

// Triangles in Morton order of their centroids, grouped into clusters of up to CLUSTER
// consecutive triangles with an axis-aligned box each (the culling hierarchy of the dense bake).
constexpr uint32_t CLUSTER = 32;
struct Clusters {
    std::vector<BakeTri> tris;      // Morton order
    std::vector<uint32_t> index;    // original index of each sorted triangle
    std::vector<float4> lo, hi;     // per cluster box (w unused)
};
Clusters build_clusters(const std::vector<BakeTri> &tris, const Domain &dm);

// Device bakes; every output voxel is pack2x16float(signed distance, 1.0) in Z, Y, X order.
struct DenseArgs {
    const BakeTri *tris;        // Morton order
    const BakeTri *otris;       // original order (coarse pass: replay when more than 8 ties)
    const uint32_t *index;      // original index per sorted triangle
    const float4 *clo, *chi;    // cluster boxes
    uint32_t ntris, nclusters;
    float lmin[3], lmax[3];
    uint32_t w, h, d;
    uint32_t topk;              // 1 (dense bake) or 8 (coarse pass of the hierarchical bake)
    uint32_t *out;              // packed distances
    uint32_t *nearest;          // topk == 8: 8 triangle indices per voxel
    unsigned long long *keys;   // per voxel (squared distance bits << 32 | original index): the nearest
                                // (topk 1) or an upper bound on the 16th nearest (topk 8)
    unsigned long long item0;   // first (brick, segment) item of a launch of the segmented traversal
    unsigned long long *part;   // topk == 8: 16 keys per cell and cluster segment, part_bricks bricks
    uint32_t part_bricks;       // bricks of 4x4x4 cells whose lists fit in `part`
    uint32_t brick0, brick_count;  // the chunk of bricks of one coarse launch
};
constexpr uint32_t SEG_CLUSTERS = 128;  // clusters per segment of the segmented traversals
// fp32 brick atlas of a dense bake: ceil(w/8) x ceil(h/8) x ceil(d/8) bricks of 8^3 voxels, bricks
// and the voxels inside a brick both in Z/Y/X order; a voxel past the domain's edge holds the value
// of the nearest voxel inside it (clamp to edge).  One wavefront per brick reads the dense bake's
// (d^2, index) keys and writes the signed fp32 distance to that triangle.
constexpr uint32_t ATLAS_BRICK = 8;
hipError_t launch_bake_atlas(hipStream_t st, const DenseArgs &a, float *atlas);
// topk 1: upper bound, (brick, cluster segment) traversal, finish; topk 8: upper bound, then per
// chunk of bricks the (brick, segment) lists and their merge.
hipError_t launch_bake_dense(hipStream_t st, const DenseArgs &a);

struct FineArgs {
    const BakeTri *tris;        // original order
    uint32_t ntris;
    const uint32_t *coarse_dist, *coarse_tri;
    uint32_t cw, ch, cd;
    float band;
    float lmin[3], lmax[3];
    uint32_t w, h, d;
    uint32_t *out;
    uint32_t *amb, *amb_count;  // voxels left to the exhaustive fallback (capacity w*h*d) and their count
};
hipError_t launch_bake_fine(hipStream_t st, const FineArgs &a);

// The fine pass's exhaustive fallback for the listed voxels: one wavefront per voxel walks the
// clusters under an upper bound and keeps the (d_sq, index) minimum, the first strict minimum
// of the reference's in-order scan.
struct ResolveArgs {
    const BakeTri *tris;        // Morton order
    const uint32_t *index;      // original index per sorted triangle
    const float4 *clo, *chi;    // cluster boxes
    uint32_t ntris, nclusters;
    const uint32_t *list;       // flat voxel indices
    uint32_t count;
    float lmin[3], lmax[3];
    uint32_t w, h, d;
    uint32_t *out;
};
hipError_t launch_bake_resolve(hipStream_t st, const ResolveArgs &a);

}  // namespace sdf

namespace bone {

// GpuBakeTriangle (sdf_baker.rs:28-38): corners in bone space with their bone weights, the unit
// face normal and the open-edge mask (bit 0: p0p1, bit 1: p1p2, bit 2: p2p0) stored as a float.
struct BakeTri {
    float p0[3], w0, p1[3], w1, p2[3], w2, n[3], bmask;
};
static_assert(sizeof(BakeTri) == 64, "GpuBakeTriangle is 64 bytes");
// GpuBakeParams (sdf_baker.rs:12-25): one bone's local box, triangle range and atlas tile.
struct Params {
    float lmin[3];
    uint32_t tri_start;
    float lmax[3];
    uint32_t tri_count;
    uint32_t col, row, layer, res, tw, th, td, pitch;
};
static_assert(sizeof(Params) == 64, "GpuBakeParams is 64 bytes");

// bake_sdf.wgsl over `nparams` bones (all of the same resolution `res`): each bone's res^3 tile of
// the atlas `out` (row pitch `pitch` texels, or the atlas width when 0).
hipError_t launch_bone_bake(hipStream_t st, const Params *params, uint32_t nparams, uint32_t res,
                            const BakeTri *tris, uint32_t *out);

}  // namespace bone
}  // namespace tc
